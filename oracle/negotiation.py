"""CPU restatement of the PathFinder negotiation loop -- TEST INFRASTRUCTURE ONLY.

Follows orthoroute/algorithms/manhattan/unified_pathfinder.py (UPF) CPU mode
(CuPy absent) statement by statement, with numpy float32 arithmetic in the
reference's operation order, for two-pin tuple requests (name, src, dst):

  EdgeAccountant              UPF:799-1043
  _pathfinder_negotiation     UPF:3548-4259
  _route_all                  UPF:4441-4823
  _order_nets_by_difficulty   UPF:4401-4439
  _build_hotset               UPF:5009-5140
  _rip_top_k_offenders        UPF:5300-5335
  _detect_barrel_conflicts    UPF:5478-5569
  via usage / keepouts / node ownership  UPF:2678-2904, 3150-3241
  via pooling / hard block (CPU)          UPF:3291-3387, 3483-3546
  _compute_layer_bias         UPF:4825-4878
  _detail_pass                UPF:4306-4399
  analyze_board_characteristics / derive / apply   board_analyzer.py:58-222,
                                                   parameter_derivation.py:51-267

The shortest-path step is pluggable: sssp="ref" emulates SimpleDijkstra
exactly (float64 heap keys, float32 comparisons, ROI-local tie order) and
reproduces the reference's own trajectories (tests/golden/negotiation*.npz);
sssp="canon" uses the FP32 lowest-predecessor semantics of the GPU kernels.

Global `random` state of the reference is reproduced with a private
random.Random: seeded 42 where PadEscapePlanner seeds at init (its only effect
on tuple requests), reseeded 42+it per iteration (UPF:3813-3818).
"""
import math
import os
import random
from collections import defaultdict

import numpy as np

import oracle

f32 = np.float32


class Config:
    """PathFinderConfig defaults (UPF:564-634) plus the getattr defaults the
    live path uses."""

    def __init__(self, **kw):
        self.max_iterations = 40
        self.pres_fac_init = 1.0
        self.pres_fac_mult = 1.10
        self.pres_fac_max = 64.0
        self.hist_gain = 0.2
        self.base_cost_weight = 0.3
        self.grid_pitch = 0.4
        self.via_cost = 0.7
        self.portal_enabled = True
        self.multi_terminal_enabled = False
        self.portal_delta_pref = 6
        self.portal_retarget_patience = 3
        self.stagnation_patience = 5
        self.hotset_cap = 512
        self.hist_cost_weight = 10.0
        self.reroute_only_offenders = True
        self.iter1_always_connect = True
        for k, v in kw.items():
            setattr(self, k, v)


# --------------------------------------------------------------------------
# board analysis / parameter derivation (board_analyzer.py, parameter_derivation.py)
# --------------------------------------------------------------------------

def analyze(lat, bounds, pitch, tasks):
    L = lat.L
    board_w = bounds[2] - bounds[0]
    board_h = bounds[3] - bounds[1]
    signal = list(range(L)) if L <= 2 else list(range(1, L - 1))
    # _assign_hv_layers_by_demand
    if len(signal) <= 2:
        h_layers, v_layers = set(signal), set(signal)
    elif not tasks:
        h_layers = set(l for l in signal if (l % 2 == 1) == True)
        v_layers = set(signal) - h_layers
    else:
        ddx = ddy = 0.0
        for _, (s, d) in tasks.items():
            sx, sy, _ = lat.coords(s)
            dx, dy, _ = lat.coords(d)
            ddx += abs(int(dx) - int(sx)); ddy += abs(int(dy) - int(sy))
        tot = ddx + ddy
        hp = 0.5 if tot < 1e-6 else ddx / tot
        n = len(signal)
        h_goal = max(1, min(n - 1, int(round(n * hp))))
        h_layers, v_layers = set(), set()
        want_h = True
        for layer in signal:
            if want_h and len(h_layers) < h_goal:
                h_layers.add(layer)
            else:
                v_layers.add(layer)
            want_h = not want_h
        while len(h_layers) < h_goal and v_layers:
            mid = sorted(v_layers)[len(v_layers) // 2]; v_layers.remove(mid); h_layers.add(mid)
        while len(h_layers) > h_goal and h_layers:
            mid = sorted(h_layers)[len(h_layers) // 2]; h_layers.remove(mid); v_layers.add(mid)
    cx = int(board_w / pitch)
    cy = int(board_h / pitch)
    tot_h = len(h_layers) * cx * cy
    tot_v = len(v_layers) * cy * cx
    hpwl = 0.0
    for _, (s, d) in tasks.items():
        sx, sy, _ = lat.coords(s)
        dx, dy, _ = lat.coords(d)
        hpwl += abs(int(dx) - int(sx)) * pitch + abs(int(dy) - int(sy)) * pitch
    eff_demand = hpwl * 1.3
    eff_cap = (tot_h + tot_v) * pitch * 0.75
    rho = eff_demand / max(1.0, eff_cap)
    via_spans = L * (L - 1) // 2
    via_flex = via_spans / max(1, L * (L - 1) // 2)
    return {"rho": rho, "layer_count": len(signal), "net_count": len(tasks), "via_flexibility": via_flex}


def derive(ch):
    rho, L, N = ch["rho"], ch["layer_count"], ch["net_count"]
    if rho < 0.6:
        mult, pmax = 1.15, 6.0
    elif rho < 0.9:
        mult, pmax = 1.12, 7.0
    elif rho < 1.2:
        mult, pmax = 1.10, 8.0
    else:
        mult, pmax = 1.08, 10.0
    if L <= 12:
        pmax *= 0.75
    elif L >= 25:
        pmax *= 1.25
    layer_bonus = 6.0 if L <= 10 else (4.0 if L <= 15 else (2.0 if L <= 25 else 0.0))
    hist_w = 8.0 + layer_bonus + 6.0 * max(0.0, min(1.0, rho - 0.7))
    hist_gain = 0.20 + 0.15 * max(0.0, min(1.0, rho - 0.8))
    if rho < 0.7:
        hp = 0.25
    elif rho < 1.0:
        hp = 0.20
    elif rho < 1.3:
        hp = 0.15
    else:
        hp = 0.12
    hotset_cap = max(32, min(int(N * hp), int(N * 0.30)))
    vf = ch["via_flexibility"]
    via_cost = 0.5 if vf > 0.9 else (0.7 if vf > 0.7 else 1.0)
    if rho < 0.8:
        patience = 5
    elif rho < 1.1:
        patience = 6
    else:
        patience = 8
    max_it = 250
    if L <= 12:
        patience = max(patience, 7)
    return dict(pres_fac_init=1.0, pres_fac_mult=mult, pres_fac_max=pmax, hist_cost_weight=hist_w,
                hist_gain=hist_gain, history_decay=1.0, hotset_cap=hotset_cap, via_cost=via_cost,
                stagnation_patience=patience, max_iterations=max_it)


def apply(cfg, d, max_iterations_cap=None):
    for k in ("pres_fac_init", "pres_fac_mult", "pres_fac_max", "hist_cost_weight", "hist_gain",
              "history_decay", "hotset_cap", "via_cost", "stagnation_patience"):
        setattr(cfg, k, d[k])
    cfg.max_iterations = max(cfg.max_iterations, d["max_iterations"])
    if max_iterations_cap is not None:
        cfg.max_iterations = max_iterations_cap


# --------------------------------------------------------------------------
# router
# --------------------------------------------------------------------------

class OraclePortal:
    """The Portal fields the router reads and retargets (PEP:55-84)."""

    def __init__(self, x_idx, y_idx, entry_layer, direction, delta_steps, retarget_count=0):
        self.x_idx, self.y_idx, self.entry_layer = int(x_idx), int(y_idx), int(entry_layer)
        self.direction, self.delta_steps = int(direction), int(delta_steps)
        self.retarget_count = int(retarget_count)


class OracleRouter:
    def __init__(self, X, Y, L, policy="reference", pitch=1.0, via_cost=3.0, sssp="ref", config=None,
                 bounds=None):
        self.cfg = config or Config(grid_pitch=pitch, via_cost=via_cost)
        self.lat = oracle.Lattice(X, Y, L, policy, pitch=pitch, via_cost=via_cost)
        self.X, self.Y, self.L = X, Y, L
        self.N, self.E = self.lat.N, self.lat.E
        self.plane = X * Y
        self.pitch = pitch
        self.bounds = bounds if bounds is not None else (0.0, 0.0, float(X - 1) * pitch, float(Y - 1) * pitch)
        self.sssp = sssp
        lat = self.lat
        # EdgeAccountant
        self.canonical = {}
        self.present = np.zeros(self.E, f32)
        self.present_ema = np.zeros(self.E, f32)
        self.history = np.zeros(self.E, f32)
        self.capacity = np.ones(self.E, f32)
        self.total_cost = None
        # via pooling (UPF:2151-2182)
        self.via_col_cap = np.full((X, Y), 4, np.int16)
        self.via_col_use = np.zeros((X, Y), np.int16)
        self.via_col_pres = np.zeros((X, Y), f32)
        self.segZ = max(L - 2, 0)
        self.has_seg = self.segZ > 0
        if self.has_seg:
            self.via_seg_cap = np.full((X, Y, self.segZ), 2, np.int8)
            self.via_seg_use = np.zeros((X, Y, self.segZ), np.int16)
            self.via_seg_pres = np.zeros((X, Y, self.segZ), f32)
            self.via_seg_prefix = np.zeros((X, Y, self.segZ), f32)
        self.layer_bias = np.ones(L, f32)
        self.node_owner = np.full(self.N, -1, np.int32)
        self.net_id_map = {}
        self.next_net_id = 0
        self.net_paths = {}
        self.net_to_edges = {}
        self.edge_to_nets = defaultdict(set)
        self.keepouts = None            # _via_keepouts_map (absent until first accumulate)
        self.iteration = 0
        self.stagnation_counter = 0
        self.fallback_count = 0
        self.last_ripped = set()
        self.freeze_until = 0
        self.locked_nets = set()
        self.net_clean_streak = defaultdict(int)
        self.rng = random.Random(42)    # PadEscapePlanner(random_seed=42) at init
        self.via_edges = lat.edge_kind.astype(bool)
        src = np.repeat(np.arange(self.N, dtype=np.int64), np.diff(lat.indptr))
        self.edge_src = src.astype(np.int32)
        vi = np.nonzero(self.via_edges)[0]
        u = src[vi]; v = lat.indices[vi].astype(np.int64)
        xu = (u % self.plane) % X; yu = (u % self.plane) // X; zu = u // self.plane; zv = v // self.plane
        self.via_meta = {"indices": vi.astype(np.int32), "x": xu.astype(np.int32), "y": yu.astype(np.int32),
                         "z_lo": np.clip(np.minimum(zu, zv), 1, L - 2).astype(np.int32),
                         "z_hi": np.clip(np.maximum(zu, zv), 1, L - 2).astype(np.int32)}
        self.trace = []   # per-iteration records for tests
        # portals (UPF:2054-2057): pad id -> OraclePortal, net -> (pad, pad)
        self.portals = {}
        self.net_pad_ids = {}
        self.net_portal_failures = defaultdict(int)
        # tree nets (cfg.multi_terminal_enabled): pins, and per routed tree
        # its branches (node lists)
        self.tree_pins = {}
        self.tree_branches = {}

    # ---------------- helpers
    def coords(self, n):
        z, r = divmod(int(n), self.plane)
        y, x = divmod(r, self.X)
        return x, y, z

    def node(self, x, y, z):
        return (z * self.Y + y) * self.X + x

    def get_net_id(self, name):
        if name not in self.net_id_map:
            self.net_id_map[name] = self.next_net_id
            self.next_net_id += 1
        return self.net_id_map[name]

    def path_to_edges(self, path):
        out = []
        ip, ix = self.lat.indptr, self.lat.indices
        for a, b in zip(path, path[1:]):
            s, e = int(ip[a]), int(ip[a + 1])
            for ei in range(s, e):
                if int(ix[ei]) == b:
                    out.append(ei)
                    break
        return out

    # ---------------- EdgeAccountant (UPF:799-1043)
    def refresh_from_canonical(self):
        self.present.fill(0)
        for idx, c in self.canonical.items():
            if 0 <= idx < self.E:
                self.present[idx] = float(c)

    def commit_path(self, edges):
        for idx in edges:
            self.canonical[idx] = self.canonical.get(idx, 0) + 1
            self.present[idx] = self.present[idx] + 1

    def clear_path(self, edges):
        for idx in edges:
            if idx in self.canonical:
                self.canonical[idx] -= 1
                if self.canonical[idx] <= 0:
                    del self.canonical[idx]
            self.present[idx] = np.maximum(0, self.present[idx] - 1)

    def compute_overuse(self):
        edge_over = np.maximum(0, self.present - self.capacity)
        s = int(edge_over.sum())
        c = int((edge_over > 0).sum())
        col = int(np.maximum(0, self.via_col_use - self.via_col_cap).sum())
        seg = int(np.maximum(0, self.via_seg_use - self.via_seg_cap).sum()) if self.has_seg else 0
        return s + col + seg, c

    def update_history(self, gain, base_costs=None, cap_mult=10.0, decay=0.98, use_raw_present=False):
        self.history *= decay
        pres = self.present if use_raw_present else self.present_ema
        over = np.maximum(0, pres - self.capacity)
        inc = gain * over
        if base_costs is not None:
            inc = np.minimum(inc, cap_mult * base_costs)
        self.history += inc

    def update_present_ema(self, beta=0.60):
        self.present_ema = beta * self.present + (1.0 - beta) * self.present_ema

    def update_costs(self, base, pres_fac, hist_weight=1.0, add_jitter=True, vcm=1.0, bcw=0.01,
                     edge_layer=None, layer_bias=None, edge_kind=None):
        over = np.maximum(0, self.present_ema - self.capacity)
        peb = 1.0
        if edge_layer is not None and layer_bias is not None and edge_kind is not None:
            peb = np.where(edge_kind == 0, layer_bias[edge_layer], 1.0)
        adjusted = base * vcm * bcw * peb
        if edge_layer is not None and layer_bias is not None and edge_kind is not None:
            inv = np.where(peb != 0, 1.0 / np.maximum(peb, 0.5), 1.0)
            inv = np.where(edge_kind == 0, inv, 1.0)
            pt = (pres_fac * inv) * over
        else:
            pt = pres_fac * over
        self.total_cost = adjusted + pt + hist_weight * self.history
        if add_jitter:
            jitter = np.arange(self.E, dtype=np.float32) % 9973
            jitter = jitter * 1e-6
            self.total_cost += jitter

    # ---------------- ROI / search
    def extract_roi(self, src, dst, bonus):
        roi = oracle.roi_bfs(self.lat, src, dst, stagnation_bonus=bonus)
        g2r = np.full(self.N, -1, np.int32)
        g2r[roi] = np.arange(len(roi), dtype=np.int32)
        return roi, g2r

    def find_path(self, src, dst, costs, roi_nodes, g2r, keepout_owner=None, net_int=-1):
        if self.sssp == "ref":
            p, _ = oracle.sssp_ref(self.lat, costs, src, dst, roi_nodes, g2r, keepout_owner, net_int)
            return p
        if g2r[src] < 0 or g2r[dst] < 0:
            return None
        mask = (g2r >= 0).astype(np.uint8)
        p, *_ = oracle.sssp_canon(self.lat, costs, src, dst, roi_mask=mask, keepout_owner=keepout_owner,
                                  net_id=net_int)
        return p

    # ---------------- via usage / ownership
    def accumulate_via_usage(self, path, net_id):
        if self.keepouts is None:
            self.keepouts = {}
        L = self.L
        for a, b in zip(path, path[1:]):
            xa, ya, za = self.coords(a)
            xb, yb, zb = self.coords(b)
            if xa == xb and ya == yb and za != zb:
                self.via_col_use[xa, ya] += 1
                if self.has_seg:
                    zlo, zhi = (za, zb) if za < zb else (zb, za)
                    zlo = max(1, min(zlo, L - 2)); zhi = max(1, min(zhi, L - 2))
                    for z in range(zlo, zhi):
                        s = z - 1
                        if 0 <= s < self.segZ:
                            self.via_seg_use[xa, ya, s] += 1
                if net_id:
                    zlo, zhi = (za, zb) if za < zb else (zb, za)
                    for z in range(zlo, zhi + 1):
                        key = (z, xa, ya)
                        if key not in self.keepouts:
                            self.keepouts[key] = net_id

    def segments(self, name, path):
        """The node paths of a committed net: a tree's branches, else its
        path."""
        return self.tree_branches.get(name, [path])

    def rebuild_via_usage(self):
        self.via_col_use.fill(0)
        if self.has_seg:
            self.via_seg_use.fill(0)
        if self.keepouts is not None:
            keep = {k: v for k, v in self.keepouts.items() if v not in self.net_paths}
            self.keepouts.clear()
            self.keepouts.update(keep)
        for nid, path in self.net_paths.items():
            if path and len(path) > 1:
                for seg in self.segments(nid, path):
                    self.accumulate_via_usage(seg, nid)
        self.rebuild_node_owner()

    def rebuild_node_owner(self):
        self.node_owner.fill(-1)
        for name, path in self.net_paths.items():
            if not path or len(path) < 2:
                continue
            nid = self.get_net_id(name)
            for seg in self.segments(name, path):
                self._mark_barrels(seg, nid)
        if self.keepouts:
            for (z, x, y), owner in self.keepouts.items():
                if z <= 0:
                    continue
                n = self.node(x, y, z)
                nid = self.get_net_id(owner)
                if self.node_owner[n] == -1:
                    self.node_owner[n] = nid

    def _mark_barrels(self, path, nid):
        for a, b in zip(path, path[1:]):
            xa, ya, za = self.coords(a)
            xb, yb, zb = self.coords(b)
            if xa == xb and ya == yb and za != zb:
                zlo, zhi = (za, zb) if za < zb else (zb, za)
                for z in range(zlo, zhi + 1):
                    self.node_owner[self.node(xa, ya, z)] = nid

    def keepout_owner_array(self):
        own = np.full(self.N, -1, np.int32)
        for (z, x, y), owner in self.keepouts.items():
            own[self.node(x, y, z)] = self.get_net_id(owner)
        return own

    def detect_barrel_conflicts(self):
        all_e, all_n = [], []
        for name, path in self.net_paths.items():
            if not path or len(path) < 2:
                continue
            nid = self.get_net_id(name)
            e = self.net_to_edges[name] if name in self.tree_branches else self.path_to_edges(path)
            all_e.extend(e); all_n.extend([nid] * len(e))
        if not all_e:
            return np.array([], np.int32), 0
        e = np.array(all_e, np.int32); n = np.array(all_n, np.int32)
        so = self.node_owner[self.edge_src[e]]
        do = self.node_owner[self.lat.indices[e]]
        m = ((so != -1) & (so != n)) | ((do != -1) & (do != n))
        ce = e[m]
        return ce, len(ce)

    # ---------------- per-iteration pieces
    def compute_layer_bias(self, alpha=0.88, max_boost=1.80):
        over = np.maximum(0, self.present_ema - self.capacity)
        per = np.bincount(self.lat.edge_layer, weights=over, minlength=self.L)
        maxv = float(per.max())
        if maxv <= 1e-9:
            raw = np.ones(self.L, np.float32)
        else:
            raw = 1.0 + (1.5 if self.L <= 2 else 0.75) * (per / maxv)
        if not hasattr(self, "_lb_ema"):
            self._lb_ema = raw.astype(np.float32)
        else:
            self._lb_ema = (alpha * self._lb_ema + (1.0 - alpha) * raw).astype(np.float32)
        self._lb_ema = np.clip(self._lb_ema, 1.0, max_boost)
        return self._lb_ema

    def apply_via_pooling(self, pres_fac):
        m = self.via_meta
        tc = self.total_cost
        pen = np.zeros(len(m["indices"]), np.float32)
        pen += 1.0 * self.via_col_pres[m["x"], m["y"]]
        if self.has_seg:
            hi = m["z_hi"] - 2; lo = m["z_lo"] - 2
            valid = m["z_hi"] > m["z_lo"]
            hv = (hi >= 0) & (hi < self.segZ); lv = (lo >= 0) & (lo < self.segZ)
            ph = np.zeros(len(pen), np.float32); pl = np.zeros(len(pen), np.float32)
            if np.any(hv):
                ph[hv] = self.via_seg_prefix[m["x"][hv], m["y"][hv], hi[hv]]
            if np.any(lv):
                pl[lv] = self.via_seg_prefix[m["x"][lv], m["y"][lv], lo[lv]]
            pen += 1.0 * ((ph - pl) * valid)
        k = 0.20
        sb = self.layer_bias[m["z_lo"]]; db = self.layer_bias[m["z_hi"]]
        disc = (1.0 - k * np.maximum(sb, 0.0)) * (1.0 + 0.5 * k * np.maximum(db, 0.0))
        pen *= disc
        mask = pen > 0
        tc[m["indices"][mask]] += pres_fac * pen[mask]

    def block_via_edges(self):
        m = self.via_meta
        for i in range(len(m["indices"])):
            x, y = int(m["x"][i]), int(m["y"][i])
            blk = self.via_col_use[x, y] >= self.via_col_cap[x, y]
            if not blk and self.has_seg:
                for z in range(int(m["z_lo"][i]), int(m["z_hi"][i])):
                    s = z - 1
                    if 0 <= s < self.segZ and self.via_seg_use[x, y, s] >= self.via_seg_cap[x, y, s]:
                        blk = True
                        break
            if blk:
                self.total_cost[m["indices"][i]] = np.float32("inf")

    def order_nets(self, tasks):
        over = np.maximum(0, self.present - self.capacity)
        scores = []
        for nid, (s, d) in tasks.items():
            sx, sy, sz = self.coords(s)
            dx, dy, dz = self.coords(d)
            dist = abs(dx - sx) + abs(dy - sy) + abs(dz - sz)
            cong = 0.0
            if nid in self.net_to_edges:
                edges = self.net_to_edges[nid]
                cong = sum(float(over[e]) for e in edges) / max(1, len(edges))
            scores.append((dist * 3 * (cong + 1), nid))
        scores.sort(reverse=True)
        rot = self.rng.randint(0, min(5, len(scores) // 10))
        ordered = [n for _, n in scores]
        if rot > 0:
            ordered = ordered[rot:] + ordered[:rot]
        return ordered

    def update_tracking(self, nid, edges):
        self.clear_tracking(nid)
        self.net_to_edges[nid] = edges
        for e in edges:
            self.edge_to_nets[e].add(nid)

    def clear_tracking(self, nid):
        self.tree_branches.pop(nid, None)
        if nid in self.net_to_edges:
            for e in self.net_to_edges[nid]:
                self.edge_to_nets[e].discard(nid)
            del self.net_to_edges[nid]

    def portal_nodes(self, nid):
        """(source seed, target seed, source entry node, target entry node)
        of a net with portals on both pads: a seed is the portal's F.Cu node
        (UPF:2495-2513), an entry node the portal on its entry layer
        (UPF:4624-4626); None without portals."""
        if not self.cfg.portal_enabled or nid not in self.net_pad_ids:
            return None
        a, b = self.net_pad_ids[nid]
        sp, dp = self.portals.get(a), self.portals.get(b)
        if sp is None or dp is None:
            return None
        return (self.node(sp.x_idx, sp.y_idx, 0), self.node(dp.x_idx, dp.y_idx, 0),
                self.node(sp.x_idx, sp.y_idx, sp.entry_layer), self.node(dp.x_idx, dp.y_idx, dp.entry_layer))

    def retarget_portals(self, nid):
        """_retarget_portals_for_net (UPF:4929-4969)."""
        for pid in self.net_pad_ids.get(nid, ()):
            p = self.portals.get(pid)
            if p is None:
                continue
            if p.retarget_count == 0:
                p.direction = -p.direction
                p.y_idx = p.y_idx - 2 * p.direction * p.delta_steps
                p.retarget_count += 1
            elif p.retarget_count == 1:
                new = self.cfg.portal_delta_pref
                if new != p.delta_steps:
                    p.y_idx += p.direction * (new - p.delta_steps)
                    p.delta_steps = new
                    p.retarget_count += 1

    def route_all(self, tasks, all_tasks, pres_fac, iteration=0):
        """_route_all (UPF:4441-4823), one net at a time in routing order.  A
        portal net's search is find_path_multisource_multisink with one seed
        and one target (UPF:4732-4738); with one of each it settles the same
        nodes in the same order as find_path_roi between them and returns the
        same path (a target is never expanded and the path is fixed when it
        is popped), so it runs as that search here."""
        self.fallback_count = 0
        bonus = self.stagnation_counter * 0.6
        ordered = self.order_nets(tasks)
        costs = self.total_cost
        total = len(tasks)
        ko_owner = None
        use_ko = total < 100 and self.keepouts is not None and len(self.keepouts) > 0
        if use_ko:
            ko_owner = self.keepout_owner_array()
        rec = []
        for nid in ordered:
            src, dst = tasks[nid]
            if nid in self.net_paths and self.net_paths[nid]:
                old = self.net_to_edges[nid] if nid in self.net_to_edges else self.path_to_edges(self.net_paths[nid])
                self.clear_path(old)
                self.clear_tracking(nid)
            if nid in self.tree_pins:
                # a tree net: grown from its first pin by canonical
                # multi-source searches on the iteration's costs (whole
                # graph, no keepouts, no fallback); committed only when every
                # pin is reached
                branches, _, unreached, nodes = oracle.steiner_canon(self.lat, costs, self.tree_pins[nid])
                if unreached:
                    self.net_paths[nid] = []
                    self.clear_tracking(nid)
                else:
                    edges = [e for b in branches for e in self.path_to_edges(b)]
                    self.commit_path(edges)
                    self.net_paths[nid] = list(nodes)
                    gid = self.get_net_id(nid)
                    for b in branches:
                        self._mark_barrels(b, gid)
                    self.update_tracking(nid, edges)
                    self.tree_branches[nid] = [list(b) for b in branches]
                rec.append((nid, False))
                continue
            sx, sy, _ = self.coords(src)
            dx, dy, _ = self.coords(dst)
            md = abs(dx - sx) + abs(dy - sy)
            pn = self.portal_nodes(nid)
            if md < 125:
                if pn is not None:
                    roi = oracle.roi_bfs(self.lat, pn[2], pn[3], stagnation_bonus=bonus, keep=(pn[0], pn[1]))
                else:
                    roi, g2r = self.extract_roi(src, dst, bonus)
            else:
                roi = np.arange(self.N, dtype=np.int32)
            if pn is not None:
                have = set(roi.tolist())
                add = [v for v in (pn[0], pn[1]) if v not in have]
                if add:
                    roi = np.append(roi, add)
            roi_set = set(roi.tolist())
            if src not in roi_set:
                roi = np.append(roi, src)
            if dst not in roi_set:
                roi = np.append(roi, dst)
            g2r = np.full(self.N, -1, np.int32)
            g2r[roi] = np.arange(len(roi), dtype=np.int32)
            net_int = self.get_net_id_for_keepout(nid) if use_ko else -1
            path = None
            if pn is not None:
                path = self.find_path(pn[0], pn[1], costs, roi, g2r, ko_owner if use_ko else None, net_int)
                if path is not None and len(path) <= 1:
                    path = None
            if pn is None or not path:
                path = self.find_path(src, dst, costs, roi, g2r, ko_owner if use_ko else None, net_int)
            used_fallback = False
            if (not path or len(path) <= 1) and self.fallback_count < 5:
                if pn is not None:
                    roi2, g2r2 = self.extract_roi(pn[2], pn[3], bonus * 2)
                    path = self.find_path(pn[0], pn[1], costs, roi2, g2r2) \
                        if g2r2[pn[0]] >= 0 and g2r2[pn[1]] >= 0 else None
                    if path is not None and len(path) <= 1:
                        path = None
                else:
                    roi2, g2r2 = self.extract_roi(src, dst, bonus * 2)
                    path = self.find_path(src, dst, costs, roi2, g2r2)
                self.fallback_count += 1
                used_fallback = True
            if path and len(path) > 1:
                edges = self.path_to_edges(path)
                self.commit_path(edges)
                self.net_paths[nid] = path
                self._mark_barrels(path, self.get_net_id(nid))
                self.update_tracking(nid, edges)
            else:
                self.net_paths[nid] = []
                self.clear_tracking(nid)
                if self.cfg.portal_enabled and nid in self.net_pad_ids:
                    self.net_portal_failures[nid] += 1
                    if self.net_portal_failures[nid] >= self.cfg.portal_retarget_patience:
                        self.retarget_portals(nid)
                        self.net_portal_failures[nid] = 0
            rec.append((nid, used_fallback))
        total_routed = sum(1 for p in self.net_paths.values() if p)
        return total_routed, len(all_tasks) - total_routed, rec

    def get_net_id_for_keepout(self, nid):
        # keepout owners compare by net name (UPF:3193); names map 1:1 to ids
        return self.get_net_id(nid)

    def build_hotset(self, tasks, ripped=None):
        ripped = ripped or set()
        over = np.maximum(0, self.present - self.capacity)
        hot = over + 0.1 * self.history
        over_idx = set(map(int, np.where(hot > 0.5)[0]))
        if not hasattr(self, "_clean_iters"):
            self._clean_iters = {}
        if len(over_idx) == 0:
            unrouted = {n for n in tasks if not self.net_paths.get(n)}
            return unrouted | ripped
        offenders = set()
        for e in over_idx:
            offenders.update(self.edge_to_nets.get(e, set()))
        for n in tasks:
            if n in offenders:
                self._clean_iters[n] = 0
            else:
                self._clean_iters[n] = self._clean_iters.get(n, 0) + 1
        frozen = {n for n in offenders if self._clean_iters.get(n, 0) >= 3}
        offenders -= frozen
        offenders |= ripped
        unrouted = {n for n in tasks if not self.net_paths.get(n)}
        scores = []
        for n in offenders:
            if n in self.net_to_edges:
                imp = sum(float(over[e]) for e in self.net_to_edges[n] if e in over_idx)
                scores.append((imp, n))
        for n in unrouted:
            if n not in offenders:
                scores.append((0.0, n))
        scores.sort(reverse=True)
        cap = min(self.cfg.hotset_cap, 100)
        rng = random.Random(42 + self.iteration)
        k_top = int(0.6 * cap)
        top = [n for _, n in scores[:k_top]]
        rest = [n for _, n in scores[k_top:]]
        k_rand = min(cap - k_top, len(rest))
        sample = rng.sample(rest, k_rand) if rest else []
        lst = top + sample
        if not hasattr(self, "_last_reroute"):
            self._last_reroute = {}
        cool = [n for n in lst if self.iteration - self._last_reroute.get(n, -999) > 1]
        for n in cool:
            self._last_reroute[n] = self.iteration
        return set(cool)

    def rip_top_k(self, k=20):
        over = np.maximum(0, self.present - self.capacity)
        over_idx = set(map(int, np.where(over > 0)[0]))
        scores = []
        for n, path in self.net_paths.items():
            if not path or n in self.locked_nets:
                continue
            if n in self.net_to_edges:
                imp = sum(float(over[e]) for e in self.net_to_edges[n] if e in over_idx)
                if imp > 0:
                    scores.append((imp, n))
        scores.sort(reverse=True)
        victims = {n for _, n in scores[:k]}
        for n in victims:
            if self.net_paths.get(n) and n in self.net_to_edges:
                self.clear_path(self.net_to_edges[n])
                self.net_paths[n] = []
                self.clear_tracking(n)
                self.net_clean_streak[n] = 0
        return victims

    def rebuild_usage(self, keep):
        self.canonical.clear()
        self.present.fill(0)
        for n in keep:
            if n in self.net_to_edges:
                for e in self.net_to_edges[n]:
                    self.canonical[e] = self.canonical.get(e, 0) + 1
                    self.present[e] += 1

    # ---------------- negotiation (UPF:3548-4259)
    def route_board_nets(self, nets, pad_to_node, max_iterations=None, iteration_cb=None):
        """route_multiple_nets on Board nets given as (name, [pad ids]) with
        the board's pad -> node map: a two-pad net's task is its pads' nodes,
        or their portals' entry-layer nodes when both pads have portals
        (_parse_requests, UPF:2600-2676)."""
        reqs = []
        for name, pads in nets:
            # nets with more than two pads are multi-terminal nets, which the
            # reference does not route (UPF:2548-2585)
            if len(pads) != 2 or pads[0] not in pad_to_node or pads[1] not in pad_to_node:
                continue
            a, b = pads[0], pads[1]
            s, d = pad_to_node[a], pad_to_node[b]
            self.net_pad_ids[name] = (a, b)
            if self.cfg.portal_enabled and a in self.portals and b in self.portals:
                pa, pb = self.portals[a], self.portals[b]
                s = self.node(pa.x_idx, pa.y_idx, pa.entry_layer)
                d = self.node(pb.x_idx, pb.y_idx, pb.entry_layer)
            reqs.append((name, int(s), int(d)))
        return self.route_multiple_nets(reqs, max_iterations, iteration_cb)

    def route_multiple_nets(self, requests, max_iterations=None, iteration_cb=None):
        """requests: (name, src, dst); with cfg.multi_terminal_enabled also
        (name, [pins]), a tree net of its distinct pins (a task of its first
        two pins for the schedule's ordering and analysis)."""
        tasks = {}
        for req in requests:
            if len(req) == 2:
                if not self.cfg.multi_terminal_enabled:
                    continue
                name, nodes = req
                pins = []
                for v in nodes:
                    if int(v) not in pins:
                        pins.append(int(v))
                if len(pins) > 2:
                    self.tree_pins[name] = pins
                    tasks[name] = (pins[0], pins[1])
                elif len(pins) == 2:
                    tasks[name] = (pins[0], pins[1])
                elif pins:
                    self.net_paths[name] = [pins[0]]
                continue
            name, s, d = req
            if s != d:
                tasks[name] = (int(s), int(d))
            else:
                self.net_paths[name] = [s]
        if not tasks:
            return {}
        return self.negotiate(tasks, max_iterations, iteration_cb)

    def negotiate(self, tasks, max_iterations=None, iteration_cb=None):
        cfg = self.cfg
        ch = analyze(self.lat, self.bounds, self.pitch, tasks)
        apply(cfg, derive(ch), max_iterations)
        n_sig = self.L - 2
        pres_fac = float(cfg.pres_fac_init)
        pres_fac_mult = float(cfg.pres_fac_mult)
        hist_gain = float(cfg.hist_gain)
        if n_sig <= 12:
            pres_fac_max, hmult = 32.0, 1.2
        elif n_sig <= 20:
            pres_fac_max, hmult = 64.0, 1.0
        else:
            pres_fac_max, hmult = 128.0, 0.8
        # the reference's environment overrides for testing (UPF:3587-3590)
        pres_fac_mult = float(os.getenv("ORTHO_PRES_FAC_MULT", pres_fac_mult))
        pres_fac_max = float(os.getenv("ORTHO_PRES_FAC_MAX", pres_fac_max))
        hist_gain = float(os.getenv("ORTHO_HIST_GAIN", hist_gain))
        best = float("inf")
        stagnant = 0
        self.net_failure = {}
        self.excluded = set()
        routed = failed = 0
        over_sum = over_cnt = 0
        jam = None
        for it in range(1, cfg.max_iterations + 1):
            self.iteration = it
            if it > 1 and cfg.reroute_only_offenders:
                self.rebuild_usage({n for n, p in self.net_paths.items() if p})
            else:
                self.refresh_from_canonical()
            self.rebuild_via_usage()
            over = np.maximum(0, self.via_col_use - self.via_col_cap).astype(np.float32)
            self.via_col_pres = 0.6 * over + 0.4 * self.via_col_pres
            if self.has_seg:
                over = np.maximum(0, self.via_seg_use - self.via_seg_cap).astype(np.float32)
                self.via_seg_pres = 0.6 * over + 0.4 * self.via_seg_pres
                np.cumsum(self.via_seg_pres, axis=2, out=self.via_seg_prefix)
            present_cache = self.present
            if it == 1:
                self.present_ema = self.present.copy()
            self.update_present_ema(beta=0.40)
            lb = self.compute_layer_bias(alpha=0.88, max_boost=1.80)
            vcm = 1.0
            if pres_fac >= 64:
                ov = np.maximum(0, present_cache - self.capacity)
                via_over = float(ov[self.via_edges[:len(ov)]].sum())
                tot_over = float(ov.sum())
                if tot_over > 0:
                    vcm = 1.5 if via_over / tot_over > 0.7 else 0.5
            self.update_costs(self.lat.costs, pres_fac, cfg.hist_cost_weight * hmult, vcm=vcm,
                              bcw=cfg.base_cost_weight, edge_layer=self.lat.edge_layer, layer_bias=lb,
                              edge_kind=self.lat.edge_kind)
            self.apply_via_pooling(pres_fac)
            if it > 1:
                self.block_via_edges()
            if it > 10 and it % 10 == 0 and self.excluded:
                self.excluded.clear()
                self.net_failure.clear()
            if it > 4:
                for n in tasks:
                    if n in self.last_ripped:
                        continue
                    if not self.net_paths.get(n):
                        self.net_failure[n] = self.net_failure.get(n, 0) + 1
                    else:
                        self.net_failure[n] = 0
                for n, f in self.net_failure.items():
                    if f >= 5 and n not in self.excluded:
                        self.excluded.add(n)
            if cfg.reroute_only_offenders and it > 1:
                offenders = self.build_hotset(tasks, ripped=self.last_ripped) - self.excluded
                sub = {k: v for k, v in tasks.items() if k in offenders and k not in self.excluded}
                self.last_ripped = set()
            else:
                sub = {k: v for k, v in tasks.items() if k not in self.excluded}
            if it > 1:
                ids = list(sub.keys())
                self.rng.seed(42 + it)
                self.rng.shuffle(ids)
                sub = {k: sub[k] for k in ids}
            cost_snapshot = self.total_cost.copy()
            routed, failed, rec = self.route_all(sub, tasks, pres_fac, it)
            self.refresh_from_canonical()
            if it <= 50 or it % 10 == 0:
                ce, cc = self.detect_barrel_conflicts()
            else:
                ce, cc = np.array([], np.int32), getattr(self, "_last_barrel", 0)
            self._last_barrel = cc
            if cc > 0 and it <= 50:
                self.present[ce] += min(10.0 * pres_fac, 100.0)
            over_sum, over_cnt = self.compute_overuse()
            over = np.maximum(0, present_cache - self.capacity)
            by_layer = None
            if over_sum > 0:
                by_layer = self.per_layer_congestion(over)
            if by_layer and sum(by_layer.values()) > 0:
                lo = np.array([by_layer.get(z, 0.0) for z in range(self.L)])
                mean = np.mean(lo[lo > 0]) + 1e-9
                pressure = lo / mean
                if n_sig <= 12:
                    a, bmin, bmax = 0.20, 0.60, 1.80
                elif n_sig <= 20:
                    a, bmin, bmax = 0.12, 0.75, 1.50
                else:
                    a, bmin, bmax = 0.08, 0.85, 1.20
                target = 1.0 + a * (pressure - 1.0)
                self.layer_bias = 0.85 * self.layer_bias + 0.15 * target
                self.layer_bias = np.clip(self.layer_bias, bmin, bmax)
                tot_l = sum(by_layer.values())
                if tot_l > 0:
                    mi, mo = max(by_layer.items(), key=lambda x: x[1])
                    pct = (mo / tot_l) * 100
                    if jam is None:
                        jam = {"layer": None, "count": 0, "boost_until": 0}
                    if pct >= 60.0:
                        if jam["layer"] == mi:
                            jam["count"] += 1
                        else:
                            jam = {"layer": mi, "count": 1, "boost_until": 0}
                        if jam["count"] >= 3 and it > jam["boost_until"]:
                            old = self.layer_bias[mi]
                            self.layer_bias[mi] = min(old * 1.8, bmax)
                            jam["boost_until"] = it + 3
                    else:
                        jam = {"layer": None, "count": 0, "boost_until": 0}
            self.trace.append({"it": it, "over_sum": over_sum, "over_cnt": over_cnt, "pres_fac": pres_fac,
                               "routed": routed, "failed": failed, "sub": list(sub.keys()), "rec": rec,
                               "cost": cost_snapshot})
            if over_sum == 0:
                unrouted = {n for n in tasks if not self.net_paths.get(n)}
                if not unrouted:
                    break
                pres_fac = min(pres_fac, 10.0)
            hg = hist_gain
            if it >= 8 and over_sum > 0:
                hg = min(1.2, hist_gain * 1.25)
            decay = 0.98 if it >= 10 else 1.0
            self.update_history(hg, base_costs=self.lat.costs, cap_mult=15.0, decay=decay, use_raw_present=True)
            if iteration_cb:
                iteration_cb(it, [], [], over_sum, over_cnt)
            if failed == 0 and over_sum == 0:
                if self._last_barrel == 0:
                    return {"success": True, "paths": self.net_paths, "converged": True}
            if over_sum < best:
                best = over_sum
                stagnant = 0
            else:
                stagnant += 1
            if stagnant == 2:
                pres_fac = min(pres_fac * 1.5, pres_fac_max)
            if stagnant >= cfg.stagnation_patience:
                self.stagnation_counter += 1
                self.last_ripped = self.rip_top_k(20)
                self.freeze_until = it + 2
                stagnant = 0
                continue
            if it <= self.freeze_until:
                pass
            else:
                pres_fac = min(pres_fac * pres_fac_mult, pres_fac_max)
        if 0 < over_sum <= 100:
            r = self.detail_pass(tasks, over_sum, over_cnt)
            if r["success"]:
                return r
        success = failed == 0 and over_sum == 0
        return {"success": success, "converged": success, "paths": self.net_paths,
                "barrel_conflicts": getattr(self, "_last_barrel", 0), "excluded_nets": len(self.excluded),
                "excluded_net_ids": list(self.excluded), "error_code": None if success else "ROUTING-FAILED",
                "overuse_sum": over_sum, "overuse_edges": over_cnt, "failed_nets": failed}

    def per_layer_congestion(self, over):
        el = self.lat.edge_layer
        m = over > 0
        if int(np.sum(m)) == 0:
            return {}
        vals = over[m]
        lay = el[m]
        out = {}
        for z in range(1, self.L + 1):
            lm = lay == z
            if np.any(lm):
                out[z] = float(np.sum(vals[lm]))
        return out

    def detail_pass(self, tasks, initial_overuse, initial_edges):
        over = np.maximum(0, self.present - self.capacity)
        conflict = set(int(e) for e in np.nonzero(over > 0)[0])
        nets = set()
        for n, path in self.net_paths.items():
            e = self.net_to_edges.get(n, []) if n in self.tree_branches else (self.path_to_edges(path) if path else [])
            if path and any(x in conflict for x in e):
                nets.add(n)
        if not nets:
            return {"success": False, "error_code": "NO-CONFLICT-NETS"}
        # the reference iterates a set of names here; sorted order is used
        # for determinism (its own order depends on the string hash seed)
        ctasks = {n: tasks[n] for n in sorted(nets) if n in tasks}
        cfg = self.cfg
        pres_fac = cfg.pres_fac_max * 0.5
        best = initial_overuse
        for _ in range(1, 11):
            self.refresh_from_canonical()
            self.update_costs(self.lat.costs, pres_fac, cfg.hist_cost_weight * 1.5, vcm=0.3, bcw=cfg.base_cost_weight)
            hs = self.build_hotset(ctasks)
            hs = set(sorted(hs)[:60])
            if not hs:
                hs = nets
            sub = {k: v for k, v in ctasks.items() if k in hs}
            old = self.stagnation_counter
            self.stagnation_counter += 3
            self.route_all(sub, tasks, pres_fac)
            self.stagnation_counter = old
            self.refresh_from_canonical()
            s, c = self.compute_overuse()
            if s == 0:
                return {"success": True, "paths": self.net_paths}
            if s < best:
                best = s
            self.update_history(cfg.hist_gain * 2.0, base_costs=self.lat.costs, cap_mult=15.0, decay=0.98)
            pres_fac = min(pres_fac * 1.5, cfg.pres_fac_max)
        return {"success": False, "error_code": "DETAIL-INCOMPLETE", "overuse_sum": best}
