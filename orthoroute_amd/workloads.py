"""Synthetic routing workloads (seeded) for the benchmark configurations.

Every generator returns (X, Y, L, requests) with requests as the tuple form
the reference router accepts, (name, src_node, dst_node) (UPF:2660-2669);
node = z*X*Y + y*X + x, pins on layer 0.

  cfg1  96x96x6, 300 two-pin nets, endpoints uniform over distinct top-layer
        cells (all pins unique).
  cfg2  600x400x18 backplane: 20 connectors of 2 rows x 200 pins at 2-cell
        pitch (400 cells wide), non-overlapping; 8,000 nets join pin k of
        connector i to pin k of connector j, (i, j, k) drawn without repeats.
  cfg3  300x300x12 BGA escape: 40x40 balls at 2-cell pitch, centred; 1,400
        nets ball -> random perimeter cell (the multi-pin part of this config
        is routed as its first two pins, as the reference's live path does).
  cfg4  1000x800x24: 40,000 nets between opposite corners of boxes whose
        centres are uniform and half-widths lognormal(ln 20, 1.0), clipped.
  cfg5  serving stream: jobs with exponential inter-arrival times, node
        counts log-uniform over a range, layer counts uniform, square boards
        X = Y = round(sqrt(N / L)), each job's nets generated like cfg1.
        `net_density` sets the net count from the board instead of the
        log-uniform draw: nets = net_density * X * Y / (X + Y) (pins are on
        the top layer, so the count follows the board's area over its
        half-perimeter, the mean net length).
"""
import numpy as np


def _node(X, Y, x, y, z=0):
    return int((z * Y + y) * X + x)


def cfg1(seed=1, X=96, Y=96, L=6, n_nets=300):
    rng = np.random.default_rng(seed)
    cells = rng.permutation(X * Y)[: 2 * n_nets]
    reqs = [(f"N{i:04d}", int(cells[2 * i]), int(cells[2 * i + 1])) for i in range(n_nets)]
    return X, Y, L, reqs


def cfg2(seed=2, X=600, Y=400, L=18, n_conn=20, pins_per_row=200, n_nets=8000):
    rng = np.random.default_rng(seed)
    width = 2 * pins_per_row               # cells spanned by one row
    # rows y0 and y0+2 per connector; stack connectors without overlapping rows
    slots = (Y - 3) // 3 + 1
    if slots < n_conn:
        raise ValueError("board too short for the connectors")
    ys = np.sort(rng.choice(slots, n_conn, replace=False)) * 3
    xs = rng.integers(0, max(1, X - width + 1), n_conn)
    pins_per = 2 * pins_per_row

    def pin(c, k):
        row, col = divmod(int(k), pins_per_row)
        return _node(X, Y, int(xs[c]) + 2 * col, int(ys[c]) + 2 * row)
    seen = set()
    reqs = []
    while len(reqs) < n_nets:
        i, j = (int(v) for v in rng.choice(n_conn, 2, replace=False))
        k = int(rng.integers(0, pins_per))
        if (i, j, k) in seen:
            continue
        seen.add((i, j, k))
        reqs.append((f"B{len(reqs):05d}", pin(i, k), pin(j, k)))
    return X, Y, L, reqs


def _perimeter(X, Y, depth=3):
    """Cells of the board's outer ring `depth` cells deep (the single outer
    ring holds 2(X+Y)-4 = 1,196 cells at 300x300, fewer than cfg3's 1,400
    two-pin nets plus the multi-terminal pads), row-major."""
    return [(x, y) for y in range(Y) for x in range(X)
            if x < depth or y < depth or x >= X - depth or y >= Y - depth]


def cfg3(seed=3, X=300, Y=300, L=12, balls=40, n_nets=1400):
    rng = np.random.default_rng(seed)
    x0 = (X - 2 * (balls - 1)) // 2
    y0 = (Y - 2 * (balls - 1)) // 2
    ball = [(x0 + 2 * i, y0 + 2 * j) for j in range(balls) for i in range(balls)]
    per = _perimeter(X, Y)
    bsel = rng.choice(len(ball), n_nets, replace=False)
    psel = rng.choice(len(per), n_nets, replace=False)
    reqs = [(f"G{i:04d}", _node(X, Y, *ball[b]), _node(X, Y, *per[p])) for i, (b, p) in enumerate(zip(bsel, psel))]
    return X, Y, L, reqs


def cfg3_multi(seed=3, X=300, Y=300, L=12, balls=40, n_nets=1400, n_multi=200):
    """cfg3's 200 multi-terminal nets: one ball and 2-7 perimeter pads each
    (3-8 pins), drawn after cfg3's two-pin nets from the same generator,
    among the balls and perimeter cells those nets left free.  Returns
    (X, Y, L, [(name, [pin nodes])])."""
    rng = np.random.default_rng(seed)
    x0 = (X - 2 * (balls - 1)) // 2
    y0 = (Y - 2 * (balls - 1)) // 2
    ball = [(x0 + 2 * i, y0 + 2 * j) for j in range(balls) for i in range(balls)]
    per = _perimeter(X, Y)
    bsel = rng.choice(len(ball), n_nets, replace=False)
    psel = rng.choice(len(per), n_nets, replace=False)
    free_b = np.setdiff1d(np.arange(len(ball)), bsel)
    free_p = np.setdiff1d(np.arange(len(per)), psel)
    mb = rng.choice(free_b, n_multi, replace=False)
    k = rng.integers(2, 8, n_multi)
    mp = rng.choice(free_p, int(k.sum()), replace=False)
    nets, off = [], 0
    for i in range(n_multi):
        pins = [_node(X, Y, *ball[mb[i]])] + [_node(X, Y, *per[q]) for q in mp[off:off + k[i]]]
        off += int(k[i])
        nets.append((f"M{i:04d}", pins))
    return X, Y, L, nets


def cfg4(seed=4, X=1000, Y=800, L=24, n_nets=40000):
    rng = np.random.default_rng(seed)
    used = set()
    reqs = []
    while len(reqs) < n_nets:
        cx, cy = rng.integers(0, X), rng.integers(0, Y)
        hx, hy = rng.lognormal(np.log(20.0), 1.0, 2)
        x1, x2 = int(np.clip(cx - hx, 0, X - 1)), int(np.clip(cx + hx, 0, X - 1))
        y1, y2 = int(np.clip(cy - hy, 0, Y - 1)), int(np.clip(cy + hy, 0, Y - 1))
        s, d = _node(X, Y, x1, y1), _node(X, Y, x2, y2)
        if s == d or s in used or d in used:
            continue
        used.add(s); used.add(d)
        reqs.append((f"S{len(reqs):05d}", s, d))
    return X, Y, L, reqs


CONFIGS = {"cfg1": cfg1, "cfg2": cfg2, "cfg3": cfg3, "cfg4": cfg4}


def cfg5(seed=5, n_jobs=64, mean_gap_s=0.5, n_min=1e5, n_max=2e7, l_min=4, l_max=24, nets_min=500,
         nets_max=30000, net_density=None, max_iterations=40, policy="adjacent"):
    """Serving jobs (serving.Job), arrival times in seconds from the start."""
    from .serving import Job
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(mean_gap_s, n_jobs)
    arrivals = np.concatenate([[0.0], np.cumsum(gaps[1:])])
    jobs = []
    for j in range(n_jobs):
        n = float(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        L = int(rng.integers(l_min, l_max + 1))
        X = Y = max(8, int(round(np.sqrt(n / L))))
        if net_density is None:
            k = int(round(np.exp(rng.uniform(np.log(nets_min), np.log(nets_max)))))
        else:
            k = int(round(net_density * X * Y / (X + Y)))
        k = max(1, min(k, X * Y // 2))
        _, _, _, reqs = cfg1(seed=seed * 1000 + j, X=X, Y=Y, L=L, n_nets=k)
        reqs = [(f"J{j:02d}_{nm}", s_, d_) for nm, s_, d_ in reqs]
        jobs.append(Job(j, X, Y, L, reqs, float(arrivals[j]), policy=policy, max_iterations=max_iterations))
    return jobs
