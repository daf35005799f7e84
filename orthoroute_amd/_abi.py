"""ctypes binding of liborgpu.so (include/orgpu.h).

The library is built in-tree by ``python -m orthoroute_amd.build``.  There is
no fallback: every compute entry point needs the HIP library and a GPU, and
raises if either is missing.
"""
import ctypes
import os

from .build import LIB_PATH

ORGPU_MAX_LAYERS = 64

ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ECAPACITY"}
ORGPU_EINVAL, ORGPU_ECAPACITY = -1, -4
REF_NOALLOC_BYTES = 64 << 20   # ORGPU_REF_NOALLOC_BYTES

# symbols include/orgpu.h declares (checked by tests/test_abi.py)
EXPORTED = [
    "orgpu_create", "orgpu_destroy", "orgpu_last_error", "orgpu_sync",
    "orgpu_lattice_count", "orgpu_lattice_build",
    "orgpu_reserve", "orgpu_route", "orgpu_roi_bfs",
    "orgpu_usage_scatter", "orgpu_present_build", "orgpu_present_ema", "orgpu_cost_update",
    "orgpu_history_update", "orgpu_overuse_sparse", "orgpu_cost_bounds", "orgpu_route_ms",
    "orgpu_set_roi_scratch", "orgpu_roi_lattice", "orgpu_reserve_ref", "orgpu_route_ref",
    "orgpu_route_ref_noalloc", "orgpu_segment_sums", "orgpu_via_pressure",
    "orgpu_reserve_solo", "orgpu_route_async",
]


class OrgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"orgpu error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class LatticeDesc(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int32), ("Y", ctypes.c_int32), ("L", ctypes.c_int32),
                ("pitch", ctypes.c_float), ("via_cost", ctypes.c_double),
                ("planar", ctypes.c_uint8 * ORGPU_MAX_LAYERS),
                ("dir", ctypes.c_uint8 * ORGPU_MAX_LAYERS),
                ("npairs", ctypes.c_int32),
                ("pair_from", ctypes.c_void_p), ("pair_to", ctypes.c_void_p)]


class Graph(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("n_edges", ctypes.c_int64),
                ("indptr", ctypes.c_void_p), ("indices", ctypes.c_void_p),
                ("cost", ctypes.c_void_p), ("edge_kind", ctypes.c_void_p),
                ("X", ctypes.c_int32), ("Y", ctypes.c_int32), ("L", ctypes.c_int32)]


class NetReq(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int32), ("dst", ctypes.c_int32), ("net_id", ctypes.c_int32),
                ("roi_index", ctypes.c_int32), ("use_keepout", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class MsPoint(ctypes.Structure):
    _fields_ = [("node", ctypes.c_int32), ("cost", ctypes.c_float)]


class MsReq(ctypes.Structure):
    _fields_ = [("seed_off", ctypes.c_int32), ("seed_n", ctypes.c_int32),
                ("target_off", ctypes.c_int32), ("target_n", ctypes.c_int32),
                ("net_id", ctypes.c_int32), ("roi_index", ctypes.c_int32),
                ("use_keepout", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class SsspParams(ctypes.Structure):
    _fields_ = [("delta", ctypes.c_float), ("n_buckets", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("rounds_per_poll", ctypes.c_int32),
                ("grid", ctypes.c_int32), ("count_minimal", ctypes.c_int32),
                ("h_xy", ctypes.c_float), ("h_z", ctypes.c_float), ("local_batches", ctypes.c_int32),
                ("launch_mode", ctypes.c_int32), ("engine", ctypes.c_int32), ("list_cap", ctypes.c_int32),
                ("lookahead", ctypes.c_int32), ("waves", ctypes.c_int32)]


class SsspStats(ctypes.Structure):
    _fields_ = [("rounds", ctypes.c_int64), ("items", ctypes.c_int64), ("edges", ctypes.c_int64),
                ("minimal_nodes", ctypes.c_int64), ("minimal_edges", ctypes.c_int64),
                ("kernel_ms", ctypes.c_double), ("launches", ctypes.c_int64), ("round_ms", ctypes.c_double),
                ("retried", ctypes.c_int64)]


class RoiReq(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int32), ("dst", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("roi_index", ctypes.c_int32), ("max_nodes", ctypes.c_int64), ("readd", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("keep", ctypes.c_int32 * 2)]


class RoiInfo(ctypes.Structure):
    _fields_ = [("seen", ctypes.c_int64), ("kept", ctypes.c_int64), ("depth", ctypes.c_int32),
                ("met", ctypes.c_int32), ("src_in", ctypes.c_int32), ("dst_in", ctypes.c_int32),
                ("front", ctypes.c_int32 * 4), ("left", ctypes.c_int32 * 4), ("appended", ctypes.c_int32 * 4)]


class RefReq(ctypes.Structure):
    _fields_ = [("src", ctypes.c_int32), ("dst", ctypes.c_int32), ("net_id", ctypes.c_int32),
                ("roi_index", ctypes.c_int32), ("use_keepout", ctypes.c_int32), ("n_rank", ctypes.c_int32),
                ("rank_node", ctypes.c_int32 * 8), ("rank_key", ctypes.c_uint32 * 8)]


class SparseOut(ctypes.Structure):
    _fields_ = [("idx", ctypes.c_void_p), ("val", ctypes.c_void_p), ("capacity", ctypes.c_int64),
                ("count", ctypes.c_int64)]


class CostParams(ctypes.Structure):
    _fields_ = [("X", ctypes.c_int32), ("Y", ctypes.c_int32), ("L", ctypes.c_int32),
                ("via_cost_mult", ctypes.c_float), ("base_cost_weight", ctypes.c_float),
                ("pres_fac", ctypes.c_float), ("hist_weight", ctypes.c_float),
                ("jitter_scale", ctypes.c_float), ("add_jitter", ctypes.c_int32),
                ("layer_bias", ctypes.c_void_p), ("via_pooling", ctypes.c_int32),
                ("seg_z", ctypes.c_int32), ("col_pres", ctypes.c_void_p), ("seg_prefix", ctypes.c_void_p),
                ("col_weight", ctypes.c_float), ("seg_weight", ctypes.c_float),
                ("via_discount", ctypes.c_void_p), ("discount_f64", ctypes.c_int32),
                ("hard_block", ctypes.c_int32), ("col_use", ctypes.c_void_p), ("seg_use", ctypes.c_void_p),
                ("col_cap", ctypes.c_int32), ("seg_cap", ctypes.c_int32)]


_lib = None


def lib_path():
    return LIB_PATH


def load():
    """Load liborgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m orthoroute_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P, i32, i64, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float
    L.orgpu_last_error.restype = ctypes.c_char_p
    L.orgpu_last_error.argtypes = []
    sigs = {
        "orgpu_create": [ctypes.c_int, P, ctypes.POINTER(P)],
        "orgpu_destroy": [P],
        "orgpu_sync": [P],
        "orgpu_lattice_count": [ctypes.POINTER(LatticeDesc), ctypes.POINTER(i64), ctypes.POINTER(i64)],
        "orgpu_lattice_build": [P, ctypes.POINTER(LatticeDesc), P, P, P, P, P],
        "orgpu_reserve": [P, i64, i32, i32, i64, i32],
        "orgpu_route": [P, ctypes.POINTER(Graph), P, i32, P, i32, P, f32, ctypes.POINTER(SsspParams),
                        P, P, i64, P, P, ctypes.POINTER(SsspStats)],
        "orgpu_route_ms": [P, ctypes.POINTER(Graph), P, i32, P, i64, P, i64, P, i32, P, f32,
                           ctypes.POINTER(SsspParams), P, P, i64, P, P, P, ctypes.POINTER(SsspStats)],
        "orgpu_roi_bfs": [P, ctypes.POINTER(Graph), P, i32, P, i32, P],
        "orgpu_reserve_ref": [P, i64, i32, i64],
        "orgpu_route_ref": [P, ctypes.POINTER(Graph), P, i32, P, i32, P, f32, i32, P, P, i64, P, P, P, P],
        "orgpu_route_ref_noalloc": [ctypes.c_int, P, ctypes.POINTER(Graph), P, i32, P, i32, P, f32, i32, i64,
                                    P, P, i64, P, P, P, P],
        "orgpu_set_roi_scratch": [P, i64],
        "orgpu_roi_lattice": [P, ctypes.POINTER(LatticeDesc), P, i32, P, i32, P],
        "orgpu_usage_scatter": [P, P, i64, P, i64, i32],
        "orgpu_present_build": [P, i64, P, P, P, i64, f32],
        "orgpu_present_ema": [P, i64, P, P, ctypes.c_int, f32, f32, ctypes.POINTER(SparseOut)],
        "orgpu_cost_update": [P, ctypes.POINTER(CostParams), i64, i64, P, P, P, P, P, P, P],
        "orgpu_cost_bounds": [P, ctypes.POINTER(Graph), P, P, P],
        "orgpu_history_update": [P, i64, P, P, P, f32, f32, f32, ctypes.c_int],
        "orgpu_overuse_sparse": [P, i64, P, P, ctypes.c_int, f32, f32, ctypes.POINTER(SparseOut)],
        "orgpu_segment_sums": [P, i64, P, P, P, i64, P],
        "orgpu_via_pressure": [P, i64, i32, P, i32, P, i32, P, P, P, P],
        "orgpu_reserve_solo": [P, i64, i32, i32, i32, i32, i32],
        "orgpu_route_async": [P, ctypes.POINTER(Graph), P, i32, P, i32, P, f32, ctypes.POINTER(SsspParams),
                              P, P, i64, P, P, P, P],
    }
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        raise OrgpuError(rc, load().orgpu_last_error().decode(errors="replace"))
    return rc
