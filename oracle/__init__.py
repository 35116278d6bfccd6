"""CPU oracle for the superstep engine — TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker (or the timed CPU baseline); the product path never imports it.  Each
function wraps a C++ routine in oracle.cpp whose comment cites the reference
file:line it restates.

Pins (tests/test_oracle.py): closed forms (paths, cycles, stars, grids, complete
graphs), invariants (BFS level/edge relations, SSSP triangle inequality and tight
parent edges, PageRank mass conservation), brute force on tiny graphs (Bellman-Ford,
exhaustive path enumeration), hand-worked golden fixtures under tests/golden/, and
agreement between the textbook and reference-path restatements.
"""
import ctypes
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32, _u64, _int, _f32, _f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_double


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            from . import build as _b
            _b.build()
        L = ctypes.CDLL(str(path))
        sigs = {
            "orc_csc": [_u32, _u64, _u32p, _u32p, _u32p, _u32p, _u32p],
            "orc_segmented_scan": [_u64, _f32p, _u32p, _int, _f32p],
            "orc_bfs_queue": [_u32, _u32p, _u32p, _u32, _i32p],
            "orc_bfs_sweep": [_u32, _u64, _u32p, _u32p, _u32, _i32p, _int],
            "orc_mbfs": [_u32, _u32p, _u32p, _u32p, _u32, _u8p, _int],
            "orc_sssp_dijkstra": [_u32, _u32p, _u32p, _u32p, _u32, _u32p],
            "orc_sssp_bellman_ford": [_u32, _u32p, _u32p, _u32p, _u32, _u32p],
            "orc_sssp_bsp": [_u32, _u64, _u32p, _u32p, _u32p, _u32, _u32p, _int, _int],
            "orc_sssp_rounds": [_u32, _u32p, _u32p, _u32p, _u32, _u32p],
            "orc_pagerank_f64": [_u32, _u64, _u32p, _u32p, _f64, _u32, _f64p, _int],
            "orc_pagerank_f32": [_u32, _u64, _u32p, _u32p, _f32, _u32, _f32p, _int],
            "orc_pagerank_ref": [_u32, _u64, _u32p, _u32p, _u32, _f32p, _int],
            "orc_mlp_loss_grad": [_f32p, _f32p, _f32p, _u32, _f64p, _f64p],
            "orc_mlp_forward": [_f32p, _f32p, _u32, _f64p],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _csr(row_ptr, col):
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    return row_ptr, col, len(row_ptr) - 1, len(col)


def csc(row_ptr, col):
    row_ptr, col, n, m = _csr(row_ptr, col)
    cptr = np.zeros(n + 1, np.uint32)
    csrc = np.zeros(max(m, 1), np.uint32)
    slot = np.zeros(max(m, 1), np.uint32)
    lib().orc_csc(n, m, row_ptr, col, cptr, csrc, slot)
    return cptr, csrc[:m], slot[:m]


def segmented_scan(values, flags, op="add"):
    values = np.ascontiguousarray(values, dtype=np.float32)
    flags = np.ascontiguousarray(flags, dtype=np.uint32)
    out = np.zeros_like(values)
    lib().orc_segmented_scan(len(values), values, flags, 0 if op == "add" else 1, out)
    return out


def bfs_queue(row_ptr, col, src):
    row_ptr, col, n, m = _csr(row_ptr, col)
    level = np.zeros(n, np.int32)
    lib().orc_bfs_queue(n, row_ptr, col, src, level)
    return level


def bfs_sweep(row_ptr, col, src, threads=0):
    """Reference CPU path; returns (levels, supersteps)."""
    row_ptr, col, n, m = _csr(row_ptr, col)
    level = np.zeros(n, np.int32)
    steps = lib().orc_bfs_sweep(n, m, row_ptr, col, src, level, threads)
    return level, steps


def mbfs(row_ptr, col, sources, threads=0):
    row_ptr, col, n, m = _csr(row_ptr, col)
    sources = np.ascontiguousarray(sources, dtype=np.uint32)
    out = np.zeros((len(sources), n), np.uint8)
    lib().orc_mbfs(n, row_ptr, col, sources, len(sources), out, threads)
    return out


def sssp_dijkstra(row_ptr, col, w, src):
    row_ptr, col, n, m = _csr(row_ptr, col)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    dist = np.zeros(n, np.uint32)
    rc = lib().orc_sssp_dijkstra(n, row_ptr, col, w, src, dist)
    if rc != 0:
        raise OverflowError("shortest distance exceeds 32 bits")
    return dist


def sssp_bellman_ford(row_ptr, col, w, src):
    row_ptr, col, n, m = _csr(row_ptr, col)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    dist = np.zeros(n, np.uint32)
    lib().orc_sssp_bellman_ford(n, row_ptr, col, w, src, dist)
    return dist


def sssp_bsp(row_ptr, col, w, src, threads=0, max_steps=0):
    """Reference CPU path; returns (dist, supersteps).  max_steps > 0 stops early."""
    row_ptr, col, n, m = _csr(row_ptr, col)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    dist = np.zeros(n, np.uint32)
    steps = lib().orc_sssp_bsp(n, m, row_ptr, col, w, src, dist, threads, max_steps)
    return dist, steps


def sssp_rounds(row_ptr, col, w, src, dist):
    """Supersteps the reference CPU path needs, derived from final distances."""
    row_ptr, col, n, m = _csr(row_ptr, col)
    w = np.ascontiguousarray(w, dtype=np.uint32)
    dist = np.ascontiguousarray(dist, dtype=np.uint32)
    return lib().orc_sssp_rounds(n, row_ptr, col, w, src, dist)


def pagerank_f64(row_ptr, col, d=0.85, steps=30, threads=0):
    row_ptr, col, n, m = _csr(row_ptr, col)
    out = np.zeros(n, np.float64)
    lib().orc_pagerank_f64(n, m, row_ptr, col, d, steps, out, threads)
    return out


def pagerank_f32(row_ptr, col, d=0.85, steps=30, threads=0):
    row_ptr, col, n, m = _csr(row_ptr, col)
    out = np.zeros(n, np.float32)
    lib().orc_pagerank_f32(n, m, row_ptr, col, d, steps, out, threads)
    return out


def pagerank_ref(row_ptr, col, steps=100, threads=0):
    row_ptr, col, n, m = _csr(row_ptr, col)
    out = np.zeros(n, np.float32)
    lib().orc_pagerank_ref(n, m, row_ptr, col, steps, out, threads)
    return out


MLP_PARAMS = 1633


def mlp_loss_grad(params, X, y):
    """FP64 loss and gradient of the scheduler MLP (mean squared error)."""
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    g = np.zeros(MLP_PARAMS, np.float64)
    loss = np.zeros(1, np.float64)
    lib().orc_mlp_loss_grad(params, X, y, len(y), g, loss)
    return float(loss[0]), g


def mlp_forward(params, X):
    params = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(len(X), np.float64)
    lib().orc_mlp_forward(params, X, len(X), out)
    return out
