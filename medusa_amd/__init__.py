"""Python binding of the MI355X superstep engine's C ABI (include/medusa_c.h).

This is a thin ctypes layer used by the tests and the benchmark; the engine
itself is libmedusa_amd.so (HIP kernels for gfx950 + C++ host code).  Nothing
here computes a result on the CPU: if the library is missing or no GPU is
visible, the calls raise.
"""
import ctypes
import math
import pathlib

import numpy as np

__all__ = [
    "lib", "MedusaError", "Opts", "Stats", "Graph", "csr_from_arcs", "gen_erdos_renyi", "gen_rmat",
    "gen_lattice", "gen_chung_lu", "gen_weights", "device_count", "DIR_AUTO", "DIR_PUSH", "DIR_PULL",
    "PR_DEFAULT", "PR_PULL", "PR_PUSH", "PR_BINNED", "LEVEL_UNREACHED", "DIST_UNREACHED", "MBFS_UNREACHED", "TRACE_FIELDS",
]

_HERE = pathlib.Path(__file__).resolve().parent
_LIBPATH = _HERE / "libmedusa_amd.so"

DIR_AUTO, DIR_PUSH, DIR_PULL = 0, 1, 2
PR_DEFAULT, PR_PUSH, PR_BINNED, PR_PULL = 0, 1, 2, 3
LEVEL_UNREACHED = -1
DIST_UNREACHED = 0xFFFFFFFF
MBFS_UNREACHED = 0xFF
TRACE_FIELDS = 9

EXPORTED = [
    "medusa_last_error", "medusa_version", "medusa_device_count", "medusa_graph_create", "medusa_graph_destroy",
    "medusa_graph_info", "medusa_graph_device_arrays", "medusa_bfs", "medusa_sssp", "medusa_pagerank",
    "medusa_mbfs", "medusa_trace_count", "medusa_trace_read", "medusa_gen_erdos_renyi", "medusa_gen_rmat",
    "medusa_gen_lattice", "medusa_gen_chung_lu", "medusa_gen_weights", "medusa_csr_from_arcs",
    "medusa_policy_create", "medusa_policy_destroy", "medusa_policy_params", "medusa_policy_train",
    "medusa_policy_predict", "medusa_policy_gradients", "medusa_policy_features", "medusa_graph_set_batch",
    "medusa_graph_batch", "medusa_bfs_multi", "medusa_autotune", "medusa_bfs_noalloc",
]
TUNE_BFS, TUNE_MBFS = 1, 2
POLICY_INPUTS = 16
POLICY_PARAMS = 1633
POLICY_DIRECTION, POLICY_DELTA, POLICY_VARIANT = 1, 2, 3


class MedusaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"medusa error {code}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [
        ("timing", ctypes.c_uint32),
        ("direction", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("delta", ctypes.c_uint32),
        ("large_degree", ctypes.c_uint32),
        ("trace", ctypes.c_uint32),
        ("policy", ctypes.c_void_p),
        ("use_graph", ctypes.c_uint32),
        ("grid", ctypes.c_uint32),
        ("far_limit", ctypes.c_uint32),
        ("local_rounds", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 4),
    ]


class Tuning(ctypes.Structure):
    _fields_ = [
        ("large_degree", ctypes.c_uint32),
        ("grid", ctypes.c_uint32),
        ("candidates", ctypes.c_uint32),
        ("reserved0", ctypes.c_uint32),
        ("best_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("reserved", ctypes.c_uint32 * 4),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("device_ms", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("kernel_launches", ctypes.c_uint64),
        ("kernel_bytes", ctypes.c_uint64),
        ("alg_bytes", ctypes.c_uint64),
        ("edges_traversed", ctypes.c_uint64),
        ("supersteps", ctypes.c_uint32),
        ("push_steps", ctypes.c_uint32),
        ("pull_steps", ctypes.c_uint32),
        ("reached", ctypes.c_uint32),
        ("residual_l1", ctypes.c_double),
        ("reserved", ctypes.c_uint32 * 8),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "reserved"}


_LIB = None
_vp = ctypes.c_void_p
_u32, _u64, _i32, _f32, _f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float, ctypes.c_double


def lib():
    """Loads libmedusa_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not _LIBPATH.exists():
            raise MedusaError(-5, f"{_LIBPATH} not built (run medusa_amd/build.py)")
        # One HIP runtime per process: when PyTorch is installed, load it first so the
        # library's libamdhip64.so.7 dependency binds to the runtime torch already holds
        # (two runtimes in one process leave one of them without a GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(_LIBPATH))
        sig = {
            "medusa_last_error": ([], ctypes.c_char_p),
            "medusa_version": ([], ctypes.c_char_p),
            "medusa_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
            "medusa_graph_create": ([_u32, _u64, _vp, _vp, _vp, ctypes.POINTER(_vp)], ctypes.c_int),
            "medusa_graph_destroy": ([_vp], ctypes.c_int),
            "medusa_graph_info": ([_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64), ctypes.POINTER(_u32)],
                                  ctypes.c_int),
            "medusa_graph_device_arrays": ([_vp] + [ctypes.POINTER(_vp)] * 6, ctypes.c_int),
            "medusa_bfs": ([_vp, _u32, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Stats)], ctypes.c_int),
            "medusa_sssp": ([_vp, _u32, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Stats)], ctypes.c_int),
            "medusa_pagerank": ([_vp, _f32, _u32, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Stats)], ctypes.c_int),
            "medusa_mbfs": ([_vp, _vp, _u32, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Stats)], ctypes.c_int),
            "medusa_trace_count": ([_vp, ctypes.POINTER(_u32)], ctypes.c_int),
            "medusa_trace_read": ([_vp, _vp, _u32], ctypes.c_int),
            "medusa_gen_erdos_renyi": ([_u32, _u64, _u64, _u64, _vp, _vp, ctypes.POINTER(_u64)], ctypes.c_int),
            "medusa_gen_rmat": ([_u32, _u64, _f64, _f64, _f64, _f64, _u64, _u64, _vp, _vp, ctypes.POINTER(_u64)],
                                ctypes.c_int),
            "medusa_gen_lattice": ([_u32, _u32, _f64, _f64, _u32, _u32, _u64, _u64, _vp, _vp, _vp,
                                    ctypes.POINTER(_u64)], ctypes.c_int),
            "medusa_gen_chung_lu": ([_u32, _u64, _f64, _u32, _f64, _u64, _u64, _vp, _vp, ctypes.POINTER(_u64)],
                                    ctypes.c_int),
            "medusa_gen_weights": ([_u64, _u32, _u32, _u64, _vp], ctypes.c_int),
            "medusa_csr_from_arcs": ([_u32, _u64, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
            "medusa_policy_create": ([_u32, _u64, ctypes.POINTER(_vp)], ctypes.c_int),
            "medusa_policy_destroy": ([_vp], ctypes.c_int),
            "medusa_policy_params": ([_vp, _vp, _vp, _u32], ctypes.c_int),
            "medusa_policy_train": ([_vp, _vp, _vp, _u32, _u32, _f32, ctypes.POINTER(_f32)], ctypes.c_int),
            "medusa_policy_predict": ([_vp, _vp, _u32, _vp], ctypes.c_int),
            "medusa_policy_gradients": ([_vp, _vp, _vp, _u32, _vp, ctypes.POINTER(_f32)], ctypes.c_int),
            "medusa_policy_features": ([_u32, _vp, _f64, _vp], ctypes.c_int),
            "medusa_graph_set_batch": ([_vp, _u32, _vp], ctypes.c_int),
            "medusa_graph_batch": ([_vp, ctypes.POINTER(_u32), _vp, _u32], ctypes.c_int),
            "medusa_bfs_multi": ([_vp, _vp, _u32, _vp, ctypes.POINTER(Opts), ctypes.POINTER(Stats)], ctypes.c_int),
            "medusa_autotune": ([_vp, _u32, _vp, _u32, _u32, ctypes.POINTER(Tuning)], ctypes.c_int),
            "medusa_bfs_noalloc": ([_vp, _u32, _vp, ctypes.POINTER(Stats)], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise MedusaError(rc, lib().medusa_last_error().decode())


def _ptr(a):
    """Address of a numpy array or a torch tensor (host or device)."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return int(a.data_ptr())


def device_count():
    c = ctypes.c_int(0)
    _check(lib().medusa_device_count(ctypes.byref(c)))
    return c.value


def _u32arr(x):
    return np.ascontiguousarray(x, dtype=np.uint32)


def gen_erdos_renyi(n, m, seed):
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    mo = _u64(0)
    _check(lib().medusa_gen_erdos_renyi(n, m, seed, m, _ptr(src), _ptr(dst), ctypes.byref(mo)))
    return src[: mo.value], dst[: mo.value]


def gen_rmat(scale, m, a=0.45, b=0.15, c=0.15, noise=0.1, seed=2):
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    mo = _u64(0)
    _check(lib().medusa_gen_rmat(scale, m, a, b, c, noise, seed, m, _ptr(src), _ptr(dst), ctypes.byref(mo)))
    return src, dst


def gen_lattice(rows, cols, p_keep=0.7, p_diag=0.02, wmin=1, wmax=1000, seed=3):
    cap = 6 * rows * cols + 6
    src = np.empty(cap, np.uint32)
    dst = np.empty(cap, np.uint32)
    w = np.empty(cap, np.uint32)
    mo = _u64(0)
    _check(lib().medusa_gen_lattice(rows, cols, p_keep, p_diag, wmin, wmax, seed, cap, _ptr(src), _ptr(dst),
                                    _ptr(w), ctypes.byref(mo)))
    k = mo.value
    return src[:k].copy(), dst[:k].copy(), w[:k].copy()


def gen_chung_lu(n, m, exponent=2.0, max_degree=100_000, p_zero=0.4, seed=5):
    src = np.empty(m, np.uint32)
    dst = np.empty(m, np.uint32)
    mo = _u64(0)
    _check(lib().medusa_gen_chung_lu(n, m, exponent, max_degree, p_zero, seed, m, _ptr(src), _ptr(dst),
                                     ctypes.byref(mo)))
    return src, dst


def gen_weights(m, wmin=1, wmax=1000, seed=0):
    w = np.empty(m, np.uint32)
    _check(lib().medusa_gen_weights(m, wmin, wmax, seed, _ptr(w)))
    return w


def csr_from_arcs(n, src, dst, w=None):
    """Stable CSR from an arc list: returns (row_ptr, col, weight-or-None)."""
    src, dst = _u32arr(src), _u32arr(dst)
    m = len(src)
    row_ptr = np.empty(n + 1, np.uint32)
    col = np.empty(max(m, 1), np.uint32)
    wo = None
    if w is not None:
        w = _u32arr(w)
        wo = np.empty(max(m, 1), np.uint32)
    _check(lib().medusa_csr_from_arcs(n, m, _ptr(src), _ptr(dst), _ptr(w), _ptr(row_ptr), _ptr(col), _ptr(wo)))
    return row_ptr, col[:m], (wo[:m] if wo is not None else None)


def pack_graphs(graphs):
    """Disjoint union of CSR graphs [(row_ptr, col[, weight]), ...] (host only): returns
    (row_ptr, col, weight-or-None, graph_begin) with graph k's vertex v at graph_begin[k] + v."""
    rps, cols, ws, begin = [], [], [], [0]
    e0 = 0
    for gr in graphs:
        rp, col = _u32arr(gr[0]), _u32arr(gr[1])
        n = len(rp) - 1
        rps.append(rp[:-1].astype(np.uint64) + e0)
        cols.append(col + np.uint32(begin[-1]))
        ws.append(_u32arr(gr[2]) if len(gr) > 2 and gr[2] is not None else None)
        e0 += len(col)
        begin.append(begin[-1] + n)
    if begin[-1] >= 2 ** 32 or e0 >= 2 ** 32:
        raise MedusaError(-1, "packed batch exceeds 32-bit ids")
    row_ptr = np.concatenate(rps + [np.array([e0], np.uint64)]).astype(np.uint32)
    col = np.concatenate(cols).astype(np.uint32) if cols else np.zeros(0, np.uint32)
    w = None
    if any(x is not None for x in ws):
        if not all(x is not None for x in ws):
            raise MedusaError(-1, "pack_graphs: either every graph or none has weights")
        w = np.concatenate(ws).astype(np.uint32)
    return row_ptr, col, w, np.array(begin, np.uint32)


def _opts(**kw):
    o = Opts()
    for k, v in kw.items():
        if k == "policy":
            o.policy = v.handle if v is not None else None
        else:
            setattr(o, k, v)
    return o


class Graph:
    """A graph resident in HBM of the current GPU (wraps medusa_graph*)."""

    def __init__(self, row_ptr, col, weight=None):
        row_ptr, col = _u32arr(row_ptr), _u32arr(col)
        self.n = len(row_ptr) - 1
        self.m = len(col)
        w = _u32arr(weight) if weight is not None else None
        h = _vp()
        _check(lib().medusa_graph_create(self.n, self.m, _ptr(row_ptr), _ptr(col) if self.m else None, _ptr(w),
                                         ctypes.byref(h)))
        self.handle = h

    @classmethod
    def packed(cls, graphs):
        """One handle holding a batch of independent graphs (see pack_graphs)."""
        row_ptr, col, w, begin = pack_graphs(graphs)
        G = cls(row_ptr, col, w)
        G.set_batch(begin)
        return G

    def set_batch(self, graph_begin):
        b = _u32arr(graph_begin)
        _check(lib().medusa_graph_set_batch(self.handle, len(b) - 1, _ptr(b)))

    def batch(self):
        """graph_begin (count+1 entries) of the packed graphs; [0, n] for a plain graph."""
        c = _u32(0)
        _check(lib().medusa_graph_batch(self.handle, ctypes.byref(c), None, 0))
        out = np.zeros(c.value + 1, np.uint32)
        _check(lib().medusa_graph_batch(self.handle, ctypes.byref(c), _ptr(out), c.value + 1))
        return out

    @classmethod
    def from_arcs(cls, n, src, dst, w=None):
        row_ptr, col, wo = csr_from_arcs(n, src, dst, w)
        return cls(row_ptr, col, wo)

    def close(self):
        if getattr(self, "handle", None):
            _check(lib().medusa_graph_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self):
        n, m, w = _u32(0), _u64(0), _u32(0)
        _check(lib().medusa_graph_info(self.handle, ctypes.byref(n), ctypes.byref(m), ctypes.byref(w)))
        return n.value, m.value, bool(w.value)

    def device_arrays(self):
        ptrs = [_vp() for _ in range(6)]
        _check(lib().medusa_graph_device_arrays(self.handle, *[ctypes.byref(p) for p in ptrs]))
        return dict(zip(["row_ptr", "col", "weight", "csc_ptr", "csc_src", "msg_slot"], [p.value for p in ptrs]))

    def bfs(self, source, out=None, **opts):
        """Returns (levels int32[n], Stats).  `out` may be a host array or a device tensor."""
        if out is None:
            out = np.empty(self.n, np.int32)
        st = Stats()
        o = _opts(**opts)
        _check(lib().medusa_bfs(self.handle, source, _ptr(out), ctypes.byref(o), ctypes.byref(st)))
        return out, st

    def bfs_noalloc(self, source, out):
        """BFS into `out` (a device tensor of n int32) with no device allocation; returns Stats."""
        st = Stats()
        _check(lib().medusa_bfs_noalloc(self.handle, source, _ptr(out), ctypes.byref(st)))
        return st

    def bfs_multi(self, sources, out=None, **opts):
        """BFS from several sources at once (all at level 0); returns (levels, Stats)."""
        sources = _u32arr(sources)
        if out is None:
            out = np.empty(self.n, np.int32)
        st = Stats()
        o = _opts(**opts)
        _check(lib().medusa_bfs_multi(self.handle, _ptr(sources), len(sources), _ptr(out), ctypes.byref(o),
                                      ctypes.byref(st)))
        return out, st

    def autotune(self, algo, sources, reps=3):
        """Times the launch-shape candidates for TUNE_BFS / TUNE_MBFS from `sources` and keeps
        the fastest in the handle; reps=0 forgets it.  Returns the Tuning record."""
        sources = _u32arr(np.atleast_1d(sources))
        t = Tuning()
        _check(lib().medusa_autotune(self.handle, algo, _ptr(sources), len(sources), reps, ctypes.byref(t)))
        return t

    def sssp(self, source, out=None, **opts):
        if out is None:
            out = np.empty(self.n, np.uint32)
        st = Stats()
        o = _opts(**opts)
        _check(lib().medusa_sssp(self.handle, source, _ptr(out), ctypes.byref(o), ctypes.byref(st)))
        return out, st

    def pagerank(self, damping=0.85, steps=30, out=None, **opts):
        if out is None:
            out = np.empty(self.n, np.float32)
        st = Stats()
        o = _opts(**opts)
        _check(lib().medusa_pagerank(self.handle, damping, steps, _ptr(out), ctypes.byref(o), ctypes.byref(st)))
        return out, st

    def mbfs(self, sources, out=None, **opts):
        sources = _u32arr(sources)
        k = len(sources)
        if out is None:
            out = np.empty((k, self.n), np.uint8)
        st = Stats()
        o = _opts(**opts)
        _check(lib().medusa_mbfs(self.handle, _ptr(sources), k, _ptr(out), ctypes.byref(o), ctypes.byref(st)))
        return out, st

    def trace(self):
        c = _u32(0)
        _check(lib().medusa_trace_count(self.handle, ctypes.byref(c)))
        out = np.zeros((c.value, TRACE_FIELDS), np.float64)
        _check(lib().medusa_trace_read(self.handle, _ptr(out), c.value))
        return out


def policy_features(kind, record, action):
    """Feature row (float32[16]) of one trace record for a candidate action (host only)."""
    rec = np.ascontiguousarray(record, np.float64)
    row = np.zeros(POLICY_INPUTS, np.float32)
    _check(lib().medusa_policy_features(kind, _ptr(rec), float(action), _ptr(row)))
    return row


def delta_samples(trace):
    """Training rows of the bucket-width policy from the trace of one fixed-width SSSP run.

    Row i pairs phase i's features and width with the run's cost from that phase to its end per
    vertex still to settle (log2 ns): a width that makes one phase cheap per vertex but leaves
    costly phases behind it, such as a huge first bucket that relaxes most vertices many times,
    is charged for those phases too.  Returns (X float32[k, 16], y float32[k])."""
    tr = np.asarray(trace, np.float64).reshape(-1, TRACE_FIELDS)
    ms_left = np.cumsum(tr[::-1, 8])[::-1]
    settled_left = np.cumsum(tr[::-1, 4])[::-1]
    keep = (tr[:, 7] > 0) & (ms_left > 0) & (settled_left > 0)
    X = np.zeros((int(keep.sum()), POLICY_INPUTS), np.float32)
    for j, r in enumerate(tr[keep]):
        X[j] = policy_features(POLICY_DELTA, r, r[7])
    y = np.log2(ms_left[keep] * 1e6 / settled_left[keep]).astype(np.float32)
    return X, y


def delta_candidates(wmean):
    """The bucket widths the width policy chooses among for a graph of (sampled) mean weight
    `wmean`, as the engine computes them: wmean * 2^k, k = -2..8, within [1, 2^32 - 2]."""
    return [int(min(4294967294.0, max(1.0, math.ldexp(float(wmean), k)))) for k in range(-2, 9)]


def delta_run_sample(trace):
    """The width policy's training row of one fixed-width SSSP run: the features at the run's
    start with its width, and the cost of the whole run per settled vertex (log2 ns), i.e. row 0
    of delta_samples.  The engine consults the policy once per run, at that point."""
    X, y = delta_samples(trace)
    return X[0], y[0]


class Policy:
    """Learned per-superstep scheduler (MLP cost model trained on the GPU)."""

    def __init__(self, kind, seed=0):
        h = _vp()
        _check(lib().medusa_policy_create(kind, seed, ctypes.byref(h)))
        self.handle = h
        self.kind = kind

    def close(self):
        if getattr(self, "handle", None):
            _check(lib().medusa_policy_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(POLICY_PARAMS, np.float32)
        _check(lib().medusa_policy_params(self.handle, _ptr(out), None, POLICY_PARAMS))
        return out

    @params.setter
    def params(self, value):
        v = np.ascontiguousarray(value, np.float32)
        _check(lib().medusa_policy_params(self.handle, None, _ptr(v), POLICY_PARAMS))

    def train(self, X, y, iters=2000, lr=3e-3):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        loss = _f32(0)
        _check(lib().medusa_policy_train(self.handle, _ptr(X), _ptr(y), len(y), iters, lr, ctypes.byref(loss)))
        return loss.value

    def predict(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros(len(X), np.float32)
        _check(lib().medusa_policy_predict(self.handle, _ptr(X), len(X), _ptr(out)))
        return out

    def gradients(self, X, y):
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(POLICY_PARAMS, np.float32)
        loss = _f32(0)
        _check(lib().medusa_policy_gradients(self.handle, _ptr(X), _ptr(y), len(y), _ptr(g), ctypes.byref(loss)))
        return g, loss.value
