"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, and only as the checker (or as the timed restatement of the reference's
MULTI_CORE engine).  The product package never imports this module.
Build with `make -C oracle`.
"""
import ctypes
import os

import numpy as np

UNREACHED = 999_999_999
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u32, u64, i64, dbl, i32 = (ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.c_double, ctypes.c_int)
        sig = {
            "or_mix64": (u64, [u64]),
            "or_rmat_edges": (None, [u32, u64, u64, dbl, dbl, dbl, u64, _u32p, _u32p]),
            "or_edge_weight": (u32, [u64, u32, u32, u32, u32, i32]),
            "or_grid_edges": (u64, [u32, dbl, u64, _u32p, _u32p]),
            "or_csr_build": (i64, [u32, u64, _u32p, _u32p, i32, _u64p, _u32p, u64]),
            "or_fill_weights": (None, [u32, _u64p, _u32p, u64, u32, u32, i32, _u32p]),
            "or_transpose": (None, [u32, _u64p, _u32p, _vp, _u64p, _u32p, _vp]),
            "or_bfs": (None, [u32, _u64p, _u32p, u32, _u32p]),
            "or_dijkstra": (None, [u32, _u64p, _u32p, _u32p, u32, _u32p]),
            "or_cc_undirected": (None, [u32, _u64p, _u32p, _u32p]),
            "or_cc_directed": (None, [u32, _u64p, _u32p, _u32p]),
            "or_pagerank64": (None, [u32, _u64p, _u32p, i32, dbl, _f64p]),
            "or_pagerank_step": (None, [u32, _u64p, _u32p, np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS"),
                                        dbl, _f64p]),
            "or_check_sssp_cert": (i64, [u32, _u64p, _u32p, _vp, u32, _u32p]),
            "or_reached_edges": (u64, [u32, _u64p, _u32p]),
            "or_max_outdeg_vertex": (u32, [u32, _u64p]),
            "or_multicore_run": (dbl, [i32, u32, _u64p, _u32p, _vp, u32, i32, _u32p,
                                       ctypes.POINTER(u32)]),
            "or_omp_max_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class CSR:
    """Host CSR with u64 offsets, u32 destinations and optional u32 weights."""

    def __init__(self, offsets, dest, weight=None):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.dest = np.ascontiguousarray(dest, dtype=np.uint32)
        self.weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.uint32)
        self.V = len(self.offsets) - 1
        self.E = int(self.offsets[-1])


def rmat_edges(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, first=0, count=None):
    m = (1 << scale) * edge_factor
    if count is None:
        count = m - first
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    lib().or_rmat_edges(scale, first, count, a, b, c, seed, src, dst)
    return src, dst


def csr_build(V, src, dst, symmetrize=False):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    cap = len(src) * (2 if symmetrize else 1)
    off = np.empty(V + 1, np.uint64)
    out = np.empty(max(cap, 1), np.uint32)
    E = lib().or_csr_build(V, len(src), src, dst, int(symmetrize), off, out, cap)
    assert E >= 0
    return CSR(off, out[:E].copy())


def fill_weights(g, seed, lo, hi, symmetric=False):
    w = np.empty(max(g.E, 1), np.uint32)
    lib().or_fill_weights(g.V, g.offsets, g.dest, seed, lo, hi, int(symmetric), w)
    g.weight = w[:g.E].copy()
    return g


def rmat_graph(scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, symmetrize=False,
               weights=(1, 100)):
    """The R-MAT workload: self-loops and duplicates removed, weights U[lo,hi]."""
    src, dst = rmat_edges(scale, edge_factor, a, b, c, seed)
    g = csr_build(1 << scale, src, dst, symmetrize)
    if weights is not None:
        fill_weights(g, seed, weights[0], weights[1], symmetric=False)
    return g


def grid_graph(n, drop=0.1, seed=1, weights=(1, 1000)):
    cap = 4 * n * (n - 1)
    src = np.empty(max(cap, 1), np.uint32)
    dst = np.empty(max(cap, 1), np.uint32)
    m = lib().or_grid_edges(n, drop, seed, src, dst)
    g = csr_build(n * n, src[:m], dst[:m], False)
    if weights is not None:
        fill_weights(g, seed, weights[0], weights[1], symmetric=True)
    return g


def transpose(g):
    in_off = np.empty(g.V + 1, np.uint64)
    in_src = np.empty(max(g.E, 1), np.uint32)
    in_w = np.empty(max(g.E, 1), np.uint32) if g.weight is not None else None
    lib().or_transpose(g.V, g.offsets, g.dest, _ptr(g.weight), in_off, in_src, _ptr(in_w))
    return CSR(in_off, in_src[:g.E].copy(), None if in_w is None else in_w[:g.E].copy())


def bfs(g, root):
    out = np.empty(g.V, np.uint32)
    lib().or_bfs(g.V, g.offsets, g.dest, root, out)
    return out


def sssp(g, root):
    out = np.empty(g.V, np.uint32)
    lib().or_dijkstra(g.V, g.offsets, g.dest, g.weight, root, out)
    return out


def cc_undirected(g):
    out = np.empty(g.V, np.uint32)
    lib().or_cc_undirected(g.V, g.offsets, g.dest, out)
    return out


def cc_directed(g):
    out = np.empty(g.V, np.uint32)
    lib().or_cc_directed(g.V, g.offsets, g.dest, out)
    return out


def pagerank64(g, iters=20, damping=0.85):
    out = np.empty(g.V, np.float64)
    lib().or_pagerank64(g.V, g.offsets, g.dest, iters, damping, out)
    return out


def pagerank_step(g, ranks32, damping=0.85):
    out = np.empty(g.V, np.float64)
    lib().or_pagerank_step(g.V, g.offsets, g.dest, np.ascontiguousarray(ranks32, np.float32), damping, out)
    return out


def check_sssp_cert(g, root, dist, unit=False):
    d = np.ascontiguousarray(dist, np.uint32)
    return int(lib().or_check_sssp_cert(g.V, g.offsets, g.dest,
                                        None if unit else _ptr(g.weight), root, d))


def reached_edges(g, val):
    return int(lib().or_reached_edges(g.V, g.offsets, np.ascontiguousarray(val, np.uint32)))


def max_outdeg_vertex(g):
    return int(lib().or_max_outdeg_vertex(g.V, g.offsets))


def multicore_run(alg, g, root=0, threads=0):
    """Restated reference MULTI_CORE engine; returns (values, ms, iterations)."""
    out = np.empty(g.V, np.uint32)
    it = ctypes.c_uint32(0)
    ms = lib().or_multicore_run(alg, g.V, g.offsets, g.dest, _ptr(g.weight), root, threads,
                                out, ctypes.byref(it))
    return out, ms, it.value


def omp_max_threads():
    """omp_get_max_threads() of the restated engine's OpenMP runtime."""
    return int(lib().or_omp_max_threads())
