"""Batched BFS/SSSP queries over graphs of uneven size (cg_batch_run), and the
mixed workload of the batched benchmark: R-MAT scales 14-22 (edge factor 16)
plus 4-neighbour lattices of 256^2 .. 2048^2 vertices with 10% of the edges
dropped, uint32 weights U[1,1000], queries spread round-robin over the graphs
with seeded uniform roots."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib
from .core import Graph

BATCHED, ONE_PER_LAUNCH = 0, 1


def run_batch(graphs, algorithms, roots, mode=BATCHED, copy_out=True, **opt):
    nq = len(roots)
    G = (ctypes.c_void_p * nq)(*[g._h.value for g in graphs])
    A = (ctypes.c_int * nq)(*[int(a) for a in algorithms])
    R = (ctypes.c_uint32 * nq)(*[int(r) for r in roots])
    outs = None
    arrays = None
    if copy_out:
        arrays = [np.empty(g.V, np.uint32) for g in graphs]
        outs = (ctypes.c_void_p * nq)(*[a.ctypes.data for a in arrays])
    o = _lib.default_options(**opt)
    st = _lib.RunStats()
    check(lib().cg_batch_run(ctypes.cast(G, ctypes.c_void_p), ctypes.cast(A, ctypes.c_void_p),
                             ctypes.cast(R, ctypes.c_void_p), nq,
                             None if outs is None else ctypes.cast(outs, ctypes.c_void_p), int(mode),
                             ctypes.byref(o), ctypes.byref(st)))
    return arrays, st.as_dict()


def run_engine_batch(engines, roots, copy_out=True):
    """cg_engine_run_batch: query i is engines[i]'s algorithm on its graph."""
    nq = len(roots)
    E = (ctypes.c_void_p * nq)(*[e._h.value for e in engines])
    R = (ctypes.c_uint32 * nq)(*[int(r) for r in roots])
    arrays, outs = None, None
    if copy_out:
        arrays = [np.empty(e.V, np.uint32) for e in engines]
        outs = (ctypes.c_void_p * nq)(*[a.ctypes.data for a in arrays])
    st = _lib.RunStats()
    check(lib().cg_engine_run_batch(ctypes.cast(E, ctypes.c_void_p), ctypes.cast(R, ctypes.c_void_p), nq,
                                    None if outs is None else ctypes.cast(outs, ctypes.c_void_p), ctypes.byref(st)))
    return arrays, st.as_dict()


def mixed_workload(n_bfs=256, n_sfs=256, rmat_scales=range(14, 23), grid_sides=(256, 512, 1024, 2048),
                   seed=1, weights=(1, 1000)):
    """Builds the graphs on the GPU and returns (graphs, per-query graph list,
    algorithms, roots)."""
    graphs = [Graph.rmat(s, 16, seed=seed, weights=weights) for s in rmat_scales]
    graphs += [Graph.grid(n, 0.1, seed=seed, weights=weights) for n in grid_sides]
    Vs = [g.V for g in graphs]
    rng = np.random.default_rng(seed)
    qg, qa, qr = [], [], []
    for i in range(n_bfs + n_sfs):
        gi = i % len(graphs)
        qg.append(graphs[gi])
        qa.append(0 if i < n_bfs else 1)
        qr.append(int(rng.integers(0, Vs[gi])))
    return graphs, qg, qa, qr
