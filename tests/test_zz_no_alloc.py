"""cg_traverse_device_noalloc (the allocation-free run) against cg_engine_run.

Two calls in a row on one process's scratch array, the second on a smaller
graph, for BFS, SSSP and CC; each output must equal cg_engine_run's (and the
oracle's) bit for bit.  The device buffers the caller owns (CSR, values) come
from the graph object and from hipMalloc in this test, never from the library
call under test.
"""
import ctypes

import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib

pytestmark = pytest.mark.gpu


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def device_csr(g):
    off, dst, w = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(_lib.lib().cg_graph_device_csr(g._h, ctypes.byref(off), ctypes.byref(dst), ctypes.byref(w)))
    return off, dst, w


def run_noalloc(hip, g, alg, root):
    inf = g.info()
    V, E = inf["V"], inf["E"]
    off, dst, w = device_csr(g)
    d_val = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_val), max(V, 1) * 4) == 0
    try:
        it = ctypes.c_uint32(0)
        print(f"noalloc alg={alg} V={V} E={E} root={root}", flush=True)
        _lib.check(_lib.lib().cg_traverse_device_noalloc(alg, off, dst, w, V, E, root, d_val, None, ctypes.byref(it)))
        out = np.empty(V, np.uint32)
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), d_val, V * 4, 2) == 0
        return out, it.value
    finally:
        hip.hipFree(d_val)


@pytest.mark.parametrize("alg", [0, 1, 2])
def test_noalloc_matches_engine_two_calls(oracle, alg):
    hip = _hip()
    big = oracle.rmat_graph(17, 8, seed=11, symmetrize=(alg == 2))
    small = oracle.rmat_graph(11, 6, seed=12, symmetrize=(alg == 2))
    for og in (big, small):  # the second call is on the smaller graph
        g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
        root = oracle.max_outdeg_vertex(og)
        e = cg.Engine(g, alg, sched_mode=1)
        ref, st = e.run(root)
        e.close()
        got, iters = run_noalloc(hip, g, alg, root)
        assert np.array_equal(got, ref), (alg, og.V)
        if alg == 0:
            assert np.array_equal(got, oracle.bfs(og, root))
        elif alg == 1:
            assert np.array_equal(got, oracle.sssp(og, root))
        else:
            assert np.array_equal(got, oracle.cc_undirected(og))
        assert iters >= 1
        g.close()


def test_noalloc_repeat_and_directed_cc(oracle):
    """Repeated calls on one graph (the scratch keeps the last call's
    contents, so each call must clear what it relies on) and label
    propagation on directed input."""
    hip = _hip()
    og = oracle.rmat_graph(14, 8, seed=5)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    for root in (oracle.max_outdeg_vertex(og), 3, 3):
        got, _ = run_noalloc(hip, g, 1, root)
        assert np.array_equal(got, oracle.sssp(og, root))
    got, _ = run_noalloc(hip, g, 2, 0)
    assert np.array_equal(got, oracle.cc_directed(og))
    g.close()
