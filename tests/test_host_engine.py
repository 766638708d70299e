"""The engine without a GPU (use_device_num 0, MULTI_CORE, SINGLE_CORE): every
iteration on the host team, with three-stage task stealing and edge donation
(TaskSteal::threeStage_taskSteal, my_project/CGgraph/taskSteal.hpp:89-274).

These run on the CPU: the library is asked for no GPU, so nothing here needs
one.  Outputs are compared bit for bit with the oracle.
"""
import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd._lib import CG_ERR_NO_DEVICE, CG_ERR_UNSUPPORTED, CgError

COORDINATION, SINGLE_CORE, MULTI_CORE, SINGLE_GPU = 0, 2, 3, 4


def hub_graph(oracle, seed, V=20000, hub_edges=15000, rest=60000):
    """Random graph whose vertex 0 has more out-edges than the 12 800-edge
    donation threshold, so the EDGE_DONATE stage runs."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([np.zeros(hub_edges, np.uint32), rng.integers(0, V, rest).astype(np.uint32)])
    dst = np.concatenate([rng.permutation(V)[:hub_edges].astype(np.uint32),
                          rng.integers(0, V, rest).astype(np.uint32)])
    g = oracle.csr_build(V, src, dst)
    g.weight = rng.integers(1, 101, g.E).astype(np.uint32)
    assert g.offsets[1] - g.offsets[0] >= 12800
    return g


def host_engine(graph, alg, kind, threads=4, **opt):
    if kind == "use_device_num=0":
        return cg.Engine(graph, alg, engine_type=COORDINATION, use_device_num=0, cpu_threads=threads, **opt)
    if kind == "MULTI_CORE":
        return cg.Engine(graph, alg, engine_type=MULTI_CORE, use_device_num=0, cpu_threads=threads, **opt)
    return cg.Engine(graph, alg, engine_type=SINGLE_CORE, use_device_num=0, **opt)


KINDS = ["use_device_num=0", "MULTI_CORE", "SINGLE_CORE"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("name", ["toy_0", "toy_1"])
def test_toys(oracle, golden, name, kind):
    t = golden["graphs"][name]
    g = cg.Graph.from_csr(np.array(t["offsets"], np.uint32), np.array(t["dest"], np.uint32),
                          np.array(t["weight"], np.uint32))
    for r, exp in t["bfs"].items():
        e = host_engine(g, 0, kind)
        got, st = e.run(int(r))
        assert got.tolist() == exp
        assert st["light_iterations"] == st["iterations"] and st["heavy_iterations"] == 0
        e.close()
    for r, exp in t["sssp"].items():
        e = host_engine(g, 1, kind)
        assert e.run(int(r))[0].tolist() == exp
        e.close()
    e = host_engine(g, 2, kind)
    assert e.run(0)[0].tolist() == t["cc_directed"]
    e.close()
    g.close()


@pytest.mark.parametrize("threads", [1, 3, 8])
@pytest.mark.parametrize("seed", range(2))
def test_random_with_donation(oracle, seed, threads):
    og = hub_graph(oracle, seed)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    for alg, ref in ((0, oracle.bfs(og, 0)), (1, oracle.sssp(og, 0)), (2, oracle.cc_directed(og))):
        e = host_engine(g, alg, "MULTI_CORE", threads=threads)
        got, st = e.run(0)
        assert np.array_equal(got, ref), (alg, threads)
        if alg != 2:
            assert st["edges_traversed"] == oracle.reached_edges(og, ref)
        assert np.array_equal(e.result(), ref)
        e.close()
    g.close()


def test_symmetric_cc_and_repeat_runs(oracle):
    og = oracle.rmat_graph(12, 8, seed=5, symmetrize=True)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    e = host_engine(g, 2, "use_device_num=0")
    ref = oracle.cc_undirected(og)
    for _ in range(3):
        assert np.array_equal(e.run(0)[0], ref)
    e.close()
    og = oracle.rmat_graph(13, 8, seed=6)
    g2 = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    e = host_engine(g2, 1, "use_device_num=0", threads=6)
    roots = [oracle.max_outdeg_vertex(og), 0, 17]
    for r in roots:  # several roots on one engine: every run starts clean
        assert np.array_equal(e.run(r)[0], oracle.sssp(og, r))
    e.close()
    g.close()
    g2.close()


def test_zero_outdeg_tail(oracle):
    """RCM_LAST_ZERO_OUTDEGREE input (V' < V): the CC frontier starts with V'
    vertices, BFS/SSSP from a root in the tail reach only the root."""
    rng = np.random.default_rng(3)
    V, tail = 3000, 700
    src = rng.integers(0, V - tail, 20000).astype(np.uint32)
    dst = rng.integers(0, V, 20000).astype(np.uint32)
    og = oracle.csr_build(V, src, dst)
    og.weight = rng.integers(1, 101, og.E).astype(np.uint32)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight, zero_outdeg_tail=tail)
    for alg, ref in ((0, oracle.bfs(og, 5)), (1, oracle.sssp(og, 5))):
        e = host_engine(g, alg, "use_device_num=0")
        assert np.array_equal(e.run(5)[0], ref)
        e.close()
    e = host_engine(g, 2, "MULTI_CORE")
    assert np.array_equal(e.run(0)[0], oracle.cc_directed(og))
    e.close()
    e = host_engine(g, 1, "use_device_num=0")
    got = e.run(V - 1)[0]
    assert got[V - 1] == 0 and np.count_nonzero(got != cg.CG_UNREACHED) == 1
    e.close()
    g.close()


def test_empty_graph_and_errors(oracle):
    g = cg.Graph.from_csr(np.zeros(9, np.uint64), np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    e = host_engine(g, 1, "use_device_num=0")
    got, st = e.run(3)
    assert got[3] == 0 and np.count_nonzero(got != cg.CG_UNREACHED) == 1 and st["iterations"] == 1
    e.close()
    with pytest.raises(CgError) as ei:
        cg.Engine(g, 3, engine_type=MULTI_CORE, use_device_num=0)
    assert ei.value.code == CG_ERR_UNSUPPORTED
    with pytest.raises(CgError) as ei:
        cg.Engine(g, 0, engine_type=1, use_device_num=0)  # COOPERATION is not provided
    assert ei.value.code == CG_ERR_UNSUPPORTED
    g.close()


def test_gpu_engine_without_gpu_fails_loudly(oracle):
    """Asking for a GPU engine where there is none is an error, never a
    silent run on the host team."""
    import ctypes
    n = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
            pytest.skip("a GPU is present")
    except OSError:
        pass
    og = oracle.rmat_graph(8, 4, seed=1)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    for et in (COORDINATION, SINGLE_GPU):
        with pytest.raises(CgError) as ei:
            cg.Engine(g, 0, engine_type=et, use_device_num=1)
        assert ei.value.code == CG_ERR_NO_DEVICE
    g.close()
