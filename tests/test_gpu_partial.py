"""Partial residency: TwoLevelPartition's placement under an HBM budget
(my_project/CGgraph/Two-level_partition.cuh:213-476, 392-445).  The 1024-vertex
segments, sorted by average out-degree, keep their edges in HBM while the
budget lasts; the host team relaxes the frontier vertices of the others every
iteration.  Outputs must equal the oracle's whatever the split, and the
resident edges must be exactly the degree-ordered prefix of segments that fits.
"""
import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib

pytestmark = pytest.mark.gpu

SEG = 1024


def placement(og, alg, budget):
    """The partition's rule restated: fixed arrays first, then segments in
    ascending average degree (ties by id) at 8 bytes per edge for SSSP, 4
    otherwise, while the budget lasts.  Returns the resident edge count."""
    V = og.V
    Vp = V - getattr(og, "zero_tail", 0)
    off = np.asarray(og.offsets, np.int64)
    nseg = (Vp + SEG - 1) // SEG
    a = np.arange(nseg) * SEG
    b = np.minimum(a + SEG, Vp)
    edges = off[b] - off[a]
    order = np.lexsort((np.arange(nseg), edges / (b - a)))
    used = 4 * V + 8 * (Vp + 1) + (1 << 30)
    per = 8 if alg == 1 else 4
    resident = 0
    for s in order:
        if used + per * edges[s] > budget:
            break
        used += per * edges[s]
        resident += int(edges[s])
    return resident


def budget_for(og, alg, frac, tail=0):
    V = og.V
    Vp = V - tail
    return 4 * V + 8 * (Vp + 1) + (1 << 30) + int(frac * og.E * (8 if alg == 1 else 4))


def reference(oracle, og, alg, root):
    return oracle.bfs(og, root) if alg == 0 else (oracle.sssp(og, root) if alg == 1 else oracle.cc_directed(og))


@pytest.mark.parametrize("frac", [0.0, 0.35, 0.8])
@pytest.mark.parametrize("alg", [0, 1, 2])
def test_partial_residency_matches_oracle(oracle, alg, frac):
    og = oracle.rmat_graph(14, 16, seed=7)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    budget = budget_for(og, alg, frac)
    want = placement(og, alg, budget)
    assert want < og.E
    for mode in (1, 0):  # GPU schedule, the reference's heuristic
        e = cg.Engine(g, alg, sched_mode=mode, hbm_budget=budget, cpu_threads=8)
        ref = reference(oracle, og, alg, root)
        for _ in range(2):
            got, st = e.run(root)
            assert np.array_equal(got, ref), (alg, frac, mode)
            assert st["resident_edges"] == want
            assert st["cpu_vertices"] > 0
            if alg != 2:
                assert st["edges_traversed"] == oracle.reached_edges(og, got)
        e.close()
    g.close()


def test_partial_residency_other_roots_and_tail(oracle):
    """Roots inside and outside the resident part, and a zero-out-degree
    tail (V' < V: CC starts from V')."""
    rng = np.random.default_rng(3)
    V, tail, m = 30000, 5000, 240000
    src = np.concatenate([rng.integers(0, 40, m // 4), rng.integers(0, V - tail, m)]).astype(np.uint32)
    dst = rng.integers(0, V, src.size).astype(np.uint32)
    og = oracle.csr_build(V, src, dst)
    og.weight = rng.integers(1, 101, og.E).astype(np.uint32)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight, zero_outdeg_tail=tail)
    for alg in (0, 1, 2):
        e = cg.Engine(g, alg, sched_mode=1, hbm_budget=budget_for(og, alg, 0.5, tail), cpu_threads=8)
        for root in (0, 12345, V - 1):
            got, st = e.run(root)
            assert np.array_equal(got, reference(oracle, og, alg, root)), (alg, root)
            assert 0 < st["resident_edges"] < og.E
            if alg == 2:
                break
        e.close()
    g.close()


def test_full_residency_and_refusals(oracle):
    og = oracle.rmat_graph(12, 8, seed=2)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    e = cg.Engine(g, 1, sched_mode=1, hbm_budget=budget_for(og, 1, 1.0))
    got, st = e.run(root)
    e.close()
    assert np.array_equal(got, oracle.sssp(og, root))
    assert st["resident_edges"] == og.E and st["cpu_vertices"] == 0
    small = budget_for(og, 1, 0.2)
    with pytest.raises(cg.CgError) as ei:
        cg.Engine(g, 3, hbm_budget=small)
    assert ei.value.code == _lib.CG_ERR_UNSUPPORTED
    with pytest.raises(cg.CgError) as ei:
        cg.Engine(g, 1, sched_mode=_lib.CG_SCHED_PROBE, trace_path="/dev/null", hbm_budget=small)
    assert ei.value.code == _lib.CG_ERR_UNSUPPORTED
    g.close()
