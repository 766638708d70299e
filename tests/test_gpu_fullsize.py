"""The benchmark configurations at full size, checked through properties that
do not need a full-size textbook run:

* R-MAT scale 18 / ef16 BFS (the reference's CPU-runnable case): bit-exact
  against the oracle's queue BFS (it is small enough).
* R-MAT scale 24 / ef16 SSSP: the shortest-path certificate (feasible edges +
  a tight in-edge for every reached vertex), and bit-exact agreement with the
  restated reference CPU engine.
* the same graph symmetrized, WCC: bit-exact against union-find.
* R-MAT scale 25 / ef44 PageRank: one float64 step from the engine's 19th
  iterate reproduces its 20th (L1 <= 1e-6), and the total rank mass equals
  (1 - d) + d * (mass on vertices with out-edges) of the previous iterate.
* R-MAT scales 20 and 22 / ef16 PageRank on either side of the automatic
  switch to propagation blocking (2^22 vertices): L1 <= 1e-6 against the
  FP64 oracle for the automatic layout and for the renumbered pull.
The GPU generator's CSR is compared with the oracle's at scale 18 bit for bit.
"""
import numpy as np
import pytest

import cggraph_amd as cg

pytestmark = pytest.mark.gpu


def host_csr(oracle, g):
    off, dst, w = g.export()
    return oracle.CSR(off, dst, w)


def test_bfs_rmat18_bit_exact(oracle):
    og = oracle.rmat_graph(18, 16, seed=1, weights=None)
    g = cg.Graph.rmat(18, 16, seed=1, weights=None)
    off, dst, _ = g.export()
    assert np.array_equal(off, og.offsets) and np.array_equal(dst, og.dest)
    root = oracle.max_outdeg_vertex(og)
    for mode in (1, 0):
        e = cg.Engine(g, 0, sched_mode=mode, cpu_threads=8)
        got, st = e.run(root)
        e.close()
        assert np.array_equal(got, oracle.bfs(og, root)), mode
        assert st["edges_traversed"] == oracle.reached_edges(og, got)


def test_sssp_rmat24_certificate_and_cpu_engine(oracle):
    g = cg.Graph.rmat(24, 16, seed=1, weights=(1, 100))
    root = g.info()["max_outdeg_vertex"]
    e = cg.Engine(g, 1)
    got, st = e.run(root)
    e.close()
    og = host_csr(oracle, g)
    assert oracle.check_sssp_cert(og, root, got) == 0
    ref, _, _ = oracle.multicore_run(1, og, root, threads=16)
    assert np.array_equal(got, ref)
    assert st["edges_traversed"] == oracle.reached_edges(og, got)


def test_wcc_rmat24_symmetric(oracle):
    g = cg.Graph.rmat(24, 16, seed=1, weights=None, symmetrize=True)
    assert g.info()["symmetric"]
    e = cg.Engine(g, 2)
    got, _ = e.run()
    e.close()
    og = host_csr(oracle, g)
    assert np.array_equal(got, oracle.cc_undirected(og))


def test_pagerank_rmat25_ef44_step_consistency(oracle):
    g = cg.Graph.rmat(25, 44, seed=1, weights=None)
    e19 = cg.Engine(g, 3, pr_iterations=19)
    r19, _ = e19.run()
    e19.close()
    e20 = cg.Engine(g, 3, pr_iterations=20)
    r20, _ = e20.run()
    e20.close()
    og = host_csr(oracle, g)
    step = oracle.pagerank_step(og, r19, 0.85)
    l1 = float(np.abs(r20.astype(np.float64) - step).sum())
    assert l1 <= 1e-6, l1
    deg = np.diff(og.offsets)
    mass = (1 - 0.85) + 0.85 * float(r19[deg > 0].astype(np.float64).sum())
    assert abs(float(r20.astype(np.float64).sum()) - mass) <= 1e-5


@pytest.mark.parametrize("scale", [20, 22])
def test_pagerank_automatic_blocking_mid_size(oracle, scale):
    """The automatic layout (the renumbered pull below 2^22 vertices,
    propagation blocking from there) and the renumbered pull: both within
    FP32 rounding of the FP64 oracle."""
    g = cg.Graph.rmat(scale, 16, seed=3, weights=None)
    e = cg.Engine(g, 3)
    auto, _ = e.run()
    e.close()
    e = cg.Engine(g, 3, pr_layout=2)
    pull, _ = e.run()
    e.close()
    ref = oracle.pagerank64(host_csr(oracle, g), 20, 0.85)
    assert float(np.abs(auto.astype(np.float64) - ref).sum()) <= 1e-6
    assert float(np.abs(pull.astype(np.float64) - ref).sum()) <= 1e-6
