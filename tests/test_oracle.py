"""Pins the CPU oracle before anything is compared against it.

Checks: the reference's own toy graphs (golden vectors computed by networkx /
numpy, tests/golden/make_golden.py), scipy on random graphs, closed forms,
brute force, an independent numpy generator, and planted mistakes that the
certificate check must reject.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

from tests import genref

U = 999_999_999


def toy(oracle, golden, name):
    t = golden["graphs"][name]
    return oracle.CSR(t["offsets"], t["dest"], t["weight"]), t


@pytest.mark.parametrize("name", ["toy_0", "toy_1"])
def test_toy_golden(oracle, golden, name):
    g, t = toy(oracle, golden, name)
    for r, exp in t["bfs"].items():
        assert oracle.bfs(g, int(r)).tolist() == exp
    for r, exp in t["sssp"].items():
        assert oracle.sssp(g, int(r)).tolist() == exp
    assert oracle.cc_directed(g).tolist() == t["cc_directed"]
    sym = oracle.csr_build(g.V, np.repeat(np.arange(g.V), np.diff(g.offsets).astype(np.int64)),
                           g.dest, symmetrize=True)
    assert oracle.cc_undirected(sym).tolist() == t["cc_symmetric"]
    assert np.allclose(oracle.pagerank64(g, 20, 0.85), t["pr"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", ["toy_0", "toy_1"])
def test_toy_multicore(oracle, golden, name):
    g, t = toy(oracle, golden, name)
    for r in t["bfs"]:
        v, _, _ = oracle.multicore_run(0, g, int(r), threads=4)
        assert v.tolist() == t["bfs"][r]
        v, _, _ = oracle.multicore_run(1, g, int(r), threads=4)
        assert v.tolist() == t["sssp"][r]
    v, _, _ = oracle.multicore_run(2, g, 0, threads=4)
    assert v.tolist() == t["cc_directed"]


def random_graph(oracle, rng, V, m, wmax=100):
    src = rng.integers(0, V, m).astype(np.uint32)
    dst = rng.integers(0, V, m).astype(np.uint32)
    g = oracle.csr_build(V, src, dst)
    g.weight = rng.integers(1, wmax + 1, g.E).astype(np.uint32)
    return g


def to_scipy(g, unit=False):
    rows = np.repeat(np.arange(g.V), np.diff(g.offsets).astype(np.int64))
    data = np.ones(g.E) if unit else g.weight.astype(np.float64)
    return sp.csr_matrix((data, (rows, g.dest.astype(np.int64))), shape=(g.V, g.V))


@pytest.mark.parametrize("seed", range(6))
def test_against_scipy(oracle, seed):
    rng = np.random.default_rng(seed)
    V = int(rng.integers(2, 400))
    g = random_graph(oracle, rng, V, int(rng.integers(0, 4 * V)))
    root = int(rng.integers(0, V))
    for unit, fn in ((True, oracle.bfs), (False, oracle.sssp)):
        d = csgraph.dijkstra(to_scipy(g, unit), directed=True, indices=root)
        exp = np.where(np.isinf(d), U, d).astype(np.uint32)
        got = fn(g, root)
        assert np.array_equal(got, exp)
        assert oracle.check_sssp_cert(g, root, got, unit=unit) == 0
    n, lab = csgraph.connected_components(to_scipy(g, True), directed=True, connection="weak")
    mins = np.full(n, V)
    np.minimum.at(mins, lab, np.arange(V))
    sym = oracle.csr_build(V, np.repeat(np.arange(V), np.diff(g.offsets).astype(np.int64)),
                           g.dest, symmetrize=True)
    assert np.array_equal(oracle.cc_undirected(sym), mins[lab])


@pytest.mark.parametrize("seed", range(4))
def test_cc_directed_bruteforce(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    V = int(rng.integers(1, 40))
    g = random_graph(oracle, rng, V, int(rng.integers(0, 3 * V)))
    A = to_scipy(g, True).toarray() > 0
    R = np.eye(V, dtype=bool) | A
    for _ in range(V):  # transitive closure by repeated squaring
        R = R | ((R.astype(np.int64) @ R.astype(np.int64)) > 0)
    exp = [min(u for u in range(V) if R[u, v]) for v in range(V)]
    assert oracle.cc_directed(g).tolist() == exp


def test_pagerank_closed_forms(oracle):
    n = 7
    cyc = oracle.csr_build(n, np.arange(n), (np.arange(n) + 1) % n)
    pr = oracle.pagerank64(cyc, 20, 0.85)
    assert np.allclose(pr, 1.0 / n, atol=1e-16)
    # star: hub 0 -> leaves 1..k; leaves dangling.  PR(0) = 0.15/N; PR(leaf) =
    # 0.15/N + 0.85 * PR_prev(0)/k, which is fixed after iteration 2.
    k = 5
    N = k + 1
    star = oracle.csr_build(N, np.zeros(k, np.uint32), np.arange(1, N))
    pr = oracle.pagerank64(star, 20, 0.85)
    assert abs(pr[0] - 0.15 / N) < 1e-17
    assert np.allclose(pr[1:], 0.15 / N + 0.85 * (0.15 / N) / k, atol=1e-17)


def test_pagerank_dense(oracle):
    rng = np.random.default_rng(7)
    V = 60
    g = random_graph(oracle, rng, V, 300)
    M = np.zeros((V, V))
    deg = np.diff(g.offsets).astype(np.int64)
    for u in range(V):
        for e in range(int(g.offsets[u]), int(g.offsets[u + 1])):
            M[g.dest[e], u] += 1.0 / deg[u]
    pr = np.full(V, 1.0 / V)
    for _ in range(20):
        pr = 0.15 / V + 0.85 * M @ pr
    assert np.allclose(oracle.pagerank64(g, 20, 0.85), pr, rtol=1e-13, atol=0)


def test_certificate_rejects_planted_mistakes(oracle):
    rng = np.random.default_rng(3)
    g = random_graph(oracle, rng, 300, 1500)
    root = oracle.max_outdeg_vertex(g)
    d = oracle.sssp(g, root)
    assert oracle.check_sssp_cert(g, root, d) == 0
    reached = np.nonzero((d != U) & (np.arange(g.V) != root))[0]
    for delta in (+1, -1):
        bad = d.copy()
        bad[reached[len(reached) // 2]] += np.uint32(delta) if delta > 0 else 0
        if delta < 0:
            bad[reached[len(reached) // 2]] -= np.uint32(1)
        assert oracle.check_sssp_cert(g, root, bad) != 0
    bad = d.copy()
    bad[reached[0]] = U
    assert oracle.check_sssp_cert(g, root, bad) != 0


@pytest.mark.parametrize("scale,ef", [(6, 4), (10, 16), (12, 8)])
def test_rmat_generator_matches_numpy(oracle, scale, ef):
    s, d = oracle.rmat_edges(scale, ef)
    s2, d2 = genref.rmat_edges(scale, ef)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    g = oracle.csr_build(1 << scale, s, d)
    off, dst = genref.csr_from_edges(1 << scale, s2, d2)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.dest, dst)
    oracle.fill_weights(g, 1, 1, 100)
    rows = np.repeat(np.arange(g.V), np.diff(g.offsets).astype(np.int64))
    assert np.array_equal(g.weight, genref.edge_weight(1, rows, g.dest, 1, 100))
    assert g.weight.min() >= 1 and g.weight.max() <= 100
    # R-MAT quadrant probabilities: the source's top bit is 0 with probability a+b.
    top = (s >> (scale - 1)) & 1
    assert abs(1 - top.mean() - 0.76) < 0.05


def test_rmat_symmetrize(oracle):
    s, d = oracle.rmat_edges(10, 8)
    g = oracle.csr_build(1 << 10, s, d, symmetrize=True)
    off, dst = genref.csr_from_edges(1 << 10, s, d, symmetrize=True)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.dest, dst)
    t = oracle.transpose(g)
    assert np.array_equal(t.offsets, g.offsets) and np.array_equal(t.dest, g.dest)


@pytest.mark.parametrize("n", [2, 5, 32])
def test_grid_generator(oracle, n):
    g = oracle.grid_graph(n, 0.1, seed=3)
    s2, d2 = genref.grid_edges(n, 0.1, seed=3)
    off, dst = genref.csr_from_edges(n * n, s2, d2)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.dest, dst)
    rows = np.repeat(np.arange(g.V), np.diff(g.offsets).astype(np.int64))
    assert np.array_equal(g.weight, genref.edge_weight(3, rows, g.dest, 1, 1000, symmetric=True))
    t = oracle.transpose(g)
    assert np.array_equal(t.dest, g.dest) and np.array_equal(t.weight, g.weight)


def test_transpose_sorted(oracle):
    rng = np.random.default_rng(5)
    g = random_graph(oracle, rng, 200, 900)
    t = oracle.transpose(g)
    A = to_scipy(g).toarray()
    B = to_scipy(t).toarray()
    assert np.array_equal(A.T, B)
    for v in range(t.V):
        seg = t.dest[int(t.offsets[v]):int(t.offsets[v + 1])]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)


@pytest.mark.parametrize("seed", range(3))
def test_multicore_random(oracle, seed):
    """Restated MULTI_CORE engine == textbook results (includes vertices above
    the 12 800-edge donation threshold to drive the EDGE_DONATE stage)."""
    rng = np.random.default_rng(seed)
    V = 20000
    src = np.concatenate([np.zeros(15000, np.uint32), rng.integers(0, V, 60000).astype(np.uint32)])
    dst = np.concatenate([rng.permutation(V)[:15000].astype(np.uint32),
                          rng.integers(0, V, 60000).astype(np.uint32)])
    g = oracle.csr_build(V, src, dst)
    g.weight = rng.integers(1, 101, g.E).astype(np.uint32)
    assert g.offsets[1] - g.offsets[0] >= 12800
    for threads in (1, 3, 8):
        v, _, _ = oracle.multicore_run(0, g, 0, threads)
        assert np.array_equal(v, oracle.bfs(g, 0))
        v, _, _ = oracle.multicore_run(1, g, 0, threads)
        assert np.array_equal(v, oracle.sssp(g, 0))
        v, _, _ = oracle.multicore_run(2, g, 0, threads)
        assert np.array_equal(v, oracle.cc_directed(g))


def test_pagerank_step_consistent(oracle):
    """One float64 step from float32 ranks reproduces the next iteration of the
    float64 power iteration to within float32 input rounding."""
    rng = np.random.default_rng(8)
    g = random_graph(oracle, rng, 500, 4000)
    p19 = oracle.pagerank64(g, 19, 0.85)
    p20 = oracle.pagerank64(g, 20, 0.85)
    step = oracle.pagerank_step(g, p19.astype(np.float32), 0.85)
    assert np.abs(step - p20).sum() < 1e-6
    assert np.abs(step - p20).sum() > 0  # the float32 rounding is visible, so inputs were used


@pytest.mark.parametrize("seed", range(3))
def test_reached_edges_and_root_helpers(oracle, seed):
    """The TEPS numerator (sum of out-degrees of reached vertices) and the
    root rule (largest out-degree, lowest id on ties) against numpy
    one-liners, on graphs with unreached vertices and tied maximum degrees."""
    rng = np.random.default_rng(seed)
    V = 3000
    src = rng.integers(0, V // 2, 9000).astype(np.uint32)  # the upper half sends nothing
    dst = rng.integers(0, V, 9000).astype(np.uint32)
    g = oracle.csr_build(V, src, dst)
    g.weight = rng.integers(1, 101, g.E).astype(np.uint32)
    deg = np.diff(g.offsets.astype(np.int64))
    for root in (0, int(rng.integers(0, V)), V - 1):
        val = oracle.sssp(g, root)
        assert oracle.reached_edges(g, val) == int(deg[val != U].sum())
    val = np.full(V, U, np.uint32)
    assert oracle.reached_edges(g, val) == 0
    assert oracle.max_outdeg_vertex(g) == int(np.argmax(deg))  # argmax takes the first maximum
    # ties: two vertices share the maximum; the lower id wins
    off = np.array([0, 1, 3, 3, 5], np.uint64)
    t = oracle.CSR(off, np.array([1, 0, 2, 0, 1], np.uint32))
    assert oracle.max_outdeg_vertex(t) == 1
