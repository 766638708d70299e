"""GPU parity of the traversal engine against the CPU oracle (bit-exact).

Every run goes through the C ABI (libcgmi.so).  Sizes stay where the oracle
finishes in seconds; the full-size R-MAT cases live in test_gpu_fullsize.py.
"""
import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib

pytestmark = pytest.mark.gpu
U = 999_999_999
BFS, SSSP, CC, PR = 0, 1, 2, 3


def gpu_graph(o):
    return cg.Graph.from_csr(o.offsets, o.dest, o.weight)


def run(g, alg, root=0, **opt):
    e = cg.Engine(g, alg, **opt)
    try:
        out, st = e.run(root)
    finally:
        e.close()
    return out, st


def sym_of(oracle, g):
    rows = np.repeat(np.arange(g.V), np.diff(g.offsets).astype(np.int64))
    return oracle.csr_build(g.V, rows, g.dest, symmetrize=True)


@pytest.mark.parametrize("name", ["toy_0", "toy_1"])
def test_toys_all_roots(oracle, golden, name):
    t = golden["graphs"][name]
    og = oracle.CSR(t["offsets"], t["dest"], t["weight"])
    g = gpu_graph(og)
    for r in range(t["V"]):
        for mode in (_lib.CG_SCHED_GPU, _lib.CG_SCHED_HEURISTIC):
            b, _ = run(g, BFS, r, sched_mode=mode)
            s, _ = run(g, SSSP, r, sched_mode=mode)
            assert b.tolist() == oracle.bfs(og, r).tolist(), (name, r, mode)
            assert s.tolist() == oracle.sssp(og, r).tolist(), (name, r, mode)
            if str(r) in t["bfs"]:
                assert b.tolist() == t["bfs"][str(r)]
                assert s.tolist() == t["sssp"][str(r)]
    c, _ = run(g, CC)
    assert c.tolist() == t["cc_directed"]
    gs = gpu_graph(sym_of(oracle, og))
    assert gs.info()["symmetric"]
    c, _ = run(gs, CC)
    assert c.tolist() == t["cc_symmetric"]


def random_graph(oracle, seed, V, m, hub=0):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, V, m).astype(np.uint32)
    dst = rng.integers(0, V, m).astype(np.uint32)
    if hub:
        src = np.concatenate([src, np.zeros(hub, np.uint32)])
        dst = np.concatenate([dst, rng.integers(0, V, hub).astype(np.uint32)])
    g = oracle.csr_build(V, src, dst)
    g.weight = rng.integers(1, 101, g.E).astype(np.uint32)
    return g


@pytest.mark.parametrize("seed,V,m,hub", [(0, 1000, 5000, 0), (1, 4099, 30000, 9000),
                                          (2, 70000, 400000, 60000), (3, 5, 0, 0)])
@pytest.mark.parametrize("mode", [_lib.CG_SCHED_GPU, _lib.CG_SCHED_HEURISTIC])
def test_random_graphs(oracle, seed, V, m, hub, mode):
    og = random_graph(oracle, seed, V, m, hub)
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    b, st = run(g, BFS, root, sched_mode=mode)
    assert np.array_equal(b, oracle.bfs(og, root))
    assert st["edges_traversed"] == oracle.reached_edges(og, b)
    s, _ = run(g, SSSP, root, sched_mode=mode)
    assert np.array_equal(s, oracle.sssp(og, root))
    c, _ = run(g, CC, sched_mode=mode)
    assert np.array_equal(c, oracle.cc_directed(og))


@pytest.mark.parametrize("choice", range(16))
def test_forced_choices(oracle, choice):
    """Every one of the 16 schedule choices reaches the same fixed point."""
    og = random_graph(oracle, 11, 20000, 150000, hub=20000)
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    for alg, ref in ((BFS, oracle.bfs(og, root)), (SSSP, oracle.sssp(og, root)),
                     (CC, oracle.cc_directed(og))):
        got, _ = run(g, alg, root, sched_mode=_lib.CG_SCHED_FORCED, forced_choice=choice)
        assert np.array_equal(got, ref), (alg, choice)


@pytest.mark.parametrize("scale,ef", [(10, 16), (14, 16), (16, 8)])
def test_rmat_small(oracle, scale, ef):
    og = oracle.rmat_graph(scale, ef, seed=1)
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    assert g.info()["max_outdeg_vertex"] == root
    b, _ = run(g, BFS, root)
    assert np.array_equal(b, oracle.bfs(og, root))
    s, _ = run(g, SSSP, root)
    assert np.array_equal(s, oracle.sssp(og, root))
    sym = sym_of(oracle, og)
    c, _ = run(gpu_graph(sym), CC)
    assert np.array_equal(c, oracle.cc_undirected(sym))


@pytest.mark.parametrize("delta", [1, 37, 1000, 10**8])
def test_sssp_delta_stepping(oracle, delta):
    """Deferring frontier vertices above a moving threshold changes the order
    of relaxations only; the distances stay the unique fixed point."""
    cases = [random_graph(oracle, 21, 30000, 200000, hub=30000), oracle.grid_graph(120, 0.1, seed=8),
             oracle.rmat_graph(15, 16, seed=4)]
    for og in cases:
        g = gpu_graph(og)
        for root in (oracle.max_outdeg_vertex(og), og.V // 3):
            for mode in (_lib.CG_SCHED_GPU, _lib.CG_SCHED_HEURISTIC):
                got, st = run(g, SSSP, root, sched_mode=mode, sssp_delta=delta)
                assert np.array_equal(got, oracle.sssp(og, root)), (delta, root, mode)


def test_fire_and_forget_atomics(oracle):
    """Activation on the pre-check (atomics without return) over-activates at
    most; the values are still the fixed point."""
    cases = [random_graph(oracle, 31, 40000, 300000, hub=40000), oracle.rmat_graph(16, 16, seed=7),
             oracle.grid_graph(90, 0.1, seed=1)]
    for og in cases:
        g = gpu_graph(og)
        root = oracle.max_outdeg_vertex(og)
        for tier in (0, 6):
            opts = dict(atomic_return=0, sched_mode=_lib.CG_SCHED_FORCED, forced_choice=tier)
            assert np.array_equal(run(g, BFS, root, **opts)[0], oracle.bfs(og, root))
            assert np.array_equal(run(g, SSSP, root, **opts)[0], oracle.sssp(og, root))
            assert np.array_equal(run(g, CC, **opts)[0], oracle.cc_directed(og))
        assert np.array_equal(run(g, SSSP, root, atomic_return=0, sssp_delta=50)[0], oracle.sssp(og, root))


def test_gpu_generator_matches_oracle(oracle):
    for scale, ef, sym in ((12, 16, False), (13, 8, True)):
        og = oracle.rmat_graph(scale, ef, seed=5, symmetrize=sym)
        g = cg.Graph.rmat(scale, ef, seed=5, symmetrize=sym)
        off, dst, w = g.export()
        assert np.array_equal(off, og.offsets) and np.array_equal(dst, og.dest)
        assert np.array_equal(w, og.weight)
        assert g.info()["symmetric"] == sym
    og = oracle.grid_graph(37, 0.1, seed=2)
    g = cg.Graph.grid(37, 0.1, seed=2)
    off, dst, w = g.export()
    assert np.array_equal(off, og.offsets) and np.array_equal(dst, og.dest)
    assert np.array_equal(w, og.weight)


def test_grid_sssp(oracle):
    og = oracle.grid_graph(200, 0.1, seed=4)
    g = cg.Graph.grid(200, 0.1, seed=4)
    for root in (0, 12345, 39999):
        s, _ = run(g, SSSP, root)
        assert np.array_equal(s, oracle.sssp(og, root))
        b, _ = run(g, BFS, root)
        assert np.array_equal(b, oracle.bfs(og, root))
    c, _ = run(g, CC)
    assert np.array_equal(c, oracle.cc_undirected(og))


def test_edge_cases(oracle):
    # no edges at all
    og = oracle.csr_build(100, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    og.weight = np.zeros(0, np.uint32)
    g = gpu_graph(og)
    b, st = run(g, BFS, 7)
    assert b[7] == 0 and np.all(np.delete(b, 7) == U) and st["iterations"] == 1
    c, _ = run(g, CC)
    assert np.array_equal(c, np.arange(100))
    # root with zero out-degree, V not a multiple of 64
    og = random_graph(oracle, 9, 131, 300)
    zero = [v for v in range(og.V) if og.offsets[v + 1] == og.offsets[v]]
    g = gpu_graph(og)
    for r in zero[:3] + [130]:
        s, _ = run(g, SSSP, r)
        assert np.array_equal(s, oracle.sssp(og, r))
    # star with a hub far above every tier threshold
    n = 300001
    og = oracle.csr_build(n, np.zeros(n - 1, np.uint32), np.arange(1, n, dtype=np.uint32))
    og.weight = np.full(og.E, 3, np.uint32)
    g = gpu_graph(og)
    s, _ = run(g, SSSP, 0)
    assert s[0] == 0 and np.all(s[1:] == 3)
    with pytest.raises(cg.CgError):
        run(g, BFS, n)  # root out of range


def test_pagerank_l1_and_determinism(oracle):
    og = oracle.rmat_graph(16, 16, seed=1, weights=None)
    g = cg.Graph.from_csr(og.offsets, og.dest)
    e = cg.Engine(g, PR)
    r1, st = e.run()
    r2, _ = e.run()
    e.close()
    assert np.array_equal(r1.view(np.uint32), r2.view(np.uint32))
    ref = oracle.pagerank64(og, 20, 0.85)
    l1 = float(np.abs(r1.astype(np.float64) - ref).sum())
    assert l1 <= 1e-6, l1
    assert st["iterations"] == 20


@pytest.mark.parametrize("layout,seg,kernel", [(0, 31, 0), (0, 8, 0), (0, 11, 0), (0, 14, 0), (0, 0, 0), (1, 0, 1),
                                               (1, 0, 2), (2, 0, 1), (2, 0, 2), (2, 0, 3), (2, 12, 3), (2, 20, 3),
                                               (3, 9, 0), (4, 0, 0), (5, 0, 0)])
def test_pagerank_layouts(oracle, layout, seg, kernel):
    """Every in-edge layout (plain, renumbered by out-degree, source windows
    smaller and larger than the graph, propagation blocking per edge and per
    (source, bin) pair) and every sweep
    (row tiers, edge tiles, two source bands) against the FP64 oracle;
    run-to-run bit-identical."""
    og = oracle.rmat_graph(17, 16, seed=2, weights=None)
    g = cg.Graph.from_csr(og.offsets, og.dest)
    e = cg.Engine(g, PR, pr_segment_log2=seg, pr_layout=layout, pr_kernel=kernel)
    r1, _ = e.run()
    r2, _ = e.run()
    e.close()
    assert np.array_equal(r1.view(np.uint32), r2.view(np.uint32))
    ref = oracle.pagerank64(og, 20, 0.85)
    l1 = float(np.abs(r1.astype(np.float64) - ref).sum())
    assert l1 <= 1e-6, l1
    # a star into vertex 0 (one long run per window) plus dangling leaves
    n = 5000
    star = oracle.csr_build(n, np.arange(1, n, dtype=np.uint32), np.zeros(n - 1, np.uint32))
    g2 = cg.Graph.from_csr(star.offsets, star.dest)
    e = cg.Engine(g2, PR, pr_segment_log2=seg, pr_layout=layout, pr_kernel=kernel)
    r, _ = e.run()
    e.close()
    assert float(np.abs(r.astype(np.float64) - oracle.pagerank64(star, 20, 0.85)).sum()) <= 1e-6
    # a graph whose first and last vertices have no edges at all
    src = np.array([3, 3, 5, 7, 8, 8, 8], np.uint32)
    dst = np.array([5, 8, 3, 3, 3, 5, 7], np.uint32)
    sparse = oracle.csr_build(12, src, dst)
    g3 = cg.Graph.from_csr(sparse.offsets, sparse.dest)
    e = cg.Engine(g3, PR, pr_segment_log2=seg, pr_layout=layout, pr_kernel=kernel)
    r, _ = e.run()
    e.close()
    assert float(np.abs(r.astype(np.float64) - oracle.pagerank64(sparse, 20, 0.85)).sum()) <= 1e-6


def test_pagerank_blocked_builds(oracle):
    """Propagation blocking sums in 64-bit fixed point, so the ranks do not
    depend on the order of the additions: two engines (whose edge slots are
    placed by atomics, so their layouts may differ) give bit-identical ranks,
    on a graph whose size is not a multiple of the 8192-vertex blocks and
    whose hot blocks are cut into several tiles (the first block receives
    about 1.2 M in-edges: three gather tiles of at most 2^19 entries)."""
    rng = np.random.default_rng(5)
    m, hot = 600000, 1000000
    src = np.concatenate([rng.integers(0, 20001, m), np.zeros(300000, np.int64), rng.integers(0, 20001, hot)])
    dst = np.concatenate([rng.integers(0, 20001, m), rng.integers(0, 20001, 300000), rng.integers(0, 8192, hot)])
    og = oracle.csr_build(20001, src.astype(np.uint32), dst.astype(np.uint32))
    assert int(np.count_nonzero(og.dest < 8192)) > 2 * 2**19
    ref = oracle.pagerank64(og, 20, 0.85)
    g = cg.Graph.from_csr(og.offsets, og.dest)
    runs = []
    for _ in range(2):
        e = cg.Engine(g, PR, pr_layout=4)
        runs.append(e.run()[0])
        e.close()
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))
    assert float(np.abs(runs[0].astype(np.float64) - ref).sum()) <= 1e-6


def test_pagerank_pair_blocking(oracle):
    """Propagation blocking over (source, bin) pairs on the graph renumbered
    by out-degree: ranks in the caller's ids within 1e-6 (L1) of the FP64
    oracle, bit-identical across engines, on a graph whose hottest sources
    also receive the most edges, so that after renumbering their bin holds
    more than 2^20 edges (a bin cut into several gather tiles) and many pairs
    carry several edges; a graph whose size is not a multiple of 8192."""
    rng = np.random.default_rng(7)
    hot, m = 2600000, 600000
    V = 30001
    src = np.concatenate([rng.integers(0, 8192, hot), rng.integers(0, V, m)])
    dst = np.concatenate([rng.integers(0, 8192, hot), rng.integers(0, V, m)])
    og = oracle.csr_build(V, src.astype(np.uint32), dst.astype(np.uint32))
    assert int(np.count_nonzero(og.dest < 8192)) > 2 * 2**20
    ref = oracle.pagerank64(og, 20, 0.85)
    g = cg.Graph.from_csr(og.offsets, og.dest)
    runs = []
    for _ in range(2):
        e = cg.Engine(g, PR, pr_layout=5)
        runs.append(e.run()[0])
        runs.append(e.run()[0])
        e.close()
    for r in runs[1:]:
        assert np.array_equal(runs[0].view(np.uint32), r.view(np.uint32))
    assert float(np.abs(runs[0].astype(np.float64) - ref).sum()) <= 1e-6
    # the same ranks as the per-edge blocking, to FP32 rounding of the ranks
    e = cg.Engine(g, PR, pr_layout=4)
    r4 = e.run()[0]
    e.close()
    assert float(np.abs(r4.astype(np.float64) - runs[0].astype(np.float64)).sum()) <= 1e-7


def test_pagerank_bad_layout(oracle):
    og = oracle.rmat_graph(8, 4, seed=1, weights=None)
    g = cg.Graph.from_csr(og.offsets, og.dest)
    with pytest.raises(cg.CgError):
        cg.Engine(g, PR, pr_layout=6)
    with pytest.raises(cg.CgError):
        cg.Engine(g, PR, pr_kernel=4)


def test_error_paths(oracle):
    og = random_graph(oracle, 1, 50, 100)
    g = cg.Graph.from_csr(og.offsets, og.dest)  # no weights
    with pytest.raises(cg.CgError) as ei:
        cg.Engine(g, SSSP)
    assert ei.value.code == _lib.CG_ERR_INVALID
    with pytest.raises(cg.CgError) as ei:
        cg.Engine(g, BFS, engine_type=5)  # BLOCK_2D
    assert ei.value.code == _lib.CG_ERR_UNSUPPORTED
    gw = gpu_graph(og)
    e = cg.Engine(gw, BFS, sched_mode=_lib.CG_SCHED_GPU, fault_flags=_lib.CG_FAULT_HIP_ERROR)
    with pytest.raises(cg.CgError) as ei:
        e.run(0)
    assert ei.value.code == _lib.CG_ERR_HIP
    with pytest.raises(cg.CgError):
        e.result()  # a failed run leaves no result behind
    e.close()


def test_edge_list_ingest(oracle, tmp_path):
    """The reference's packed {src, dst, w} edge records -> CSR on the GPU:
    self-loops dropped, parallel edges merged with their smallest weight."""
    rng = np.random.default_rng(12)
    V, m = 3000, 40000
    src = rng.integers(0, V, m).astype(np.uint32)
    dst = rng.integers(0, V, m).astype(np.uint32)
    src[:500] = dst[:500]  # self-loops
    src[500:1500], dst[500:1500] = src[1500:2500], dst[1500:2500]  # parallel edges
    w = rng.integers(1, 1000, m).astype(np.uint32)
    rec = np.stack([src, dst, w], axis=1).astype("<u4")
    path = str(tmp_path / "edges.bin")
    rec.tofile(path)
    for sym in (False, True):
        g = cg.Graph.load_edge_file(path, V, has_weights=True, symmetrize=sym)
        off, d2, w2 = g.export()
        ref = oracle.csr_build(V, src, dst, symmetrize=sym)
        assert np.array_equal(off, ref.offsets) and np.array_equal(d2, ref.dest)
        # expected weight: minimum over all copies of (u, v) (and (v, u) when symmetrized)
        ss, dd, ww = (np.concatenate([src, dst]), np.concatenate([dst, src]), np.concatenate([w, w])) if sym \
            else (src, dst, w)
        best = {}
        for a, b, c in zip(ss.tolist(), dd.tolist(), ww.tolist()):
            if a != b:
                best[(a, b)] = min(best.get((a, b), 1 << 31), c)
        rows = np.repeat(np.arange(V), np.diff(off).astype(np.int64))
        assert [best[(a, b)] for a, b in zip(rows.tolist(), d2.tolist())] == w2.tolist()
        ref.weight = w2
        root = int(src[2000])
        s, _ = run(g, SSSP, root)
        assert np.array_equal(s, oracle.sssp(ref, root))
    g = cg.Graph.from_edges(V, src, dst)  # no weights
    assert np.array_equal(g.export()[1], oracle.csr_build(V, src, dst).dest)


@pytest.mark.parametrize("sched", [0, 1, 2])
@pytest.mark.parametrize("relabel", [1, 2])
def test_renumbered_engine(oracle, sched, relabel):
    """BFS / SSSP on the engine-owned copy renumbered by out-degree: values
    come back in the caller's ids, bit-exact, under the GPU, heuristic
    (host team on the renumbered copy) and learned schedules."""
    og = oracle.rmat_graph(14, 16, seed=9, weights=(1, 100))
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    rng = np.random.default_rng(relabel * 10 + sched)
    roots = [oracle.max_outdeg_vertex(og)] + [int(x) for x in rng.integers(0, og.V, 3)]
    for alg in (BFS, SSSP):
        e = cg.Engine(g, alg, sched_mode=sched, relabel=relabel)
        for r in roots:
            got, st = e.run(r)
            ref = oracle.bfs(og, r) if alg == BFS else oracle.sssp(og, r)
            assert np.array_equal(got, ref), (alg, r)
            assert np.array_equal(e.result(), ref)
            assert st["edges_traversed"] == oracle.reached_edges(og, ref)
        e.close()


def test_renumbered_lattice_and_tiny(oracle):
    """Forced renumbering on a lattice (no skew) and on a graph with isolated
    vertices; automatic mode on the same inputs."""
    og = oracle.grid_graph(40, 0.1, seed=3)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    for relabel in (0, 2):
        for alg in (BFS, SSSP):
            got, _ = run(g, alg, 37, relabel=relabel)
            assert np.array_equal(got, oracle.bfs(og, 37) if alg == BFS else oracle.sssp(og, 37))
    src = np.array([3, 3, 5, 7, 8, 8, 8], np.uint32)
    dst = np.array([5, 8, 3, 3, 3, 5, 7], np.uint32)
    w = np.array([4, 1, 2, 2, 9, 1, 1], np.uint32)
    sp = oracle.csr_build(12, src, dst)
    # csr_build sorts rows; weights follow the edge's (src, dst) order
    order = np.lexsort((dst, src))
    g2 = cg.Graph.from_csr(sp.offsets, sp.dest, w[order])
    sp_w = oracle.CSR(sp.offsets, sp.dest, w[order])
    for root in (0, 3, 8, 11):
        for alg in (BFS, SSSP):
            got, _ = run(g2, alg, root, relabel=2)
            assert np.array_equal(got, oracle.bfs(sp_w, root) if alg == BFS else oracle.sssp(sp_w, root))
    with pytest.raises(cg.CgError):
        cg.Engine(g2, BFS, relabel=3)


@pytest.mark.parametrize("min_pull", [1, 60, 250, 1000])
def test_min_pull_rule(oracle, min_pull):
    """Bottom-up SSSP / label-propagation steps chosen by the frontier-edge
    rule (GPU schedule, and renumbered engines) reach the push fixed point."""
    og = oracle.rmat_graph(14, 16, seed=4, weights=(1, 100))
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    for relabel in (1, 2):
        got, st = run(g, SSSP, root, min_pull=min_pull, relabel=relabel)
        assert np.array_equal(got, oracle.sssp(og, root)), (min_pull, relabel)
        if min_pull == 1:
            assert st["pull_iterations"] >= 1
        got, st = run(g, CC, 0, min_pull=min_pull)
        assert np.array_equal(got, oracle.cc_directed(og)), min_pull
    lat = oracle.grid_graph(50, 0.1, seed=8)
    gl = gpu_graph(lat)
    got, _ = run(gl, SSSP, 77, min_pull=min_pull, sssp_delta=300)
    assert np.array_equal(got, oracle.sssp(lat, 77))


@pytest.mark.parametrize("hub_degree,hub_delta", [(1, 1), (16, 5), (128, 1), (128, 0), (4096, 50)])
def test_sssp_hub_deferral(oracle, hub_degree, hub_delta):
    """High-degree frontier vertices wait until their distance is within
    hub_delta of the frontier's smallest: the order of relaxations changes,
    the distances stay the unique fixed point (GPU and learned schedules,
    plain and renumbered engines)."""
    cases = [random_graph(oracle, 22, 30000, 200000, hub=30000), oracle.rmat_graph(15, 16, seed=6)]
    if hub_degree > 1:
        # (every lattice vertex waiting would take one iteration per distance)
        cases.append(oracle.grid_graph(60, 0.1, seed=2))
    for og in cases:
        g = gpu_graph(og)
        for root in (oracle.max_outdeg_vertex(og), og.V // 3):
            ref = oracle.sssp(og, root)
            for mode, relabel in ((_lib.CG_SCHED_GPU, 1), (_lib.CG_SCHED_GPU, 2), (_lib.CG_SCHED_LEARNED, 1)):
                got, st = run(g, SSSP, root, sched_mode=mode, relabel=relabel, hub_degree=hub_degree,
                              hub_delta=hub_delta)
                assert np.array_equal(got, ref), (hub_degree, hub_delta, root, mode, relabel)
                assert st["edges_relaxed"] >= st["edges_traversed"]


def test_sssp_hub_deferral_accounting(oracle):
    """With every vertex waiting until its distance is the frontier's smallest
    (hub_degree 1, hub_delta 1, integer weights >= 1) each vertex is expanded
    once, at its final distance: the relaxed-edge count the engine reports
    (expansions minus the vertices kept waiting) equals the out-degree sum of
    the reached vertices, and the BFS schedule ignores the option."""
    og = oracle.rmat_graph(13, 16, seed=3, weights=(1, 20))
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    got, st = run(g, SSSP, root, hub_degree=1, hub_delta=1)
    assert np.array_equal(got, oracle.sssp(og, root))
    assert st["edges_relaxed"] == st["edges_traversed"] == oracle.reached_edges(og, got)
    _, plain = run(g, SSSP, root)
    assert plain["edges_relaxed"] > st["edges_relaxed"]
    got, _ = run(g, BFS, root, hub_degree=1, hub_delta=1)
    assert np.array_equal(got, oracle.bfs(og, root))
    with pytest.raises(cg.CgError):
        cg.Engine(g, SSSP, hub_degree=64, sssp_delta=10)


@pytest.mark.parametrize("cap", [1, 7, 100000])
def test_expand_grid_caps(oracle, cap):
    """However few workgroups the push expansion gets, its wavefronts cover
    every work item."""
    og = random_graph(oracle, 12, 20000, 150000, hub=20000)
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    for alg, ref in ((BFS, oracle.bfs(og, root)), (SSSP, oracle.sssp(og, root)), (CC, oracle.cc_directed(og))):
        got, _ = run(g, alg, root, expand_grid=cap, sched_mode=_lib.CG_SCHED_FORCED, forced_choice=2)
        assert np.array_equal(got, ref), (alg, cap)
        got, _ = run(g, alg, root, expand_grid=cap)
        assert np.array_equal(got, ref), (alg, cap)


def test_gpu_schedule_pipelines(oracle):
    """The GPU schedule's queued-ahead pipeline (statistics, compaction and
    expansion issued without waiting for the host), with and without
    delta-stepping."""
    og = random_graph(oracle, 21, 30000, 200000, hub=30000)
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    for alg, ref in ((BFS, oracle.bfs(og, root)), (SSSP, oracle.sssp(og, root)), (CC, oracle.cc_directed(og))):
        got, _ = run(g, alg, root)
        assert np.array_equal(got, ref), alg
    for delta in (1, 40, 5000):
        got, _ = run(g, SSSP, root, sssp_delta=delta)
        assert np.array_equal(got, oracle.sssp(og, root)), delta
    lat = oracle.grid_graph(70, 0.1, seed=9)
    gl = gpu_graph(lat)
    for alg in (BFS, SSSP):
        got, _ = run(gl, alg, 3, sssp_delta=200)
        assert np.array_equal(got, oracle.bfs(lat, 3) if alg == BFS else oracle.sssp(lat, 3))


def test_back_to_back_runs(oracle):
    """A run that leaves only no-op iterations on its stream returns without
    draining it; the next run is queued behind them, and the records they
    publish late carry the earlier run's tag.  Successive queries with
    different roots, their values read or not in between, each give that
    root's values and counts, on the caller's ids and on a renumbered copy,
    and device-pointer reads see the finished values."""
    og = random_graph(oracle, 23, 20000, 160000, hub=20000)
    g = gpu_graph(og)
    roots = [oracle.max_outdeg_vertex(og), 0, 17, 19999, 5]
    for alg in (BFS, SSSP, CC):
        for relabel in ((1, 2) if alg != CC else (1,)):  # never / always renumbered
            e = cg.Engine(g, alg, sched_mode=_lib.CG_SCHED_GPU, relabel=relabel)
            for i, r in enumerate(roots * 2):
                ref = oracle.bfs(og, r) if alg == BFS else (oracle.sssp(og, r) if alg == SSSP else None)
                if alg == CC:
                    ref = oracle.cc_directed(og)
                got, st = e.run(r, copy_out=(i % 3 == 2))
                if got is not None:
                    assert np.array_equal(got, ref), (alg, relabel, r)
                if alg != CC:
                    assert st["edges_traversed"] == oracle.reached_edges(og, ref), (alg, relabel, r)
            assert np.array_equal(e.result(), ref), (alg, relabel)
            e.close()
    g.close()


@pytest.mark.parametrize("wmax", [255, 256, 65535, 65536, 10**8])
def test_sssp_weight_widths(oracle, wmax):
    """The push expansion of SSSP reads an 8-bit or 16-bit copy of the weights
    when every weight fits (largest 255 / 65535), the 32-bit array otherwise:
    distances bit-exact against the oracle at each boundary, under the GPU
    schedule with and without hub deferral, on the caller's ids and on a
    renumbered copy."""
    og = random_graph(oracle, 31, 30000, 240000, hub=20000)
    rng = np.random.default_rng(wmax % 1000)
    og.weight = rng.integers(1, wmax + 1, og.E).astype(np.uint32)
    og.weight[0] = wmax  # the largest weight is present
    g = gpu_graph(og)
    root = oracle.max_outdeg_vertex(og)
    ref = oracle.sssp(og, root)
    for opt in ({}, {"hub_degree": 16, "hub_delta": 3}, {"relabel": 2}):
        got, st = run(g, SSSP, root, sched_mode=_lib.CG_SCHED_GPU, **opt)
        assert np.array_equal(got, ref), (wmax, opt)
        assert st["edges_traversed"] == oracle.reached_edges(og, ref)
    g.close()


@pytest.mark.parametrize("tiers", [(4, 32, 64), (16, 1024, 2048), (64, 128, 64)])
def test_propagated_vertices_dropped(oracle, tiers):
    """SSSP and CC under the GPU schedule skip a frontier vertex whose value is
    one its out-edges were already all relaxed with (the expansion records
    it; a vertex split into chunks keeps the largest value its chunks read).
    Tiers that put many vertices into chunks, with and without hub deferral,
    and several roots on one engine (the records are cleared per query):
    values bit-exact against the oracle, reached edges exact."""
    og = random_graph(oracle, 37, 25000, 220000, hub=25000)
    g = gpu_graph(og)
    bw, bc, ce = tiers
    roots = [oracle.max_outdeg_vertex(og), 3, 24999]
    for alg in (SSSP, CC):
        for hub in ((0, 1), (8, 2)) if alg == SSSP else ((0, 1),):
            e = cg.Engine(g, alg, sched_mode=_lib.CG_SCHED_GPU, bin_wave=bw, bin_chunk=bc, chunk_edges=ce,
                          hub_degree=hub[0], hub_delta=hub[1])
            for r in roots + roots[:1]:
                ref = oracle.sssp(og, r) if alg == SSSP else oracle.cc_directed(og)
                got, st = e.run(r)
                assert np.array_equal(got, ref), (alg, hub, r)
                if alg == SSSP:
                    assert st["edges_traversed"] == oracle.reached_edges(og, ref), (hub, r)
            e.close()
    g.close()
