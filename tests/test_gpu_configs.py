"""Every benchmark configuration at its stated size, with the options the
benchmark runs (tuning table, learned schedule, the reference's LIGHT/HEAVY
rule with the host team), plus the cases the smaller suites leave open:
the expansion grid's stride path, a zero-out-degree tail (V' < V) on the
GPU, and the autotuner itself.

* BFS R-MAT s18 ef16: bit-exact against the oracle's queue BFS for the tuned
  entry, the learned schedule and the heuristic schedule.
* SSSP R-MAT s24 ef16: the tuned entry (hub deferral, tiers, expansion grid),
  the learned schedule and the heuristic (co-executed) schedule, each checked
  by the shortest-path certificate and against the restated reference CPU
  engine's output.
* The batched workload in full (256 BFS + 256 SSSP queries over R-MAT s14-s22
  and lattices of 256^2-2048^2 vertices, one batch): one BFS and one SSSP per
  graph bit-exact against the oracle, every query by the certificate, and the
  TEPS numerator against the oracle's count.
"""
import os

import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import autotune, batch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MLP = os.path.join(os.path.dirname(HERE), "cggraph_amd", "data", "sched_mlp.bin")


def host_csr(oracle, g):
    off, dst, w = g.export()
    return oracle.CSR(off, dst, w)


def tuned(g, alg, kind="rmat"):
    inf = g.info()
    return autotune.best_options(autotune.engine_signature(kind, alg, inf["V"], inf["E"]))


@pytest.fixture(scope="module")
def rmat24(oracle):
    g = cg.Graph.rmat(24, 16, seed=1, weights=(1, 100))
    og = host_csr(oracle, g)
    root = g.info()["max_outdeg_vertex"]
    ref, _, _ = oracle.multicore_run(1, og, root, threads=16)
    assert oracle.check_sssp_cert(og, root, ref) == 0
    yield g, og, root, ref
    g.close()


@pytest.mark.parametrize("schedule", ["tuned", "learned", "heuristic"])
def test_sssp_rmat24_schedules(oracle, rmat24, schedule):
    g, og, root, ref = rmat24
    opt = tuned(g, 1)
    assert opt.get("hub_degree") and opt.get("expand_grid"), "the tuning table has the s24 SSSP entry"
    if schedule == "tuned":
        e = cg.Engine(g, 1, sched_mode=1, **opt)
    elif schedule == "learned":
        e = cg.Engine(g, 1, sched_mode=2, mlp_path=MLP, **opt)
    else:
        e = cg.Engine(g, 1, sched_mode=0, cpu_threads=16)
    got, st = e.run(root)
    if schedule == "heuristic":
        assert st["light_iterations"] >= 1 and st["heavy_iterations"] >= 1
        assert st["cpu_vertices"] > 0 and st["gpu_vertices"] > 0
    got2, _ = e.run(root)
    e.close()
    assert np.array_equal(got, ref), schedule
    assert np.array_equal(got2, ref), schedule
    assert st["edges_traversed"] == oracle.reached_edges(og, got)


@pytest.mark.parametrize("schedule", ["tuned", "learned", "heuristic"])
def test_bfs_rmat18_schedules(oracle, schedule):
    og = oracle.rmat_graph(18, 16, seed=1, weights=None)
    g = cg.Graph.rmat(18, 16, seed=1, weights=None)
    root = oracle.max_outdeg_vertex(og)
    if schedule == "tuned":
        opt = tuned(g, 0)
        assert opt, "the tuning table has the s18 BFS entry"
        e = cg.Engine(g, 0, sched_mode=1, **opt)
    elif schedule == "learned":
        e = cg.Engine(g, 0, sched_mode=2, mlp_path=MLP)
    else:
        e = cg.Engine(g, 0, sched_mode=0, cpu_threads=16)
    ref = oracle.bfs(og, root)
    for _ in range(2):
        got, st = e.run(root)
        assert np.array_equal(got, ref), schedule
        assert st["edges_traversed"] == oracle.reached_edges(og, got)
    e.close()
    g.close()


@pytest.mark.parametrize("grid", [1, 7, 4096, 16384])
@pytest.mark.parametrize("alg", [0, 1, 2])
def test_expand_grid_stride(oracle, alg, grid):
    """Expansion grids from one workgroup (every wavefront strides over many
    items) to far more workgroups than items."""
    og = oracle.rmat_graph(15, 12, seed=21, symmetrize=(alg == 2))
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    for hub in ((0, 0), (64, 2)) if alg == 1 else ((0, 0),):
        e = cg.Engine(g, alg, sched_mode=1, expand_grid=grid, bin_wave=16, bin_chunk=256, chunk_edges=128,
                      hub_degree=hub[0], hub_delta=hub[1], relabel=1)
        got, _ = e.run(root)
        e.close()
        ref = oracle.bfs(og, root) if alg == 0 else (oracle.sssp(og, root) if alg == 1 else oracle.cc_undirected(og))
        assert np.array_equal(got, ref), (alg, grid, hub)
    g.close()


def tail_graph(oracle, seed, V=40000, tail=9000, m=300000):
    """Vertices [V - tail, V) have no out-edges (RCM_LAST_ZERO_OUTDEGREE);
    a few sources are hubs so the heuristic sees HEAVY iterations."""
    rng = np.random.default_rng(seed)
    src = np.concatenate([rng.integers(0, 50, m // 3), rng.integers(0, V - tail, m)]).astype(np.uint32)
    dst = rng.integers(0, V, src.size).astype(np.uint32)
    og = oracle.csr_build(V, src, dst)
    og.weight = rng.integers(1, 101, og.E).astype(np.uint32)
    assert int(og.offsets[V - tail]) == og.E
    return og


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_zero_outdeg_tail_gpu(oracle, mode):
    """V' < V on the GPU: BFS / SSSP from inside V' and from the tail, CC
    initialised over V', under the heuristic (host team and co-execution),
    GPU and learned schedules, on the graph and on a renumbered copy."""
    og = tail_graph(oracle, 4 + mode)
    V, tail = og.V, og.V - int(np.searchsorted(og.offsets, og.E))
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight, zero_outdeg_tail=9000)
    extra = {"mlp_path": MLP} if mode == 2 else {}
    for relabel in (1, 2):
        for alg in (0, 1):
            e = cg.Engine(g, alg, sched_mode=mode, relabel=relabel, cpu_threads=8,
                          heavy_active_vertices=64, heavy_active_edges=2000, **extra)
            for root in (0, 17, V - 1):
                got, _ = e.run(root)
                ref = oracle.bfs(og, root) if alg == 0 else oracle.sssp(og, root)
                assert np.array_equal(got, ref), (mode, relabel, alg, root)
            e.close()
    e = cg.Engine(g, 2, sched_mode=mode, cpu_threads=8, heavy_active_vertices=64, heavy_active_edges=2000, **extra)
    got, _ = e.run(0)
    e.close()
    assert np.array_equal(got, oracle.cc_directed(og))
    assert tail >= 9000
    g.close()


def test_heuristic_co_execution_split(oracle):
    """HEAVY iterations shared with the host team: the GPU agent claims from
    the head of the small tier, the host from the tail, with the host's region
    sized by the split model (larger when the GPU is reported slow).  Outputs
    are bit-exact whichever side got how much."""
    og = oracle.rmat_graph(16, 16, seed=9)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    ref = oracle.sssp(og, root)
    for fault in (0, 4):
        e = cg.Engine(g, 1, sched_mode=0, cpu_threads=8, heavy_active_vertices=256, heavy_active_edges=4096,
                      fault_flags=fault)
        for _ in range(3):
            got, st = e.run(root)
            assert np.array_equal(got, ref), fault
            assert st["heavy_iterations"] >= 1 and st["cpu_vertices"] > 0 and st["gpu_vertices"] > 0
        e.close()


def test_autotuner_tiny_space(oracle, tmp_path):
    """tune_engine / tune_batch over a two-value space on a tiny graph: the
    table entry is written, and the tuned options run bit-exact."""
    og = oracle.rmat_graph(12, 8, seed=2)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    table = str(tmp_path / "tuning.json")
    sig, rec = autotune.tune_engine(g, 1, root, kind="test", reps=1, path=table,
                                    space={"bin_wave": [16, 64], "expand_grid": [0, 7]})
    assert sig.startswith("engine/test/alg1/") and set(rec["options"]) == {"bin_wave", "expand_grid"}
    opt = autotune.best_options(sig, path=table)
    assert opt == rec["options"]
    e = cg.Engine(g, 1, **opt)
    assert np.array_equal(e.run(root)[0], oracle.sssp(og, root))
    e.close()
    qg, qa, qr = [g] * 4, [0, 1, 0, 1], [root, root, 1, 2]
    bsig, brec = autotune.tune_batch(qg, qa, qr, signature="batch/test", path=table)
    assert autotune.best_options("batch/test", path=table) == brec["options"]
    outs, _ = batch.run_batch(qg, qa, qr, **brec["options"])
    for a, r, got in zip(qa, qr, outs):
        assert np.array_equal(got, oracle.bfs(og, r) if a == 0 else oracle.sssp(og, r))
    g.close()


# ---- the batched workload in full ----

@pytest.fixture(scope="module")
def mixed(oracle):
    graphs, qg, qa, qr = batch.mixed_workload()
    assert len(qr) == 512 and len(graphs) == 13
    opts = autotune.best_options("batch/mixed") or {"sssp_delta": 10000}
    print(f"batched mixed workload: {len(qr)} queries over {len(graphs)} graphs, options {opts}", flush=True)
    outs, st = batch.run_batch(qg, qa, qr, **opts)
    yield graphs, qg, qa, qr, outs, st
    for g in graphs:
        g.close()


@pytest.mark.parametrize("gi", range(13))
def test_mixed_batch_full_size(oracle, mixed, gi):
    graphs, qg, qa, qr, outs, st = mixed
    g = graphs[gi]
    og = host_csr(oracle, g)
    mine = [i for i in range(len(qr)) if qg[i] is g]
    assert mine
    checked_exact = set()
    for i in mine:
        a, r, got = qa[i], qr[i], outs[i]
        if a not in checked_exact:
            ref = oracle.bfs(og, r) if a == 0 else oracle.sssp(og, r)
            assert np.array_equal(got, ref), (gi, i, a, r)
            checked_exact.add(a)
        else:
            assert oracle.check_sssp_cert(og, r, got, unit=(a == 0)) == 0, (gi, i, a, r)
    assert checked_exact == {0, 1}


def test_mixed_batch_teps_numerator(oracle, mixed):
    graphs, qg, qa, qr, outs, st = mixed
    csr = {id(g): host_csr(oracle, g) for g in graphs}
    total = sum(oracle.reached_edges(csr[id(g)], o) for g, o in zip(qg, outs))
    assert st["edges_traversed"] == total
