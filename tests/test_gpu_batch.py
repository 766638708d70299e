"""Batched queries (one HIP-graph-captured level loop for many queries over
uneven graphs): every query's output is bit-exact against the oracle and
identical to the one-query-per-launch path."""
import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import batch

pytestmark = pytest.mark.gpu


def test_batch_matches_oracle(oracle):
    ogs = [oracle.rmat_graph(s, 16, seed=s, weights=(1, 1000)) for s in (9, 11, 13)]
    ogs += [oracle.grid_graph(n, 0.1, seed=n) for n in (20, 45)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(0)
    qg, qo, qa, qr = [], [], [], []
    for i in range(40):
        k = i % len(gs)
        qg.append(gs[k])
        qo.append(ogs[k])
        qa.append(i % 2)
        qr.append(int(rng.integers(0, ogs[k].V)))
    outs, st = batch.run_batch(qg, qa, qr, mode=batch.BATCHED, batch_levels=4, batch_grid=64)
    for o, a, r, got in zip(qo, qa, qr, outs):
        ref = oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)
        assert np.array_equal(got, ref), (a, r)
    assert st["edges_traversed"] == sum(oracle.reached_edges(o, x) for o, x in zip(qo, outs))
    outs1, st1 = batch.run_batch(qg, qa, qr, mode=batch.ONE_PER_LAUNCH)
    for a, b in zip(outs, outs1):
        assert np.array_equal(a, b)
    assert st1["edges_traversed"] == st["edges_traversed"]


def test_batches_from_two_threads(oracle):
    """Two batches over the same graphs at once, from two host threads: the
    graphs' engines serve one query at a time, so both batches' outputs are
    the oracle's."""
    import threading
    ogs = [oracle.rmat_graph(12, 16, seed=6, weights=(1, 1000)), oracle.grid_graph(30, 0.1, seed=6)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(9)
    qs = []
    for t in range(2):
        qg, qo, qa, qr = [], [], [], []
        for i in range(20):
            k = i % 2
            qg.append(gs[k])
            qo.append(ogs[k])
            qa.append((i // 2) % 2)
            qr.append(int(rng.integers(0, ogs[k].V)))
        qs.append((qg, qo, qa, qr))
    results = [None, None]

    def work(t):
        qg, _, qa, qr = qs[t]
        results[t] = batch.run_batch(qg, qa, qr, batch_engine_v=1, batch_levels=4, batch_grid=64)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for t in range(2):
        _, qo, qa, qr = qs[t]
        outs, st = results[t]
        assert st["heavy_iterations"] == 10
        for o, a, r, got in zip(qo, qa, qr, outs):
            assert np.array_equal(got, oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)), (t, a, r)
    for g in gs:
        g.close()


@pytest.mark.parametrize("engine_v", [1, 3000, 0xFFFFFFFF])
def test_batch_engine_split(oracle, engine_v):
    """Queries on skewed graphs of at least batch_engine_v vertices run on an
    engine of their graph (a host thread of their own, values copied into
    the batch's slots) while the rest run level by level: every output is
    the oracle's, the same with every query batched, repeated roots on one
    engine included, and the counts add up."""
    ogs = [oracle.rmat_graph(s, 16, seed=s, weights=(1, 1000)) for s in (10, 12)]
    ogs += [oracle.grid_graph(n, 0.1, seed=n) for n in (25, 40)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(3)
    qg, qo, qa, qr = [], [], [], []
    for i in range(36):
        k = i % len(gs)
        qg.append(gs[k])
        qo.append(ogs[k])
        qa.append((i // 4) % 2)
        qr.append(int(rng.integers(0, ogs[k].V)) if i % 9 else 0)
    outs, st = batch.run_batch(qg, qa, qr, batch_engine_v=engine_v, batch_levels=4, batch_grid=64)
    for o, a, r, got in zip(qo, qa, qr, outs):
        ref = oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)
        assert np.array_equal(got, ref), (engine_v, a, r)
    assert st["edges_traversed"] == sum(oracle.reached_edges(o, x) for o, x in zip(qo, outs))
    def small_world(o):  # largest out-degree >= 16 x the average, rounded up
        deg = np.diff(o.offsets.astype(np.int64))
        return deg.max() >= 16 * max(-(-o.E // o.V), 1)
    assert small_world(ogs[0]) and small_world(ogs[1]) and not small_world(ogs[2]) and not small_world(ogs[3])
    on_engines = sum(1 for o in qo if engine_v != 0xFFFFFFFF and o.V >= engine_v and small_world(o))
    assert st["heavy_iterations"] == on_engines
    outs_b, st_b = batch.run_batch(qg, qa, qr, batch_engine_v=0xFFFFFFFF, batch_levels=4, batch_grid=64)
    assert st_b["heavy_iterations"] == 0
    for a, b in zip(outs, outs_b):
        assert np.array_equal(a, b)
    for g in gs:
        g.close()


@pytest.mark.parametrize("block,grid,levels,delta,group", [(64, 1, 1, 0, 64), (1024, 512, 32, 0, 1),
                                                           (256, 64, 8, 1, 16), (256, 2048, 16, 700, 32),
                                                           (512, 128, 4, 10**7, 7)])
def test_batch_launch_shapes(oracle, block, grid, levels, delta, group):
    o = oracle.grid_graph(30, 0.1, seed=5)
    g = cg.Graph.from_csr(o.offsets, o.dest, o.weight)
    roots = [0, 17, 899, 450]
    outs, _ = batch.run_batch([g] * 4, [0, 1, 0, 1], roots, batch_block=block, batch_grid=grid,
                              batch_levels=levels, sssp_delta=delta, batch_group=group)
    for a, r, got in zip([0, 1, 0, 1], roots, outs):
        assert np.array_equal(got, oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r))


def test_batch_rejects_bad_queries(oracle):
    o = oracle.rmat_graph(8, 4, seed=1, weights=None)
    g = cg.Graph.from_csr(o.offsets, o.dest)
    with pytest.raises(cg.CgError):
        batch.run_batch([g], [1], [0])  # SSSP without weights
    with pytest.raises(cg.CgError):
        batch.run_batch([g], [0], [o.V])  # root out of range
    with pytest.raises(cg.CgError):
        batch.run_batch([g], [2], [0])  # CC is not a batched query


def test_batch_delta_stepping_mixed(oracle):
    ogs = [oracle.rmat_graph(12, 16, seed=3, weights=(1, 1000)), oracle.grid_graph(64, 0.1, seed=2)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(4)
    qg, qo, qa, qr = [], [], [], []
    for i in range(16):
        k = i % 2
        qg.append(gs[k])
        qo.append(ogs[k])
        qa.append(1 if i >= 4 else 0)
        qr.append(int(rng.integers(0, ogs[k].V)))
    for delta in (250, 4000):
        outs, st = batch.run_batch(qg, qa, qr, sssp_delta=delta, batch_levels=8)
        for o, a, r, got in zip(qo, qa, qr, outs):
            assert np.array_equal(got, oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)), (delta, a, r)


def test_engine_handle_batch(oracle):
    """cg_engine_run_batch: engines repeat in the list, each query takes its
    engine's algorithm and graph; a PR engine is refused."""
    ogs = [oracle.rmat_graph(10, 8, seed=7, weights=(1, 100)), oracle.grid_graph(24, 0.1, seed=3)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    engs = [cg.Engine(gs[0], 0), cg.Engine(gs[0], 1, relabel=2), cg.Engine(gs[1], 1)]
    owner = [0, 1, 2, 1, 0, 2, 2]
    rng = np.random.default_rng(11)
    roots = [int(rng.integers(0, ogs[0 if k < 2 else 1].V)) for k in owner]
    outs, st = batch.run_engine_batch([engs[k] for k in owner], roots)
    for k, r, got in zip(owner, roots, outs):
        o = ogs[0 if k < 2 else 1]
        assert np.array_equal(got, oracle.bfs(o, r) if k == 0 else oracle.sssp(o, r)), (k, r)
    assert st["edges_traversed"] == sum(oracle.reached_edges(ogs[0 if k < 2 else 1], x)
                                        for k, x in zip(owner, outs))
    pr = cg.Engine(gs[0], 3)
    with pytest.raises(cg.CgError):
        batch.run_engine_batch([engs[0], pr], [0, 0])


def test_batch_renumbered_graph(oracle):
    """A skewed graph of 2^20 vertices runs its batched queries on a copy
    renumbered by out-degree (built on first use, kept with the graph); roots
    map in and values map back, so the outputs and traversed-edge counts are
    the oracle's, on the first batch and on a second one, mixed with a small
    graph that is not renumbered."""
    big = oracle.rmat_graph(20, 16, seed=4, weights=(1, 1000))
    small = oracle.rmat_graph(12, 16, seed=4, weights=(1, 1000))
    gb = cg.Graph.from_csr(big.offsets, big.dest, big.weight)
    gs = cg.Graph.from_csr(small.offsets, small.dest, small.weight)
    rng = np.random.default_rng(7)
    roots = [oracle.max_outdeg_vertex(big)] + [int(x) for x in rng.integers(0, big.V, 3)]
    for rep in range(2):
        qg = [gb] * 4 + [gs] * 2
        qo = [big] * 4 + [small] * 2
        qa = [0, 1, 0, 1, 0, 1]
        qr = (roots if rep == 0 else roots[::-1]) + [3, 5]
        outs, st = batch.run_batch(qg, qa, qr, mode=batch.BATCHED)
        for o, a, r, got in zip(qo, qa, qr, outs):
            ref = oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)
            assert np.array_equal(got, ref), (rep, a, r)
        assert st["edges_traversed"] == sum(oracle.reached_edges(o, x) for o, x in zip(qo, outs))


@pytest.mark.parametrize("hub", [(1, 1), (16, 20), (64, 200), (8, 5)])
@pytest.mark.parametrize("delta", [0, 3000])
def test_batch_hub_deferral(oracle, hub, delta):
    """SSSP hub deferral in the batch (a frontier vertex of out-degree >=
    hub_degree waits while its distance is >= its query's frontier minimum +
    hub_delta), alone and with delta-stepping: every output bit-exact, and
    BFS queries in the same batch unaffected."""
    ogs = [oracle.rmat_graph(s, 16, seed=s + 3, weights=(1, 1000)) for s in (10, 13)]
    ogs += [oracle.grid_graph(40, 0.1, seed=2)]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(hub[0] + delta)
    qg, qo, qa, qr = [], [], [], []
    for i in range(24):
        k = i % len(gs)
        qg.append(gs[k])
        qo.append(ogs[k])
        qa.append(1 if i % 3 else 0)
        qr.append(int(rng.integers(0, ogs[k].V)))
    outs, st = batch.run_batch(qg, qa, qr, hub_degree=hub[0], hub_delta=hub[1], sssp_delta=delta,
                               batch_levels=8, batch_grid=256)
    for o, a, r, got in zip(qo, qa, qr, outs):
        assert np.array_equal(got, oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)), (a, r, hub, delta)
    assert st["edges_traversed"] == sum(oracle.reached_edges(o, x) for o, x in zip(qo, outs))


def test_batch_many_queries(oracle):
    """More queries than the one-kernel-per-level form takes (4096): the
    batch plans each level in a kernel of its own; every output is the
    oracle's."""
    ogs = [oracle.grid_graph(9, 0.1, seed=1), oracle.rmat_graph(7, 4, seed=2, weights=(1, 50))]
    gs = [cg.Graph.from_csr(o.offsets, o.dest, o.weight) for o in ogs]
    rng = np.random.default_rng(5)
    qg, qo, qa, qr = [], [], [], []
    for i in range(4200):
        k = i % 2
        qg.append(gs[k])
        qo.append(ogs[k])
        qa.append((i // 2) % 2)
        qr.append(int(rng.integers(0, ogs[k].V)))
    outs, st = batch.run_batch(qg, qa, qr, batch_levels=8, batch_grid=256)
    refs = {}
    for o, a, r, got in zip(qo, qa, qr, outs):
        key = (id(o), a, r)
        if key not in refs:
            refs[key] = oracle.bfs(o, r) if a == 0 else oracle.sssp(o, r)
        assert np.array_equal(got, refs[key]), (a, r)
    assert st["edges_traversed"] == sum(oracle.reached_edges(o, x) for o, x in zip(qo, outs))
    for g in gs:
        g.close()
