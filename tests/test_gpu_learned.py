"""Learned scheduler on the GPU: the HIP forward/backward kernels against the
CPU oracle (oracle/mlp_ref.py), probe-trace collection, training, and what the
trained network gains over the fixed rules -- with outputs identical to the
oracle whatever it chooses."""
import os

import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib, scheduler
from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def test_forward_matches_oracle():
    rng = np.random.default_rng(1)
    P = rng.normal(0, 0.3, 6032).astype(np.float32)
    X = rng.uniform(0, 1, (300, 12)).astype(np.float32)
    got = scheduler.mlp_forward(P, X)
    ref, _ = mlp_ref.forward(P.astype(np.float64), X.astype(np.float64))
    assert np.allclose(got, ref, atol=2e-6, rtol=1e-4)


def test_gradient_matches_oracle():
    rng = np.random.default_rng(2)
    P = rng.normal(0, 0.3, 6032).astype(np.float32)
    X = rng.uniform(0, 1, (257, 12)).astype(np.float32)  # several 64-sample chunks, ragged tail
    Y = rng.integers(0, 16, 257).astype(np.int32)
    loss, g = scheduler.mlp_loss_grad(P, X, Y)
    rl, rg = mlp_ref.loss_grad(P.astype(np.float64), X.astype(np.float64), Y)
    assert abs(loss - rl) <= 1e-5 * max(1.0, abs(rl))
    assert np.allclose(g, rg, atol=2e-6, rtol=2e-4), float(np.max(np.abs(g - rg)))


def test_nan_weights_fall_back(oracle):
    og = oracle.rmat_graph(12, 8, seed=3)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    e = cg.Engine(g, 1, sched_mode=_lib.CG_SCHED_LEARNED, fault_flags=_lib.CG_FAULT_MLP_NAN)
    got, st = e.run(root)
    e.close()
    assert np.array_equal(got, oracle.sssp(og, root))
    assert st["learned_fallbacks"] == st["iterations"]


def test_probe_train_and_gain(oracle, tmp_path):
    trace = str(tmp_path / "probe.jsonl")
    weights = str(tmp_path / "sched.bin")
    # training graphs: R-MAT scales 12-14 (seeds 1, 2) and a lattice
    train = [cg.Graph.rmat(s, 16, seed=sd) for s in (12, 14) for sd in (1, 2)] + [cg.Graph.grid(96, 0.1, seed=1)]
    for g in train:
        r = g.info()["max_outdeg_vertex"]
        for alg in (0, 1):
            scheduler.collect_probe_traces(g, alg, [r], trace, cpu_threads=4)
    recs = scheduler.read_probe_records(trace)
    assert len(recs) > 20
    X, Y, T, _ = scheduler.dataset(recs)
    loss = scheduler.train(trace, weights, epochs=600, lr=3e-3, seed=7)
    P = scheduler.load_weights(weights)
    l0, _ = mlp_ref.loss_grad(np.zeros(6032), X.astype(np.float64), Y)  # uniform predictor
    assert loss < l0, (loss, l0)
    # held-out seed: probe it, then compare the probe-measured cost of each policy
    held = str(tmp_path / "held.jsonl")
    og = oracle.rmat_graph(14, 16, seed=9)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    root = oracle.max_outdeg_vertex(og)
    for alg in (0, 1):
        scheduler.collect_probe_traces(g, alg, [root], held, cpu_threads=4)
    hr = scheduler.read_probe_records(held)
    Xh, Yh, Th, Rh = scheduler.dataset(hr)
    learned = np.argmax(scheduler.mlp_forward(P, Xh), axis=1)
    heur = [scheduler.heuristic_choice(r) for r in Rh]
    t_learn, t_heur, t_best = (scheduler.chosen_time(Th, learned), scheduler.chosen_time(Th, heur),
                               scheduler.chosen_time(Th, Yh))
    print(f"held-out probe cost: learned {t_learn:.0f} us, heuristic {t_heur:.0f} us, oracle-best {t_best:.0f} us")
    assert t_learn <= t_heur
    # the learned engine's outputs are bit-exact, whatever it chose
    for alg, ref in ((0, oracle.bfs(og, root)), (1, oracle.sssp(og, root))):
        e = cg.Engine(g, alg, sched_mode=_lib.CG_SCHED_LEARNED, mlp_path=weights, cpu_threads=4)
        got, st = e.run(root)
        e.close()
        assert np.array_equal(got, ref)
        assert st["learned_fallbacks"] <= st["iterations"]
