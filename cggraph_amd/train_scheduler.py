"""Collects probe traces, trains the learned scheduler on the GPU and writes
cggraph_amd/data/sched_mlp.bin, plus a held-out evaluation.

  python -m cggraph_amd.train_scheduler [--quick] [--out PATH] [--trace PATH]

Training workloads: R-MAT scales 18-24 (edge factor 16, weights U[1,100],
seeds 2 and 3) and 4-neighbour lattices of 1024^2-4096^2 vertices (10% of
edges dropped, weights U[1,1000], seeds 2 and 3); BFS, SSSP and directed CC
(label propagation) from the max-out-degree vertex.  Held-out evaluation:
seed 1 of the same families (the benchmark graph's seed).  On lattices,
whose queries run thousands of iterations, only iteration pairs (k, k + 1)
with k a multiple of the probe stride are probed (`probe_stride`); the
network picks iteration k + 1's schedule from iteration k's frontier, so a
training sample needs both.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _lib, scheduler
from .core import Engine, Graph

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "sched_mlp.bin")


def workloads(seeds, scales, sides):
    for sd in seeds:
        for s in scales:
            yield f"rmat{s}", sd, 1, (lambda s=s, sd=sd: Graph.rmat(s, 16, seed=sd, weights=(1, 100)))
        for n in sides:
            stride = max(1, n // 128)
            yield f"grid{n}", sd, stride, (lambda n=n, sd=sd: Graph.grid(n, 0.1, seed=sd, weights=(1, 1000)))


def probe(trace, seeds, scales, sides, algs=(0, 1, 2), log=print):
    for name, sd, stride, make in workloads(seeds, scales, sides):
        g = make()
        r = g.info()["max_outdeg_vertex"]
        for alg in algs:
            t = time.time()
            e = Engine(g, alg, sched_mode=_lib.CG_SCHED_PROBE, trace_path=trace, probe_stride=stride)
            try:
                _, st = e.run(int(r), copy_out=False)
            finally:
                e.close()
            log(f"probe {name}/seed{sd} alg={alg} stride={stride}: {st['iterations']} iterations, "
                f"{time.time() - t:.1f}s", flush=True)
        g.close()


def evaluate(weights, trace):
    recs = scheduler.read_probe_records(trace)
    X, Y, T, R = scheduler.dataset(recs)
    P = scheduler.load_weights(weights)
    learned = np.argmax(scheduler.mlp_forward(P, X), axis=1)
    heur = [scheduler.heuristic_choice(r) for r in R]
    res = {"records": len(recs), "pairs": len(R), "accuracy": float(np.mean(learned == Y)),
           "us_learned": scheduler.chosen_time(T, learned), "us_heuristic": scheduler.chosen_time(T, heur),
           "us_gpu_push_default": scheduler.chosen_time(T, [0] * len(R)),
           "us_best": float(np.where(T > 0, T, np.inf).min(axis=1).sum())}
    # per family (graph kind and size) and algorithm
    fam = {}
    for i, r in enumerate(R):
        k = f"{'grid' if r['E'] < 5 * r['V'] else 'rmat'}/V{r['V']}/alg{r['alg']}"
        f = fam.setdefault(k, {"pairs": 0, "us_learned": 0.0, "us_gpu_push_default": 0.0, "us_best": 0.0})
        f["pairs"] += 1
        f["us_learned"] += float(T[i, learned[i]])
        f["us_gpu_push_default"] += float(T[i, 0])
        f["us_best"] += float(np.where(T[i] > 0, T[i], np.inf).min())
    res["families"] = fam
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--trace", default="sched_train.jsonl")
    ap.add_argument("--held", default="sched_held.jsonl")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--no-probe", action="store_true", help="train on existing trace files")
    args = ap.parse_args()
    scales = (14, 16) if args.quick else (18, 20, 22, 24)
    sides = (256,) if args.quick else (1024, 2048, 4096)
    if not args.no_probe:
        for p in (args.trace, args.held):
            if os.path.exists(p):
                os.remove(p)
        probe(args.trace, (2, 3), scales, sides)
        probe(args.held, (1,), scales, sides)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    t = time.time()
    loss = scheduler.train(args.trace, args.out, epochs=args.epochs, lr=3e-3, seed=1)
    print(f"trained in {time.time() - t:.1f}s, loss {loss:.4f}", flush=True)
    res = {"train_loss": loss, "train": evaluate(args.out, args.trace), "held_out": evaluate(args.out, args.held)}
    print(json.dumps(res, indent=1))
    with open(os.path.splitext(args.out)[0] + "_eval.json", "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
