"""Collects probe traces, trains the learned scheduler on the GPU and writes
cggraph_amd/data/sched_mlp.bin, plus a held-out evaluation.

  python -m cggraph_amd.train_scheduler [--quick] [--out PATH] [--trace PATH]

Training workloads: R-MAT scales 16-20 (edge factor 16, seeds 1-2) and
4-neighbour lattices of 256^2-1024^2 vertices (10% of edges dropped), BFS,
SSSP (weights U[1,100] / U[1,1000]) and directed-CC queries from the
max-out-degree vertex.  Held-out evaluation: seed 3 of the same families.
"""
import argparse
import json
import os
import time

import numpy as np

from . import _lib, scheduler
from .core import Graph

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "sched_mlp.bin")


def workloads(seeds, scales, sides):
    for sd in seeds:
        for s in scales:
            yield f"rmat{s}/seed{sd}", (lambda s=s, sd=sd: Graph.rmat(s, 16, seed=sd, weights=(1, 100)))
        for n in sides:
            yield f"grid{n}/seed{sd}", (lambda n=n, sd=sd: Graph.grid(n, 0.1, seed=sd, weights=(1, 1000)))


def probe(trace, seeds, scales, sides, algs=(0, 1, 2), log=print):
    for name, make in workloads(seeds, scales, sides):
        g = make()
        r = g.info()["max_outdeg_vertex"]
        for alg in algs:
            t = time.time()
            scheduler.collect_probe_traces(g, alg, [r], trace)
            log(f"probe {name} alg={alg}: {time.time() - t:.1f}s")
        g.close()


def evaluate(weights, trace):
    recs = scheduler.read_probe_records(trace)
    X, Y, T = scheduler.dataset(recs)
    P = scheduler.load_weights(weights)
    learned = np.argmax(scheduler.mlp_forward(P, X), axis=1)
    heur = [scheduler.heuristic_choice(r) for r in recs]
    gpu_rule = [0] * len(recs)
    return {"records": len(recs), "accuracy": float(np.mean(learned == Y)),
            "us_learned": scheduler.chosen_time(T, learned), "us_heuristic": scheduler.chosen_time(T, heur),
            "us_gpu_push_default": scheduler.chosen_time(T, gpu_rule), "us_best": scheduler.chosen_time(T, Y)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--trace", default="sched_train.jsonl")
    ap.add_argument("--held", default="sched_held.jsonl")
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=1500)
    args = ap.parse_args()
    scales = (14, 16) if args.quick else (16, 18, 20)
    sides = (128,) if args.quick else (256, 512, 1024)
    for p in (args.trace, args.held):
        if os.path.exists(p):
            os.remove(p)
    probe(args.trace, (1, 2), scales, sides)
    probe(args.held, (3,), scales, sides)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    loss = scheduler.train(args.trace, args.out, epochs=args.epochs, lr=3e-3, seed=1)
    res = {"train_loss": loss, "train": evaluate(args.out, args.trace), "held_out": evaluate(args.out, args.held)}
    print(json.dumps(res, indent=1))
    with open(os.path.splitext(args.out)[0] + "_eval.json", "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
