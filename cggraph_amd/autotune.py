"""Autotuner for the traversal engine's degree tiers, delta-stepping width,
hub deferral and renumbering, and the batched kernel's launch shape.

It times real runs through the C ABI (the outputs do not depend on any of the
tuned values, so any choice is correct) and keeps the fastest configuration
per workload signature in a small JSON table (cggraph_amd/data/tuning.json)
that the benchmark and callers read back with best_options().

Search: coordinate descent from the defaults -- each parameter in turn is
swept with the others fixed, two passes -- so a tuning run costs tens of
timed queries, not the full cross product.
"""
import json
import os
import time

from . import _lib
from .batch import run_batch
from .core import Engine

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "data", "tuning.json")

ENGINE_SPACE = {
    "bin_wave": [8, 16, 32, 64, 128],
    "bin_chunk": [512, 1024, 4096, 16384],
    "chunk_edges": [512, 1024, 2048, 4096, 8192],
    "sssp_delta": [0, 16, 32, 64, 128, 256, 1024, 4096],
    "hub_degree": [0, 64, 128, 256, 512],
    "hub_delta": [1, 2, 3, 5, 8],
    "relabel": [0, 1],
    "atomic_return": [1, 0],
    "expand_grid": [0, 4096, 8192, 16384],
}
BATCH_SPACE = {
    "sssp_delta": [1000, 3000, 10000, 30000, 100000],
    "hub_degree": [0, 32, 64, 128, 512],
    "hub_delta": [5, 20, 50, 200, 1000],
    "batch_grid": [256, 512, 1024, 2048, 4096],
    "batch_block": [256, 512, 1024],
    "batch_levels": [8, 32, 128],
    "batch_group": [16, 32, 64],
    "batch_engine_v": [1 << 17, 1 << 18, 1 << 19, 1 << 20, 0xFFFFFFFF],
}


def load_table(path=TABLE):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def save_table(table, path=TABLE):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)


def engine_signature(kind, algorithm, V, E):
    """Workloads are keyed by family, algorithm and size class (log2 V)."""
    return f"engine/{kind}/alg{int(algorithm)}/v{max(int(V) - 1, 1).bit_length()}/e{max(int(E) - 1, 1).bit_length()}"


def best_options(signature, path=TABLE):
    return dict(load_table(path).get(signature, {}).get("options", {}))


def _coordinate_descent(space, measure, start, passes=2, log=None):
    cur = dict(start)
    best_t = measure(cur)
    tried = {json.dumps(cur, sort_keys=True): best_t}
    for _ in range(passes):
        improved = False
        for key, values in space.items():
            for v in values:
                cand = dict(cur, **{key: v})
                k = json.dumps(cand, sort_keys=True)
                if k not in tried:
                    tried[k] = measure(cand)
                    if log:
                        log(f"  {cand} -> {tried[k]:.3f} ms")
                if tried[k] < best_t:
                    best_t, cur, improved = tried[k], cand, True
        if not improved:
            break
    return cur, best_t, len(tried)


def tune_engine(graph, algorithm, root, kind="rmat", reps=3, space=None, log=None, path=TABLE, save=True):
    space = dict(space or ENGINE_SPACE)
    if int(algorithm) != 1:
        for k in ("sssp_delta", "hub_degree", "hub_delta"):
            space.pop(k, None)
    o = _lib.default_options()
    start = {k: getattr(o, k) for k in space}

    def measure(cfg):
        if cfg.get("sssp_delta") and cfg.get("hub_degree"):
            return float("inf")  # alternatives, not combined
        e = Engine(graph, algorithm, sched_mode=_lib.CG_SCHED_GPU, **cfg)
        try:
            e.run(root, copy_out=False)
            return min(e.run(root, copy_out=False)[1]["ms"] for _ in range(reps))
        finally:
            e.close()

    t0 = time.time()
    best, ms, n = _coordinate_descent(space, measure, start, log=log)
    base = measure(start)
    inf = graph.info()
    sig = engine_signature(kind, algorithm, inf["V"], inf["E"])
    rec = {"options": best, "ms": round(ms, 4), "default_ms": round(base, 4), "configs_timed": n,
           "tuning_s": round(time.time() - t0, 1)}
    if save:
        table = load_table(path)
        table[sig] = rec
        save_table(table, path)
    return sig, rec


def tune_batch(qg, qa, qr, signature="batch/mixed", log=None, path=TABLE, save=True):
    o = _lib.default_options()
    start = {k: getattr(o, k) for k in BATCH_SPACE}
    start["sssp_delta"] = 10000

    def measure(cfg):
        _, st = run_batch(qg, qa, qr, copy_out=False, **cfg)
        return st["ms"]

    t0 = time.time()
    best, ms, n = _coordinate_descent(BATCH_SPACE, measure, start, passes=1, log=log)
    rec = {"options": best, "ms": round(ms, 3), "configs_timed": n, "tuning_s": round(time.time() - t0, 1)}
    if save:
        table = load_table(path)
        table[signature] = rec
        save_table(table, path)
    return signature, rec
