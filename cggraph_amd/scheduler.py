"""Learned per-iteration scheduler: trace collection, training and inference
helpers over libcgmi (the network and its training run in HIP kernels).

Workflow:
  1. collect_probe_traces(): engines in probe mode time all 16 schedule
     choices from the same state at every iteration and append one JSON line
     per iteration (features + 16 times) to a trace file;
  2. train(): cg_mlp_train fits the 12-64-64-16 network on the GPU (label =
     fastest choice) and writes a weight file;
  3. Engine(..., sched_mode=CG_SCHED_LEARNED, mlp_path=weights) runs with the
     network's choice each iteration (inference fused into the statistics
     kernel), falling back to the fixed rule when the network is missing,
     non-finite or not confident.
"""
import ctypes
import json

import numpy as np

from . import _lib
from ._lib import check, lib
from .core import Engine

CHOICE_NAMES = [f"{'pull' if c & 1 else 'push'}/tier{(c >> 1) & 3}/{'host' if c & 8 else 'gpu'}"
                for c in range(16)]


def decode_choice(c):
    return {"pull": c & 1, "tier": (c >> 1) & 3, "host": (c >> 3) & 1}


def collect_probe_traces(graph, algorithm, roots, trace_path, cpu_threads=0):
    """Runs one probe-mode query per root; returns the per-run stats."""
    out = []
    for r in roots:
        e = Engine(graph, algorithm, sched_mode=_lib.CG_SCHED_PROBE, trace_path=trace_path,
                   cpu_threads=cpu_threads)
        try:
            _, st = e.run(int(r), copy_out=False)
        finally:
            e.close()
        out.append(st)
    return out


def read_probe_records(trace_path):
    recs = []
    with open(trace_path) as f:
        for line in f:
            d = json.loads(line)
            if d.get("probe"):
                recs.append(d)
    return recs


def _run_key(r):
    return (r["alg"], r["V"], r["E"], r["root"])


def dataset(records):
    """Training / evaluation pairs of the lagged scheduler: the network sees
    the features of iteration k's frontier and picks iteration k + 1's
    schedule.  X = features of record k, T = probe times of record k + 1 of the
    same run (same algorithm, graph, root; next iteration), Y = the fastest
    choice in T, R = the records k + 1 (for rules that look at the iteration
    itself)."""
    X, T, R = [], [], []
    for a, b in zip(records, records[1:]):
        if _run_key(a) == _run_key(b) and b["iter"] == a["iter"] + 1:
            X.append(a["features"])
            T.append(b["us_all"])
            R.append(b)
    X = np.array(X, np.float32).reshape(-1, 12)
    T = np.array(T, np.float64).reshape(-1, 16)
    Y = labels(T)
    return X, Y, T, R


LABEL_TOLERANCE = 0.03  # cg_mlp_train's kLabelTolerance


def labels(T):
    """The lowest-numbered choice within LABEL_TOLERANCE of the fastest
    (choices that close are equal within the timing noise)."""
    Tp = np.where(T > 0, T, np.inf)
    lim = Tp.min(axis=1, keepdims=True) * (1.0 + LABEL_TOLERANCE)
    return np.argmax(Tp <= lim, axis=1).astype(np.int32)


def train(trace_path, out_path, epochs=400, lr=3e-3, seed=1):
    loss = ctypes.c_float(0)
    check(lib().cg_mlp_train(trace_path.encode(), out_path.encode(), int(epochs), float(lr), int(seed),
                             ctypes.byref(loss)))
    return float(loss.value)


def load_weights(path):
    with open(path, "rb") as f:
        magic = f.read(8)
        if magic != b"CGMLP001":
            raise ValueError(f"{path} is not a scheduler weight file")
        P = np.frombuffer(f.read(), dtype=np.float32)
    if P.size != _lib.MLP_PARAMS:
        raise ValueError("wrong parameter count")
    return P.copy()


def mlp_forward(P, X):
    P = np.ascontiguousarray(P, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    n = X.shape[0]
    probs = np.empty((n, 16), np.float32)
    check(lib().cg_mlp_forward(P.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p), n,
                               probs.ctypes.data_as(ctypes.c_void_p)))
    return probs


def mlp_loss_grad(P, X, Y):
    P = np.ascontiguousarray(P, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.int32)
    g = np.empty(_lib.MLP_PARAMS, np.float32)
    loss = ctypes.c_float(0)
    check(lib().cg_mlp_loss_grad(P.ctypes.data_as(ctypes.c_void_p), X.ctypes.data_as(ctypes.c_void_p),
                                 Y.ctypes.data_as(ctypes.c_void_p), len(Y), ctypes.byref(loss),
                                 g.ctypes.data_as(ctypes.c_void_p)))
    return float(loss.value), g


def heuristic_choice(record, heavy_vertices=16384, heavy_edges=25_000_000):
    """The reference LIGHT/HEAVY rule expressed in the 16-choice space
    (LIGHT -> host team, HEAVY -> GPU push with the default tiers)."""
    n, m = record.get("n_active"), record.get("m_active")
    if n is None:
        return 0
    heavy = n != 1 and n >= heavy_vertices and m >= heavy_edges
    return 0 if heavy else 8


def chosen_time(T, choices):
    """Total probe-measured time of a sequence of choices."""
    return float(sum(T[i, c] for i, c in enumerate(choices)))
