"""Benchmark: SSSP GTEPS on R-MAT scale 24 / edge factor 16 on one MI355X.

A step is one SSSP query (root = max out-degree vertex, uint32 weights
U[1,100]) over the whole graph, resident in HBM, through libcgmi's C ABI.
GTEPS = sum of out-degrees of the reached vertices / seconds per query.
The line also carries: the roofline of the dominant kernel (push expansion)
with its practical ceiling (the measured rate of relaxations at data-dependent
addresses, cggraph_amd/gather_ceiling), the restated reference CPU engine
timed on this host (cpu_baseline), and the
other headline workloads (BFS scale 18, WCC scale 24 symmetrized, PageRank
scale 25 / edge factor 44) in "extra".

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-extra] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GTEPS (BFS/SSSP, bit-exact) learned+batched vs plain, 1 MI355X; PR HBM GB/s"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gather_ceiling():
    """Rates of 4-byte gathers and relaxations at data-dependent addresses on
    this GPU (cggraph_amd/gather_ceiling, a child process): the practical
    ceiling of the push expansion and the PageRank pull."""
    exe = os.path.join(ROOT, "cggraph_amd", "gather_ceiling")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built; run `make -C cggraph_amd`")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"gather_ceiling failed ({r.returncode}): {r.stderr.strip()}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def host_cores():
    """The host cores this process may use: its CPU affinity, capped by a
    cgroup CPU quota when one is set, and the socket count."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(period)))
    except (OSError, ValueError):
        pass
    sockets = set()
    base = "/sys/devices/system/cpu"
    for c in (os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else range(nproc)):
        try:
            with open(f"{base}/cpu{c}/topology/physical_package_id") as f:
                sockets.add(f.read().strip())
        except OSError:
            pass
    usable = min(nproc, quota) if quota else nproc
    return {"nproc": nproc, "os_cpu_count": os.cpu_count(), "cgroup_cpu_quota": quota,
            "sockets": len(sockets) or 1, "usable": usable}


def trimmed_mean(times):
    """The reference's run average: sorted, min and max dropped when there
    are more than two runs (main_partition.hpp:209-227)."""
    t = sorted(times)
    return sum(t[1:-1]) / (len(t) - 2) if len(t) > 2 else sum(t) / len(t)


_HIP = None


def device_sync():
    """hipDeviceSynchronize: a query may return while iterations that find
    nothing to do are still queued on its engine's stream."""
    global _HIP
    if _HIP is None:
        import ctypes
        _HIP = ctypes.CDLL("libamdhip64.so")
    if _HIP.hipDeviceSynchronize() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


def timed_runs(eng, root, steps, warmup):
    for _ in range(warmup):
        eng.run(root, copy_out=False)
    device_sync()
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        _, st = eng.run(root, copy_out=False)
    device_sync()
    dt = (time.perf_counter() - t0) / steps
    return dt, st


def profile_run(cg, g, alg, root, **opt):
    eng = cg.Engine(g, alg, profile=1, **opt)
    eng.run(root, copy_out=False)
    _, st = eng.run(root, copy_out=False)
    eng.close()
    return st


def profile_workload(args):
    """For rocprofv3 counter runs: one 1 GiB device-to-device copy (the
    calibration kernel of known size: 1 GiB read + 1 GiB written) and one
    SSSP query with the benchmark's settings."""
    import ctypes
    import cggraph_amd as cg
    from cggraph_amd import autotune
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), n) or hip.hipMalloc(ctypes.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemcpy(b, a, n, 3)
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    g = cg.Graph.rmat(args.scale, args.edge_factor, seed=1, weights=(1, 100))
    info = g.info()
    tuned = autotune.best_options(autotune.engine_signature("rmat", 1, info["V"], info["E"]))
    e = cg.Engine(g, 1, sched_mode=1, **tuned)
    _, st = e.run(info["max_outdeg_vertex"], copy_out=False)
    e.close()
    print(json.dumps({"profile_workload": "1 GiB D2D copy + one SSSP query", "expand_launches": st["expand_launches"],
                      "expand_alg_bytes": st["expand_alg_bytes"], "iterations": st["iterations"]}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--profile-run", action="store_true",
                    help="run the calibration copy and one query only (for rocprofv3 counter runs)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (default: every host core this process may use)")
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="CPU baseline runs; min and max are dropped from the mean when more than two")
    ap.add_argument("--profile-batch", action="store_true",
                    help="run the batched mixed workload once (for rocprofv3 kernel traces)")
    ap.add_argument("--profile-pr", action="store_true",
                    help="run PageRank on R-MAT scale 25 / ef44 once (for rocprofv3 kernel traces)")
    ap.add_argument("--profile-bfs", action="store_true",
                    help="run 3 + 5 tuned BFS queries on R-MAT scale 18 / ef16 (for rocprofv3 kernel traces)")
    ap.add_argument("--profile-wcc", action="store_true",
                    help="run WCC on symmetrized R-MAT scale 24 / ef16, 2 + 5 times (for rocprofv3 kernel traces)")
    ap.add_argument("--profile-cc", action="store_true",
                    help="run directed CC (label propagation) on R-MAT scale 24 / ef16, 2 + 5 times "
                         "(for rocprofv3 kernel traces)")
    args = ap.parse_args()
    if args.profile_run:
        profile_workload(args)
        return
    if args.profile_bfs or args.profile_wcc or args.profile_cc:
        import cggraph_amd as cg
        from cggraph_amd import autotune
        if args.profile_bfs:
            g = cg.Graph.rmat(18, 16, seed=1, weights=None)
            info = g.info()
            opt = autotune.best_options(autotune.engine_signature("rmat", 0, info["V"], info["E"]))
            e = cg.Engine(g, 0, sched_mode=1, **opt)
            dt, st = timed_runs(e, info["max_outdeg_vertex"], 5, 3)
        elif args.profile_wcc:
            g = cg.Graph.rmat(24, 16, seed=1, weights=None, symmetrize=True)
            e = cg.Engine(g, 2, sched_mode=1)
            dt, st = timed_runs(e, 0, 5, 2)
        else:
            g = cg.Graph.rmat(24, 16, seed=1, weights=None)
            e = cg.Engine(g, 2, sched_mode=1)
            dt, st = timed_runs(e, 0, 5, 2)
        e.close()
        name = "bfs_s18" if args.profile_bfs else ("wcc_s24" if args.profile_wcc else "cc_s24_directed")
        print(json.dumps({"profile": name, "ms": round(dt * 1e3, 4), "iterations": st["iterations"]}), flush=True)
        return
    if args.profile_batch or args.profile_pr:
        import cggraph_amd as cg
        from cggraph_amd import autotune, batch
        if args.profile_batch:
            graphs, qg, qa, qr = batch.mixed_workload()
            opts = autotune.best_options("batch/mixed") or {"sssp_delta": 10000}
            _, st = batch.run_batch(qg, qa, qr, copy_out=False, **opts)
            print(json.dumps({"profile_batch": st["ms"], "levels": st["iterations"]}), flush=True)
        if args.profile_pr:
            g = cg.Graph.rmat(25, 44, seed=1, weights=None)
            e = cg.Engine(g, 3)
            _, st = e.run(0, copy_out=False)
            print(json.dumps({"profile_pr_ms": st["ms"]}), flush=True)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X; runs on more than one GPU are not implemented "
              "(sharding across GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import cggraph_amd as cg

    ceiling = gather_ceiling()
    log(f"gather ceiling: {ceiling}")

    t = time.perf_counter()
    g = cg.Graph.rmat(args.scale, args.edge_factor, seed=1, weights=(1, 100))
    info = g.info()
    root = info["max_outdeg_vertex"]
    log(f"R-MAT s{args.scale} ef{args.edge_factor}: V={info['V']} E={info['E']} root={root} "
        f"built in {time.perf_counter() - t:.2f}s")

    from cggraph_amd import autotune
    sig = autotune.engine_signature("rmat", 1, info["V"], info["E"])
    tuned = autotune.best_options(sig)
    log(f"tuned options for {sig}: {tuned or 'none (defaults)'}")
    eng = cg.Engine(g, 1, sched_mode=1, **tuned)
    dt, st = timed_runs(eng, root, args.steps, args.warmup)
    gpu_vals = eng.result()
    eng.close()
    m_q = st["edges_traversed"]
    gteps = m_q / dt / 1e9
    log(f"SSSP: {dt * 1e3:.3f} ms/query, iterations {st['iterations']}, reached edges {m_q}, "
        f"{gteps:.2f} GTEPS")

    learned = run_learned(cg, g, root, tuned, gpu_vals)

    pst = profile_run(cg, g, 1, root, sched_mode=1, **tuned)
    kms = pst["expand_kernel_ms"] / max(pst["expand_launches"], 1)
    alg_bytes = pst["expand_alg_bytes"] / max(pst["expand_launches"], 1)
    achieved = alg_bytes / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    # HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    # (scaled by the same counters on a 1 GiB copy), kept in profiles/
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_expand_sssp_rmat24.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")
    roofline = {"bound": "hbm", "kernel": "k_expand<SSSP>", "achieved": round(achieved, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic, "avg_launch_ms": round(kms, 5),
                "alg_bytes_per_launch": round(alg_bytes, 1), "launches": pst["expand_launches"]}
    # every relaxed edge is one 4-byte gather plus a conditional atomicMin at a
    # data-dependent address: the measured rate of that pattern is the ceiling
    relax_rate = pst["edges_relaxed"] / (pst["expand_kernel_ms"] * 1e-3) / 1e9 if pst["expand_kernel_ms"] else 0.0
    relax_ceiling = ceiling["relax_skewed_v24_G_per_s"]
    roofline["practical_ceiling"] = {
        "pattern": "relaxation (read, strict-less check, atomicMin) at degree-skewed addresses over 2^24 "
                   "values, measured by cggraph_amd/gather_ceiling",
        "unit": "G edges/s", "achieved": round(relax_rate, 2), "ceiling": relax_ceiling,
        "frac": round(relax_rate / relax_ceiling, 4), "edges_relaxed": pst["edges_relaxed"]}
    log(f"practical ceiling: {roofline['practical_ceiling']}")

    cpu = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle as orc
        off, dst, w = g.export()
        og = orc.CSR(off, dst, w)
        hc = host_cores()
        threads = args.cpu_threads or hc["usable"]
        times, parity = [], True
        for _ in range(max(1, args.cpu_runs)):
            vals, ms, iters = orc.multicore_run(1, og, root, threads)
            times.append(ms)
            parity = parity and bool(np.array_equal(vals, gpu_vals))
        ms = trimmed_mean(times)
        cpu_m = orc.reached_edges(og, vals)
        cpu = {"value": round(cpu_m / (ms * 1e-3) / 1e9, 4), "unit": "GTEPS", "cores": threads,
               "kind": "port",
               "sample": f"{len(times)} SSSP queries from the bench root on the same R-MAT s{args.scale} graph "
                         f"(restated reference MULTI_CORE engine: Graph_Numa + threeStage_taskSteal, one socket), "
                         f"{iters} iterations each; value from the mean of the runs without the fastest and "
                         f"slowest", "ms_per_query": round(ms, 2), "run_ms": [round(t, 2) for t in times],
               "omp_max_threads": orc.omp_max_threads(), "nproc": hc["nproc"],
               "os_cpu_count": hc["os_cpu_count"], "cgroup_cpu_quota": hc["cgroup_cpu_quota"],
               "sockets": hc["sockets"], "bit_exact_vs_gpu": parity}
        log(f"CPU baseline: {cpu}")

    extra = {"learned_sssp_s24": learned, "gather_ceiling": ceiling}
    if not args.no_extra:
        extra.update(run_extra(cg, np, info, ceiling))
    g.close()

    line = {
        "metric": METRIC, "value": round(gteps, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"SSSP on R-MAT scale {args.scale} edge factor {args.edge_factor} "
                               f"(a/b/c/d .57/.19/.19/.05, seed 1, self-loops and duplicates removed, "
                               f"weights U[1,100]), root = max out-degree vertex",
                   "V": info["V"], "E": info["E"], "root": root, "schedule": "gpu",
                   "tuned_options": tuned, "iterations": st["iterations"], "reached_edges": m_q},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


def run_learned(cg, g, root, tuned, ref_vals):
    """SSSP with the learned per-iteration scheduler (weights trained on the
    GPU from the engine's own probe traces, cggraph_amd/data/sched_mlp.bin)."""
    import numpy as np
    wpath = os.path.join(ROOT, "cggraph_amd", "data", "sched_mlp.bin")
    if not os.path.exists(wpath):
        return None
    eng = cg.Engine(g, 1, sched_mode=2, mlp_path=wpath, **tuned)
    dt, st = timed_runs(eng, root, 10, 3)
    same = bool(np.array_equal(eng.result(), ref_vals))
    eng.close()
    res = {"ms": round(dt * 1e3, 4), "gteps": round(st["edges_traversed"] / dt / 1e9, 3),
           "iterations": st["iterations"], "learned_fallbacks": st["learned_fallbacks"],
           "light_iterations": st["light_iterations"], "bit_exact_vs_plain": same}
    # the reference's own LIGHT / HEAVY rule (host team for light iterations
    # and the small-degree tail), the plain heuristic the network replaces
    eng = cg.Engine(g, 1, sched_mode=0)
    hdt, hst = timed_runs(eng, root, 3, 1)
    hsame = bool(np.array_equal(eng.result(), ref_vals))
    eng.close()
    res["heuristic"] = {"ms": round(hdt * 1e3, 4), "gteps": round(hst["edges_traversed"] / hdt / 1e9, 3),
                        "iterations": hst["iterations"], "light_iterations": hst["light_iterations"],
                        "bit_exact_vs_plain": hsame}
    res["speedup_vs_heuristic"] = round(hdt / dt, 2)
    log(f"learned SSSP s24: {res}")
    return res


def run_batched(cg, np):
    """256 BFS + 256 SSSP queries over R-MAT s14-22 and lattices 256^2-2048^2:
    the batched run (the R-MAT s18-22 queries one by one on engines of their
    graphs from a host thread, the rest level by level in one HIP-graph-
    captured loop, both at once) vs one query per engine launch."""
    from cggraph_amd import autotune, batch
    graphs, qg, qa, qr = batch.mixed_workload()
    opts = autotune.best_options("batch/mixed") or {"sssp_delta": 10000}
    batch.run_batch(qg[:32], qa[:32], qr[:32], copy_out=False, **opts)  # warm-up
    outs_b, st_b = batch.run_batch(qg, qa, qr, copy_out=True, **opts)
    t = time.perf_counter()
    outs_1, st_1 = batch.run_batch(qg, qa, qr, mode=batch.ONE_PER_LAUNCH, copy_out=True,
                                   sssp_delta=opts.get("sssp_delta", 0))
    wall_1 = time.perf_counter() - t
    same = all(np.array_equal(a, b) for a, b in zip(outs_b, outs_1))
    del outs_b, outs_1
    res = {"queries": len(qr), "graphs": len(graphs), "options": opts,
           "batched_ms": round(st_b["ms"], 2), "batched_gteps": round(st_b["edges_traversed"] / st_b["ms"] / 1e6, 3),
           "batched_levels": st_b["iterations"], "graph_replays": st_b["expand_launches"],
           "queries_on_engines": st_b["heavy_iterations"],
           "one_per_launch_ms": round(st_1["ms"], 2),
           "one_per_launch_gteps": round(st_1["edges_traversed"] / st_1["ms"] / 1e6, 3),
           "one_per_launch_wall_s": round(wall_1, 2), "speedup": round(st_1["ms"] / st_b["ms"], 2),
           "reached_edges": st_b["edges_traversed"], "outputs_identical": same}
    log(f"batched: {res}")
    for gr in graphs:
        gr.close()
    return res


def run_extra(cg, np, info24, ceiling):
    extra = {"batched_mixed": run_batched(cg, np)}
    # BFS R-MAT scale 18 (the reference's own CPU-runnable configuration)
    from cggraph_amd import autotune
    g = cg.Graph.rmat(18, 16, seed=1, weights=None)
    inf = g.info()
    eng = cg.Engine(g, 0, sched_mode=1, **autotune.best_options(autotune.engine_signature("rmat", 0, inf["V"], inf["E"])))
    dt, st = timed_runs(eng, inf["max_outdeg_vertex"], 20, 3)
    eng.close()
    extra["bfs_s18"] = {"gteps": round(st["edges_traversed"] / dt / 1e9, 3), "ms": round(dt * 1e3, 4),
                        "E": inf["E"], "iterations": st["iterations"], "pull_iterations": st["pull_iterations"]}
    log(f"BFS s18: {extra['bfs_s18']}")
    g.close()
    # BFS on scale 24 as well (tuned tiers for this graph kind, when the table has them)
    g = cg.Graph.rmat(24, 16, seed=1, weights=None)
    inf = g.info()
    eng = cg.Engine(g, 0, sched_mode=1, **autotune.best_options(autotune.engine_signature("rmat", 0, inf["V"], inf["E"])))
    dt, st = timed_runs(eng, inf["max_outdeg_vertex"], 5, 2)
    eng.close()
    extra["bfs_s24"] = {"gteps": round(st["edges_traversed"] / dt / 1e9, 3), "ms": round(dt * 1e3, 4),
                        "iterations": st["iterations"], "pull_iterations": st["pull_iterations"]}
    log(f"BFS s24: {extra['bfs_s24']}")
    g.close()
    # WCC on the symmetrized scale-24 graph
    g = cg.Graph.rmat(24, 16, seed=1, weights=None, symmetrize=True)
    inf = g.info()
    eng = cg.Engine(g, 2, sched_mode=1)
    dt, st = timed_runs(eng, 0, 5, 2)
    eng.close()
    extra["wcc_s24_sym"] = {"gteps": round(inf["E"] / dt / 1e9, 3), "ms": round(dt * 1e3, 4), "E": inf["E"]}
    log(f"WCC s24: {extra['wcc_s24_sym']}")
    g.close()
    # PageRank on R-MAT scale 25 edge factor 44
    t = time.perf_counter()
    g = cg.Graph.rmat(25, 44, seed=1, weights=None)
    inf = g.info()
    log(f"R-MAT s25 ef44: V={inf['V']} E={inf['E']} built in {time.perf_counter() - t:.2f}s")
    alg = 8.0 * inf["E"] + 8.0 * (inf["V"] + 1) + 4.0 * inf["V"]

    def pr_time(**opt):
        eng = cg.Engine(g, 3, profile=1, **opt)
        eng.run(0, copy_out=False)
        _, st = eng.run(0, copy_out=False)
        eng.close()
        return st["expand_kernel_ms"] / max(st["iterations"], 1), st

    layouts = {4: "propagation blocking per edge", 5: "propagation blocking per (source, bin) pair, renumbered"}
    # default (automatic: pair blocking at this size); the bytes its layout
    # streams per iteration by design come with the run's statistics
    per_it, st = pr_time()
    gbs = alg / (per_it * 1e-3) / 1e9
    streamed = st["pr_stream_bytes"]
    extra["pr_s25_ef44"] = {"layout": layouts.get(st["pr_layout"], str(st["pr_layout"])) + " (automatic)",
                            "ms_per_iteration": round(per_it, 4),
                            "alg_bytes_per_iteration": alg, "GBps": round(gbs, 1),
                            "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "E": inf["E"],
                            "iterations": st["iterations"], "streamed_bytes_per_iteration": streamed,
                            "streamed_GBps": round(streamed / (per_it * 1e-3) / 1e9, 1),
                            "frac_of_copy_rate": round(streamed / (per_it * 1e-3) / 1e9 / ceiling["copy_GBps"], 4)}
    # the per-edge blocking it replaces
    edge_it, st4 = pr_time(pr_layout=4)
    extra["pr_s25_ef44_per_edge"] = {"ms_per_iteration": round(edge_it, 4),
                                     "frac_of_peak": round(alg / (edge_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                     "streamed_bytes_per_iteration": st4["pr_stream_bytes"]}
    # the renumbered pull: one 4-byte gather per in-edge, against the measured
    # gather rate over 2^25 values
    pull_it, _ = pr_time(pr_layout=2)
    gathers = inf["E"] / (pull_it * 1e-3) / 1e9
    pr_ceiling = ceiling["gather_skewed_v25_G_per_s"]
    extra["pr_s25_ef44_pull"] = {"ms_per_iteration": round(pull_it, 4),
                                 "frac_of_peak": round(alg / (pull_it * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "gathers_G_per_s": round(gathers, 2), "gather_ceiling_G_per_s": pr_ceiling,
                                 "frac_of_gather_ceiling": round(gathers / pr_ceiling, 4)}
    log(f"PR s25 pull: {extra['pr_s25_ef44_pull']}")
    log(f"PR s25: {extra['pr_s25_ef44']}")
    g.close()
    return extra


if __name__ == "__main__":
    main()
