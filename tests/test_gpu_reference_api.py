"""The reference's driver-level names over the GPU engine (main_partition,
TwoLevelPartition, CGgraphEngine.CG_co_execution, checkBinResult) and the
kernel-level seam cg_push_step_device (bfs_device / sssp_device)."""
import ctypes
import os

import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib
from cggraph_amd.reference_api import (Algorithm_type, CGgraphEngine, Engine_type, OrderMethod, TwoLevelPartition,
                                       checkBinResult, getToyGraph, main_partition, save_binArray)

pytestmark = pytest.mark.gpu


def test_engine_on_toys_matches_golden(golden):
    for name in ("toy_0", "toy_1"):
        t = golden["graphs"][name]
        csr = getToyGraph(name)
        for alg, key in ((Algorithm_type.BFS, "bfs"), (Algorithm_type.SSSP, "sssp")):
            part = TwoLevelPartition(csr, 0, 1, alg).getPartitionResult()
            assert part["segment_num"] == 1
            eng = CGgraphEngine(csr, part, alg, 1)
            for r, exp in t[key].items():
                ms = eng.CG_co_execution(int(r))
                assert ms >= 0
                assert eng.vertexValue.tolist() == exp


def test_main_partition_trimmed_mean_and_result_check(golden, tmp_path):
    res = main_partition("toy_1", OrderMethod.NATIVE, root=3, runs=5, algorithm=Algorithm_type.SSSP)
    assert len(res["times_ms"]) == 5
    ts = sorted(res["times_ms"])
    assert abs(res["total_run_time_ms"] - sum(ts[1:-1]) / 3) < 1e-9
    assert res["vertexValue"].tolist() == golden["graphs"]["toy_1"]["sssp"]["3"]
    # logResult path: <graph>_<ALG>_<root>.bin compared exactly (util.cuh:24-118)
    os.environ["CG_RESULT_DIR"] = str(tmp_path)
    save_binArray(str(tmp_path / "toy_1_SSSP_3.bin"), np.array(golden["graphs"]["toy_1"]["sssp"]["3"], np.uint32))
    main_partition("toy_1", OrderMethod.NATIVE, root=3, runs=1, algorithm=Algorithm_type.SSSP, logResult=True)
    bad = np.array(golden["graphs"]["toy_1"]["sssp"]["3"], np.uint32)
    bad[5] += 1
    with pytest.raises(cg.CgError):
        checkBinResult(str(tmp_path / "toy_1_SSSP_3.bin"), bad, 12)
    # MULTI_CORE runs on the host team (no GPU), bit-exact; BLOCK_2D is refused
    res = main_partition("toy_1", OrderMethod.NATIVE, root=3, runs=3, algorithm=Algorithm_type.SSSP,
                         engine_type=Engine_type.MULTI_CORE)
    assert res["vertexValue"].tolist() == golden["graphs"]["toy_1"]["sssp"]["3"]
    with pytest.raises(cg.CgError):
        main_partition("toy_0", engine_type=Engine_type.BLOCK_2D)


def _hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipDeviceSynchronize.argtypes = []
    return h


def test_push_step_device_seam(oracle):
    """Level-synchronous BFS driven step by step through cg_push_step_device
    on caller-owned device buffers equals the oracle."""
    og = oracle.rmat_graph(12, 8, seed=6)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    V = og.V
    words = (V + 63) // 64
    hip = _hip()
    H2D, D2H = 1, 2
    ptrs = [ctypes.c_void_p() for _ in range(3)]
    assert hip.hipMalloc(ctypes.byref(ptrs[0]), V * 4) == 0
    assert hip.hipMalloc(ctypes.byref(ptrs[1]), words * 8) == 0
    assert hip.hipMalloc(ctypes.byref(ptrs[2]), words * 8) == 0
    try:
        root = oracle.max_outdeg_vertex(og)
        for alg, ref in ((0, oracle.bfs(og, root)), (1, oracle.sssp(og, root))):
            val = np.full(V, 999_999_999, np.uint32)
            val[root] = 0
            fin = np.zeros(words, np.uint64)
            fin[root >> 6] = np.uint64(1) << np.uint64(root & 63)
            hip.hipMemcpy(ptrs[0], val.ctypes.data, V * 4, H2D)
            hip.hipMemcpy(ptrs[1], fin.ctypes.data, words * 8, H2D)
            cur, nxt = ptrs[1], ptrs[2]
            for _ in range(10 * V):
                zero = np.zeros(words, np.uint64)
                hip.hipMemcpy(nxt, zero.ctypes.data, words * 8, H2D)
                _lib.check(_lib.lib().cg_push_step_device(alg, g._h, ptrs[0], cur, nxt, None))
                hip.hipDeviceSynchronize()
                out = np.empty(words, np.uint64)
                hip.hipMemcpy(out.ctypes.data, nxt, words * 8, D2H)
                if not out.any():
                    break
                cur, nxt = nxt, cur
            hip.hipMemcpy(val.ctypes.data, ptrs[0], V * 4, D2H)
            assert np.array_equal(val, ref), alg
    finally:
        for p in ptrs:
            hip.hipFree(p)
