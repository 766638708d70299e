"""C-ABI conformance on the CPU: the library loads, exports exactly what
include/cgmi.h declares, and its host-side logic (graph validation, file
formats, partition, error codes) behaves; no GPU compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "cgmi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", txt)))


def test_exports_match_header():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert b"gfx950" in L.cg_version()


def test_options_defaults_match_reference_constants():
    o = _lib.default_options()
    assert o.segment_size == 1024          # Two-level_partition.cuh:7-8
    assert o.heavy_active_vertices == 16384  # CGgraph_engine.cuh:21
    assert o.heavy_active_edges == 25_000_000  # CGgraph_engine.cuh:22
    assert o.pr_iterations == 20 and abs(o.pr_damping - 0.85) < 1e-7
    sizes = [ctypes.c_size_t() for _ in range(3)]
    _lib.lib().cg_abi_sizes(*[ctypes.byref(s) for s in sizes])
    assert [s.value for s in sizes] == [ctypes.sizeof(_lib.Options), ctypes.sizeof(_lib.RunStats),
                                        ctypes.sizeof(_lib.PartitionInfo)]


def test_tuning_table_names_real_options():
    """Every option the tuning table stores (engine and batch entries) is a
    field of cg_options, so the tuned runs apply all of it; the autotuner's
    search spaces name only real fields too."""
    from cggraph_amd import autotune
    fields = {f for f, _ in _lib.Options._fields_}
    table = autotune.load_table()
    assert table, "tuning table missing"
    for sig, rec in table.items():
        assert set(rec["options"]) <= fields, (sig, set(rec["options"]) - fields)
        _lib.default_options(**rec["options"])  # every value fits its field
    assert set(autotune.ENGINE_SPACE) <= fields and set(autotune.BATCH_SPACE) <= fields


def test_graph_validation_and_export():
    off = np.array([0, 2, 3, 3], np.uint32)
    dst = np.array([1, 2, 0], np.uint32)
    w = np.array([5, 6, 7], np.uint32)
    g = cg.Graph.from_csr(off, dst, w)
    inf = g.info()
    assert inf == {"V": 3, "E": 3, "has_weights": True, "symmetric": False, "max_outdeg_vertex": 0}
    o2, d2, w2 = g.export()
    assert o2.tolist() == [0, 2, 3, 3] and d2.tolist() == [1, 2, 0] and w2.tolist() == [5, 6, 7]
    for bad_off, bad_dst in (([1, 2, 3, 3], dst), ([0, 2, 1, 3], dst), ([0, 2, 3, 3], [1, 9, 0])):
        with pytest.raises(cg.CgError) as ei:
            cg.Graph.from_csr(np.array(bad_off, np.uint32), np.array(bad_dst, np.uint32))
        assert ei.value.code == _lib.CG_ERR_INVALID
    with pytest.raises(cg.CgError):  # weight large enough to wrap value + weight
        cg.Graph.from_csr(off, dst, np.array([5, 2**32 - 1, 7], np.uint32))
    with pytest.raises(cg.CgError):  # tail declared zero-out-degree but has edges
        cg.Graph.from_csr(off, dst, None, zero_outdeg_tail=2)
    sym = cg.Graph.from_csr(np.array([0, 1, 2], np.uint64), np.array([1, 0], np.uint32))
    assert sym.info()["symmetric"]


def test_load_files_and_result_files(tmp_path):
    off = np.array([0, 2, 3, 3], np.uint32)
    dst = np.array([1, 2, 0], np.uint32)
    w = np.array([5, 6, 7], np.uint32)
    off.tofile(tmp_path / "o.bin")
    dst.tofile(tmp_path / "d.bin")
    w.tofile(tmp_path / "w.bin")
    g = cg.Graph.load_files(str(tmp_path / "o.bin"), str(tmp_path / "d.bin"), str(tmp_path / "w.bin"), 3, 3)
    assert g.export()[1].tolist() == [1, 2, 0]
    with pytest.raises(cg.CgError) as ei:  # wrong size, as load_binFile asserts
        cg.Graph.load_files(str(tmp_path / "o.bin"), str(tmp_path / "d.bin"), None, 3, 4)
    assert ei.value.code == _lib.CG_ERR_IO
    vals = np.array([0, 1, 999_999_999], np.uint32)
    p = str(tmp_path / "r.bin").encode()
    _lib.check(_lib.lib().cg_save_bin_u32(p, vals.ctypes.data_as(ctypes.c_void_p), 3))
    bad = ctypes.c_uint64(0)
    assert _lib.lib().cg_check_result_file(p, vals.ctypes.data_as(ctypes.c_void_p), 3, None,
                                           ctypes.byref(bad)) == 0
    other = np.array([0, 2, 999_999_999], np.uint32)
    assert _lib.lib().cg_check_result_file(p, other.ctypes.data_as(ctypes.c_void_p), 3, None,
                                           ctypes.byref(bad)) == _lib.CG_ERR_MISMATCH
    assert bad.value == 1
    # through old2new: the engine's vertex old2new[i] is the file's vertex i
    o2n = np.array([2, 0, 1], np.uint32)
    mine = np.array([1, 999_999_999, 0], np.uint32)
    assert _lib.lib().cg_check_result_file(p, mine.ctypes.data_as(ctypes.c_void_p), 3,
                                           o2n.ctypes.data_as(ctypes.c_void_p), ctypes.byref(bad)) == 0


def test_partition_buckets(oracle):
    og = oracle.rmat_graph(14, 16, seed=1)
    g = cg.Graph.from_csr(og.offsets, og.dest, og.weight)
    p = g.partition(1)
    nseg = (og.V + 1023) // 1024
    assert p["segment_num"] == nseg
    deg = np.array([(int(og.offsets[min(og.V, (s + 1) * 1024)]) - int(og.offsets[s * 1024])) /
                    (min(og.V, (s + 1) * 1024) - s * 1024) for s in range(nseg)])
    b = p["bucket_start"]
    assert b[0] == 0 and b[4] == nseg
    assert b[1] == int(np.sum(deg < 17)) and b[2] == int(np.sum(deg < 33)) and b[3] == int(np.sum(deg < 257))


def test_engine_needs_a_gpu_here():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    g = cg.Graph.from_csr(np.array([0, 1, 1], np.uint32), np.array([1], np.uint32))
    with pytest.raises(cg.CgError) as ei:
        cg.Engine(g, 0)
    assert ei.value.code == _lib.CG_ERR_NO_DEVICE
    # use_device_num = 0 asks for the engine without a GPU: it runs on the host team
    e = cg.Engine(g, 0, use_device_num=0)
    vals, st = e.run(0)
    assert vals.tolist() == [0, 1] and st["heavy_iterations"] == 0
    e.close()


def test_noalloc_rejects_before_touching_a_device():
    """The allocation-free run checks its arguments and the scratch size
    before any HIP call (fake device pointers are never dereferenced)."""
    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    it = ctypes.c_uint32(7)
    # 2^24 vertices need more than the 64 MiB scratch
    assert L.cg_traverse_device_noalloc(1, fake, fake, fake, 1 << 24, 0, 0, fake, None, ctypes.byref(it)) == \
        _lib.CG_ERR_INVALID
    assert b"scratch" in L.cg_last_error()
    assert it.value == 0
    # an edge count whose chunk queue does not fit
    assert L.cg_traverse_device_noalloc(0, fake, fake, None, 1000, 1 << 36, 0, fake, None, None) == \
        _lib.CG_ERR_INVALID
    assert L.cg_traverse_device_noalloc(3, fake, fake, fake, 10, 10, 0, fake, None, None) == _lib.CG_ERR_UNSUPPORTED
    assert L.cg_traverse_device_noalloc(0, fake, fake, None, 10, 10, 10, fake, None, None) == _lib.CG_ERR_INVALID
    assert L.cg_traverse_device_noalloc(1, fake, fake, None, 10, 10, 0, fake, None, None) == _lib.CG_ERR_INVALID
    assert L.cg_traverse_device_noalloc(0, None, fake, None, 10, 10, 0, fake, None, None) == _lib.CG_ERR_INVALID
    assert L.cg_traverse_device_noalloc(0, fake, fake, None, 0, 0, 0, fake, None, None) == _lib.CG_ERR_INVALID
    assert _lib.CG_NOALLOC_SCRATCH_BYTES == 64 << 20
