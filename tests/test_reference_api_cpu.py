"""Host-side parts of the reference's driver surface (no GPU): the graph
registry, raw array files, neighbour sort, enums."""
import numpy as np
import pytest

import cggraph_amd as cg
from cggraph_amd import reference_api as ra


def test_enums_match_reference_values():
    # Basic/Graph/basic_struct.hpp:144-151, 165-178, 41-...
    assert [int(a) for a in ra.Algorithm_type] == [0, 1, 2, 3, 4]
    assert ra.Engine_type.COORDINATION == 0 and ra.Engine_type.MULTI_CORE == 3 and ra.Engine_type.SINGLE_GPU == 4
    assert ra.OrderMethod.NATIVE == 0 and ra.OrderMethod.RCM_LAST_ZERO_OUTDEGREE == 6
    assert ra.VertexValue_MAX == 999_999_999 and ra.SEGMENT_SIZE == 1024


def test_registry_sizes_and_paths(monkeypatch, tmp_path):
    monkeypatch.setenv("CG_GRAPH_DIR", str(tmp_path))
    gf = ra.getGraphFile("twitter2010")
    assert (gf.vertices, gf.edges) == (61578415, 1468364884)  # graphFileList.hpp:219-220
    assert gf.csrOffsetFile.startswith(str(tmp_path)) and gf.old2newFile == ""
    gf = ra.getGraphFile("uk-union", ra.OrderMethod.RCM_LAST_ZERO_OUTDEGREE)
    assert gf.edges == 5475109924 and gf.common_root == 3 and gf.addtitionFile.endswith("_addition.bin")
    with pytest.raises(cg.CgError):
        ra.getGraphFile("no-such-graph")


def test_load_bin_file_size_check(tmp_path):
    a = np.arange(10, dtype=np.uint32)
    p = str(tmp_path / "a.bin")
    ra.save_binArray(p, a)
    assert np.array_equal(ra.load_binFile(p, 10), a)
    with pytest.raises(cg.CgError):
        ra.load_binFile(p, 11)


def test_toy_graph_and_neighbour_sort():
    csr = ra.getToyGraph("toy_0")
    assert (csr.vertexNum, csr.edgeNum) == (12, 26)
    shuffled = ra.CSR_Result_type(vertexNum=3, edgeNum=4, csr_offset=np.array([0, 3, 4, 4], np.uint32),
                                  csr_dest=np.array([2, 0, 1, 0], np.uint32),
                                  csr_weight=np.array([7, 8, 9, 5], np.uint32))
    ra.nbrSort_taskSteal(shuffled)
    assert shuffled.csr_dest.tolist() == [0, 1, 2, 0] and shuffled.csr_weight.tolist() == [8, 9, 7, 5]
    dup = ra.CSR_Result_type(vertexNum=1, edgeNum=2, csr_offset=np.array([0, 2], np.uint32),
                             csr_dest=np.array([0, 0], np.uint32), csr_weight=None)
    with pytest.raises(cg.CgError):  # strictly ascending, as sortNbr.hpp:84-91 asserts
        ra.nbrSort_taskSteal(dup)


@pytest.mark.parametrize("engine_type", [ra.Engine_type.MULTI_CORE, ra.Engine_type.SINGLE_CORE,
                                         ra.Engine_type.COORDINATION])
def test_main_partition_without_gpu(golden, engine_type):
    """main_partition with a CPU engine type (main_partition.hpp:248-317), or
    COORDINATION with useDeviceNum = 0: the host team runs every iteration;
    runs = 4 gives the trimmed mean of the two middle times."""
    t = golden["graphs"]["toy_0"]
    for alg, key in ((ra.Algorithm_type.BFS, "bfs"), (ra.Algorithm_type.SSSP, "sssp")):
        for r, exp in t[key].items():
            res = ra.main_partition("toy_0", root=int(r), runs=4, algorithm=alg, engine_type=engine_type,
                                    useDeviceNum=0, cpu_threads=3)
            assert res["vertexValue"].tolist() == exp
            assert len(res["times_ms"]) == 4
            assert res["stats"]["light_iterations"] == res["stats"]["iterations"]
    res = ra.main_partition("toy_0", root=0, algorithm=ra.Algorithm_type.CC, engine_type=engine_type,
                            useDeviceNum=0)
    assert res["vertexValue"].tolist() == t["cc_directed"]
