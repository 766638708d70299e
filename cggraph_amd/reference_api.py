"""The reference's driver-level surface, mirrored over libcgmi.

Names, argument meaning and error behaviour follow the reference (paths
relative to CGgraph/); errors raise CgError instead of exit(1):

  Algorithm_type, Engine_type, OrderMethod   Basic/Graph/basic_struct.hpp:41-178
  CSR_Result_type                            Basic/Graph/basic_struct.hpp:7-16
  GraphFile_type, getGraphFile               Basic/Graph/basic_struct.hpp:99-111,
                                             Basic/Graph/graphFileList.hpp:11-504
  getToyGraph                                Basic/Graph/graphFileList.hpp:510-570
  load_binFile / save_binArray               Basic/Other/IO.hpp:44-114
  checkBinResult, getResultFilePath          Basic/Graph/util.cuh:24-118
  nbrSort_taskSteal                          my_project/CGgraph/sortNbr.hpp:8-100
  TwoLevelPartition                          my_project/CGgraph/Two-level_partition.cuh:106-748
  CGgraphEngine.CG_co_execution              my_project/CGgraph/CGgraph_engine.cuh:154-974
  main_partition                             my_project/CGgraph/main_partition.hpp:18-243

The reference elides its graph paths ("...", graphFileList.hpp:8); here a
graph named G is looked up under $CG_GRAPH_DIR/G/ with the file names listed in
_file_names() below.
"""
import ctypes
import os
from dataclasses import dataclass, field
from enum import IntEnum
from typing import Optional

import numpy as np

from . import _lib
from ._lib import CgError, check, lib
from .core import Engine, Graph

VertexValue_MAX = _lib.CG_UNREACHED
SEGMENT_SIZE = 1024
BUCKET_NUM = 4


class Algorithm_type(IntEnum):
    BFS = 0
    SSSP = 1
    CC = 2
    PR = 3
    UNDEFINE = 4


class Engine_type(IntEnum):
    COORDINATION = 0
    COOPERATION = 1
    SINGLE_CORE = 2
    MULTI_CORE = 3
    SINGLE_GPU = 4
    BLOCK_2D = 5
    COMPUTE_INDEGREE = 6
    REORDER = 7
    ACTIVEBLOCK = 8


class OrderMethod(IntEnum):
    NATIVE = 0
    OUTDEGREE_NODE = 1
    OUTDEGREE_SOCKET = 2
    OUTDEGREE_THREAD = 3
    GORDER = 4
    RCM = 5
    RCM_LAST_ZERO_OUTDEGREE = 6
    RCM_LAST_ONE_ZERO_DEGREE = 7
    OUTDEGREE_NODE_dec = 8
    DEGREE_NODE = 9
    RCM_INDEGREE = 10
    RCM_OUTDEGREE = 11
    NORMAL_OUTDEGREE_DEC = 12
    UnDirected_NATIVE = 13
    UnDirected_RCM = 14
    UnDirected_RCM_LAST_ZERO_OUTDEGREE = 15
    MY_BFS_OUT = 16
    MY_BFS_IN = 17
    MY_BFS_OUT_IN = 18
    MY_IMAX_BFS_OUT = 19
    MY_IMAX_BFS_IN = 20
    IN_DEGREE_DEC = 21


# orderings whose files carry the zero-out-degree tail count ("addition" file)
_TAIL_ORDERS = {OrderMethod.RCM_LAST_ZERO_OUTDEGREE, OrderMethod.RCM_INDEGREE,
                OrderMethod.RCM_OUTDEGREE, OrderMethod.NORMAL_OUTDEGREE_DEC}


@dataclass
class CSR_Result_type:
    vertexNum: int = 0
    edgeNum: int = 0
    csr_offset: Optional[np.ndarray] = None
    csr_dest: Optional[np.ndarray] = None
    csr_weight: Optional[np.ndarray] = None
    noZeroOutDegreeNum: int = 0


@dataclass
class GraphFile_type:
    graphFile: str = ""
    vertices: int = 0
    common_root: int = 0
    edges: int = 0
    old2newFile: str = ""
    addtitionFile: str = ""
    csrOffsetFile: str = ""
    csrDestFile: str = ""
    csrWeightFile: str = ""


# |V|, |E| of the reference registry (Basic/Graph/graphFileList.hpp:18-413)
GRAPH_REGISTRY = {
    "cusha": (8, 14),
    "enwiki": (4206785, 101311614),
    "twitter2010": (61578415, 1468364884),
    "friendster": (124836180, 1806067135),
    "uk-union": (133633040, 5475109924),
}


def _file_names(order):
    tag = "native" if order == OrderMethod.NATIVE else OrderMethod(order).name.lower()
    return {"csrOffsetFile": f"{tag}_csrOffset_u32.bin", "csrDestFile": f"{tag}_csrDest_u32.bin",
            "csrWeightFile": f"{tag}_csrWeight_u32.bin", "old2newFile": f"{tag}_old2new.bin",
            "addtitionFile": f"{tag}_addition.bin", "graphFile": f"{tag}_edges.bin"}


def getGraphFile(graphName, orderMethod=OrderMethod.NATIVE):
    if graphName not in GRAPH_REGISTRY:
        raise CgError(_lib.CG_ERR_INVALID, f"no graph named [{graphName}] in the registry")
    V, E = GRAPH_REGISTRY[graphName]
    base = os.path.join(os.environ.get("CG_GRAPH_DIR", ""), graphName)
    gf = GraphFile_type(vertices=V, edges=E,
                        common_root=3 if orderMethod == OrderMethod.RCM_LAST_ZERO_OUTDEGREE else 0)
    for k, v in _file_names(orderMethod).items():
        setattr(gf, k, os.path.join(base, v))
    if orderMethod == OrderMethod.NATIVE:
        gf.old2newFile = gf.addtitionFile = ""
    return gf


_TOYS = {
    "toy_0": ([0, 3, 6, 6, 8, 8, 11, 15, 16, 21, 22, 24, 26],
              [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5],
              [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5]),
    "toy_1": ([0, 3, 6, 6, 8, 8, 11, 16, 17, 21, 24, 26, 28],
              [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 10, 5, 0, 4, 9, 11, 5, 6, 8, 2, 8, 2, 5],
              [1, 3, 7, 3, 5, 9, 5, 6, 7, 9, 11, 0, 5, 8, 9, 5, 0, 2, 4, 9, 11, 5, 2, 8, 2, 5, 2, 5]),
}


def getToyGraph(graphName, orderMethod=OrderMethod.NATIVE):
    if graphName not in _TOYS:
        raise CgError(_lib.CG_ERR_INVALID, "no toy graph of that name")
    off, dst, w = _TOYS[graphName]
    return CSR_Result_type(vertexNum=12, edgeNum=len(dst), csr_offset=np.array(off, np.uint32),
                           csr_dest=np.array(dst, np.uint32), csr_weight=np.array(w, np.uint32),
                           noZeroOutDegreeNum=12)


def load_binFile(path, length, dtype=np.uint32):
    """Raw little-endian array; the size must match exactly (IO.hpp:51)."""
    dtype = np.dtype(dtype)
    if not os.path.exists(path):
        raise CgError(_lib.CG_ERR_IO, f"cannot open {path}")
    n = os.path.getsize(path)
    if n != length * dtype.itemsize:
        raise CgError(_lib.CG_ERR_IO, f"(total_bytes / utilSize != length) for {path}")
    return np.fromfile(path, dtype=dtype, count=length)


def save_binArray(outFile, array):
    np.ascontiguousarray(array).tofile(outFile)


def getAlgName(algorithm):
    return Algorithm_type(algorithm).name


def getResultFilePath(graphName, algorithm, root, base=None):
    base = base if base is not None else os.environ.get("CG_RESULT_DIR", "")
    return os.path.join(base, f"{graphName}_{getAlgName(algorithm)}_{root}.bin")


def checkBinResult(resultFilePath, current_array, length, orderMethod=OrderMethod.NATIVE, old2new=None):
    """Exact compare against a result file; through old2new unless NATIVE."""
    cur = np.ascontiguousarray(current_array, np.uint32)
    o2n = None if orderMethod == OrderMethod.NATIVE or old2new is None else \
        np.ascontiguousarray(old2new, np.uint32)
    bad = ctypes.c_uint64(0)
    check(lib().cg_check_result_file(resultFilePath.encode(), cur.ctypes.data_as(ctypes.c_void_p), int(length),
                                     None if o2n is None else o2n.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.byref(bad)))
    return True


def nbrSort_taskSteal(csr, ascending=True, check_result=True):
    """Sorts every adjacency list (carrying weights); the check asserts
    strictly ascending neighbours, as sortNbr.hpp:84-91 does."""
    off = np.asarray(csr.csr_offset, np.int64)
    for v in range(csr.vertexNum):
        a, b = off[v], off[v + 1]
        if b - a < 2:
            continue
        order = np.argsort(csr.csr_dest[a:b], kind="stable")
        if not ascending:
            order = order[::-1]
        csr.csr_dest[a:b] = csr.csr_dest[a:b][order]
        if csr.csr_weight is not None:
            csr.csr_weight[a:b] = csr.csr_weight[a:b][order]
        if check_result and ascending and np.any(np.diff(csr.csr_dest[a:b].astype(np.int64)) <= 0):
            raise CgError(_lib.CG_ERR_INVALID, f"vertex {v}: neighbours are not strictly ascending")
    return csr


class TwoLevelPartition:
    """Segments of 1024 vertices in four degree buckets (17 / 33 / 257),
    charged against this GPU's HBM; the graph then lives in HBM."""

    def __init__(self, csrResult, zeroOutDegreeNum, deviceNum, algorithm):
        if deviceNum > 1:
            raise CgError(_lib.CG_ERR_UNSUPPORTED, "libcgmi drives one GPU per engine")
        self.csr = csrResult
        self.algorithm = Algorithm_type(algorithm)
        w = csrResult.csr_weight if self.algorithm == Algorithm_type.SSSP else None
        self.graph = Graph.from_csr(csrResult.csr_offset, csrResult.csr_dest, w,
                                    zero_outdeg_tail=int(zeroOutDegreeNum))
        self._info = self.graph.partition(int(self.algorithm))

    def getPartitionResult(self):
        return dict(self._info, graph=self.graph)


class CGgraphEngine:
    """CGgraphEngine(csr, partition, algorithm, useDeviceNum); CG_co_execution
    returns milliseconds and leaves the result in .vertexValue."""

    def __init__(self, csrResult, partitionResult, algorithm, useDeviceNum=1, engine_type=Engine_type.COORDINATION,
                 **options):
        self.algorithm = Algorithm_type(algorithm)
        self.graph = partitionResult["graph"]
        self.engine = Engine(self.graph, int(self.algorithm), int(engine_type), int(useDeviceNum), **options)
        self.vertexNum = csrResult.vertexNum
        self.noZeroOutDegreeNum = csrResult.noZeroOutDegreeNum or csrResult.vertexNum
        self.vertexValue = None
        self.last_stats = None

    def CG_co_execution(self, root=0):
        if self.algorithm in (Algorithm_type.BFS, Algorithm_type.SSSP) and root >= self.vertexNum:
            raise CgError(_lib.CG_ERR_INVALID, f"root {root} >= V")
        vals, st = self.engine.run(int(root))
        self.vertexValue = vals
        self.last_stats = st
        return st["ms"]


def main_partition(graphName="toy_0", orderMethod=OrderMethod.NATIVE, root=0, runs=1, algorithm=Algorithm_type.SSSP,
                   engine_type=Engine_type.COORDINATION, useDeviceNum=1, logResult=False, rootTranfer=True,
                   RCM_LAST_ZERO_OUTDEGREE_check=True, csr=None, **options):
    """The reference driver: load, partition, run `runs` times, report the
    trimmed mean (min and max dropped when runs > 2), optionally check the
    result file.  `csr` supplies a CSR_Result_type directly."""
    zeroOutDegreeNum = 0
    old2new = None
    if csr is None:
        if graphName.startswith("toy"):
            csr = getToyGraph(graphName, orderMethod)
        else:
            gf = getGraphFile(graphName, orderMethod)
            csr = CSR_Result_type(vertexNum=gf.vertices, edgeNum=gf.edges,
                                  csr_offset=load_binFile(gf.csrOffsetFile, gf.vertices + 1),
                                  csr_dest=load_binFile(gf.csrDestFile, gf.edges),
                                  csr_weight=load_binFile(gf.csrWeightFile, gf.edges))
            if orderMethod in _TAIL_ORDERS:
                zeroOutDegreeNum = int(load_binFile(gf.addtitionFile, 1)[0])
                if RCM_LAST_ZERO_OUTDEGREE_check and \
                        int(csr.csr_offset[csr.vertexNum - zeroOutDegreeNum]) != csr.edgeNum:
                    raise CgError(_lib.CG_ERR_INVALID, "zero-out-degree tail check failed")
            csr.noZeroOutDegreeNum = csr.vertexNum - zeroOutDegreeNum
            if orderMethod != OrderMethod.NATIVE and rootTranfer:
                old2new = load_binFile(gf.old2newFile, gf.vertices)
            nbrSort_taskSteal(csr)
    if engine_type not in (Engine_type.COORDINATION, Engine_type.SINGLE_GPU, Engine_type.MULTI_CORE,
                           Engine_type.SINGLE_CORE):
        raise CgError(_lib.CG_ERR_UNSUPPORTED, f"engine type {Engine_type(engine_type).name} is not provided")
    if engine_type in (Engine_type.MULTI_CORE, Engine_type.SINGLE_CORE):
        # Graph_Numa (main_partition.hpp:248-317): no partition, no GPU
        useDeviceNum = 0
    part = TwoLevelPartition(csr, zeroOutDegreeNum, useDeviceNum, algorithm).getPartitionResult()
    eng = CGgraphEngine(csr, part, algorithm, useDeviceNum, engine_type, **options)
    if root < 0 or root >= csr.vertexNum:
        raise CgError(_lib.CG_ERR_INVALID, f"root {root} outside [0, V)")
    switch_root = int(old2new[root]) if old2new is not None else int(root)
    times = sorted(eng.CG_co_execution(switch_root) for _ in range(runs))
    mean = sum(times[1:-1]) / (runs - 2) if runs > 2 else sum(times) / runs
    if logResult:
        checkBinResult(getResultFilePath(graphName, algorithm, root), eng.vertexValue, csr.vertexNum, orderMethod,
                       old2new)
    return {"times_ms": times, "total_run_time_ms": mean, "vertexValue": eng.vertexValue,
            "stats": eng.last_stats, "partition": {k: v for k, v in part.items() if k != "graph"}}
