"""ctypes binding of libcgmi.so (include/cgmi.h).

The library is built in-tree (`make -C cggraph_amd`, or __graft_entry__.build()).
Loading fails loudly when it is missing: there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgmi.so")

CG_OK = 0
CG_ERR_INVALID = 1
CG_ERR_NO_DEVICE = 2
CG_ERR_HIP = 3
CG_ERR_OOM = 4
CG_ERR_IO = 5
CG_ERR_UNSUPPORTED = 6
CG_ERR_INTERNAL = 7
CG_ERR_MISMATCH = 8

CG_UNREACHED = 999_999_999
CG_NUM_CHOICES = 16
CG_NUM_FEATURES = 12

CG_SCHED_HEURISTIC, CG_SCHED_GPU, CG_SCHED_LEARNED, CG_SCHED_FORCED, CG_SCHED_PROBE = 0, 1, 2, 3, 4
MLP_PARAMS = 6032
CG_NOALLOC_SCRATCH_BYTES = 64 << 20
CG_FAULT_HIP_ERROR, CG_FAULT_MLP_NAN, CG_FAULT_SLOW_GPU = 1, 2, 4


class Options(ctypes.Structure):
    _fields_ = [
        ("segment_size", ctypes.c_uint32),
        ("heavy_active_vertices", ctypes.c_uint32),
        ("heavy_active_edges", ctypes.c_uint64),
        ("bin_wave", ctypes.c_uint32),
        ("bin_chunk", ctypes.c_uint32),
        ("chunk_edges", ctypes.c_uint32),
        ("sched_mode", ctypes.c_int),
        ("forced_choice", ctypes.c_int),
        ("cpu_threads", ctypes.c_int),
        ("pr_iterations", ctypes.c_int),
        ("pr_damping", ctypes.c_float),
        ("profile", ctypes.c_int),
        ("trace_path", ctypes.c_char_p),
        ("mlp_path", ctypes.c_char_p),
        ("fault_flags", ctypes.c_uint32),
        ("batch_block", ctypes.c_uint32),
        ("batch_grid", ctypes.c_uint32),
        ("batch_levels", ctypes.c_uint32),
        ("sssp_delta", ctypes.c_uint32),
        ("pr_segment_log2", ctypes.c_uint32),
        ("atomic_return", ctypes.c_uint32),
        ("pr_layout", ctypes.c_uint32),
        ("relabel", ctypes.c_uint32),
        ("pr_kernel", ctypes.c_uint32),
        ("min_pull", ctypes.c_uint32),
        ("expand_grid", ctypes.c_uint32),
        ("batch_group", ctypes.c_uint32),
        ("hub_degree", ctypes.c_uint32),
        ("hub_delta", ctypes.c_uint32),
        ("probe_stride", ctypes.c_uint32),
        ("hbm_budget", ctypes.c_uint64),
        ("batch_engine_v", ctypes.c_uint32),
    ]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_uint32),
        ("heavy_iterations", ctypes.c_uint32),
        ("light_iterations", ctypes.c_uint32),
        ("pull_iterations", ctypes.c_uint32),
        ("ms", ctypes.c_double),
        ("edges_traversed", ctypes.c_uint64),
        ("edges_relaxed", ctypes.c_uint64),
        ("bytes_h2d", ctypes.c_uint64),
        ("bytes_d2h", ctypes.c_uint64),
        ("cpu_vertices", ctypes.c_uint64),
        ("gpu_vertices", ctypes.c_uint64),
        ("expand_launches", ctypes.c_uint32),
        ("expand_kernel_ms", ctypes.c_double),
        ("expand_alg_bytes", ctypes.c_double),
        ("learned_fallbacks", ctypes.c_uint32),
        ("resident_edges", ctypes.c_uint64),
        ("pr_layout", ctypes.c_uint32),
        ("pr_stream_bytes", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PartitionInfo(ctypes.Structure):
    _fields_ = [
        ("segment_num", ctypes.c_uint64),
        ("bucket_start", ctypes.c_uint64 * 5),
        ("can_hold_all", ctypes.c_int),
        ("device_bytes", ctypes.c_uint64),
    ]


# name -> (restype, argtypes); mirrors include/cgmi.h one to one.
_vp, _u32, _u64, _int, _dbl, _flt = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_float)
_cp = ctypes.c_char_p
_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "cg_options_default": (_int, [ctypes.POINTER(Options)]),
    "cg_graph_from_csr": (_int, [_u32, _u64, _vp, _int, _vp, _vp, _u32, _pp]),
    "cg_graph_load_files": (_int, [_cp, _int, _cp, _cp, _u32, _u64, _pp]),
    "cg_graph_from_edges": (_int, [_u32, _u64, _vp, _vp, _vp, _int, _pp]),
    "cg_graph_load_edge_file": (_int, [_cp, _u32, _int, _int, _pp]),
    "cg_graph_generate_rmat": (_int, [_u32, _u32, _dbl, _dbl, _dbl, _u64, _u32, _u32, _int, _pp]),
    "cg_graph_generate_grid": (_int, [_u32, _dbl, _u64, _u32, _u32, _pp]),
    "cg_graph_info": (_int, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64), ctypes.POINTER(_int),
                             ctypes.POINTER(_int), ctypes.POINTER(_u32)]),
    "cg_graph_export": (_int, [_vp, _vp, _vp, _vp]),
    "cg_graph_build_transpose": (_int, [_vp]),
    "cg_graph_destroy": (None, [_vp]),
    "cg_partition": (_int, [_vp, _int, ctypes.POINTER(PartitionInfo)]),
    "cg_engine_create": (_int, [_vp, _int, _int, _int, ctypes.POINTER(Options), _pp]),
    "cg_engine_run": (_int, [_vp, _u32, _vp, ctypes.POINTER(RunStats)]),
    "cg_engine_result": (_int, [_vp, _vp]),
    "cg_engine_result_device": (_int, [_vp, _pp]),
    "cg_engine_destroy": (None, [_vp]),
    "cg_push_step_device": (_int, [_int, _vp, _vp, _vp, _vp, _vp]),
    "cg_graph_device_csr": (_int, [_vp, _pp, _pp, _pp]),
    "cg_batch_run": (_int, [_vp, _vp, _vp, _int, _vp, _int, ctypes.POINTER(Options),
                            ctypes.POINTER(RunStats)]),
    "cg_engine_run_batch": (_int, [_vp, _vp, _int, _vp, ctypes.POINTER(RunStats)]),
    "cg_traverse_device_noalloc": (_int, [_int, _vp, _vp, _vp, _u32, _u64, _u32, _vp, _vp,
                                          ctypes.POINTER(_u32)]),
    "cg_mlp_train": (_int, [_cp, _cp, _int, _flt, _u64, ctypes.POINTER(_flt)]),
    "cg_mlp_forward": (_int, [_vp, _vp, _int, _vp]),
    "cg_mlp_loss_grad": (_int, [_vp, _vp, _vp, _int, _vp, _vp]),
    "cg_check_result_file": (_int, [_cp, _vp, _u32, _vp, ctypes.POINTER(_u64)]),
    "cg_save_bin_u32": (_int, [_cp, _vp, _u64]),
    "cg_abi_sizes": (_int, [ctypes.POINTER(ctypes.c_size_t)] * 3),
    "cg_last_error": (_cp, []),
    "cg_version": (_cp, []),
}

_LIB = None


class CgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cgmi error {code}: {msg}")
        self.code = code


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `make -C cggraph_amd` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(code):
    if code != CG_OK:
        raise CgError(code, lib().cg_last_error().decode(errors="replace"))
    return code


def default_options(**kw):
    o = Options()
    check(lib().cg_options_default(ctypes.byref(o)))
    for k, v in kw.items():
        if isinstance(v, str):
            v = v.encode()
        setattr(o, k, v)
    return o
