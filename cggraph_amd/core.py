"""Python handles over libcgmi's graph and engine objects."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib


def _arr_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Graph:
    """A CSR graph owned by libcgmi (host and/or HBM copy)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle

    # ---- constructors ----
    @classmethod
    def from_csr(cls, offsets, dest, weight=None, zero_outdeg_tail=0):
        off = np.ascontiguousarray(offsets)
        u64 = off.dtype == np.uint64 or off.dtype == np.int64
        off = np.ascontiguousarray(off, dtype=np.uint64 if u64 else np.uint32)
        dst = np.ascontiguousarray(dest, dtype=np.uint32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.uint32)
        V = len(off) - 1
        h = ctypes.c_void_p()
        check(lib().cg_graph_from_csr(V, len(dst), _arr_ptr(off), int(u64), _arr_ptr(dst), _arr_ptr(w),
                                      int(zero_outdeg_tail), ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_files(cls, offset_file, dst_file, weight_file, V, E, offsets_u64=False):
        h = ctypes.c_void_p()
        check(lib().cg_graph_load_files(offset_file.encode(), int(offsets_u64), dst_file.encode(),
                                        None if weight_file is None else weight_file.encode(), V, E,
                                        ctypes.byref(h)))
        return cls(h)

    @classmethod
    def from_edges(cls, V, src, dst, weight=None, symmetrize=False):
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.uint32)
        h = ctypes.c_void_p()
        check(lib().cg_graph_from_edges(int(V), len(s), _arr_ptr(s), _arr_ptr(d), _arr_ptr(w), int(symmetrize),
                                        ctypes.byref(h)))
        return cls(h)

    @classmethod
    def load_edge_file(cls, path, V, has_weights=True, symmetrize=False):
        h = ctypes.c_void_p()
        check(lib().cg_graph_load_edge_file(path.encode(), int(V), int(has_weights), int(symmetrize),
                                            ctypes.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor, a=0.57, b=0.19, c=0.19, seed=1, weights=(1, 100), symmetrize=False):
        lo, hi = weights if weights is not None else (0, 0)
        h = ctypes.c_void_p()
        check(lib().cg_graph_generate_rmat(scale, edge_factor, a, b, c, seed, lo, hi, int(symmetrize),
                                           ctypes.byref(h)))
        return cls(h)

    @classmethod
    def grid(cls, n, drop=0.1, seed=1, weights=(1, 1000)):
        lo, hi = weights if weights is not None else (0, 0)
        h = ctypes.c_void_p()
        check(lib().cg_graph_generate_grid(n, drop, seed, lo, hi, ctypes.byref(h)))
        return cls(h)

    # ---- queries ----
    def info(self):
        V, E, hw, sym, mv = (ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_int(), ctypes.c_int(),
                             ctypes.c_uint32())
        check(lib().cg_graph_info(self._h, ctypes.byref(V), ctypes.byref(E), ctypes.byref(hw),
                                  ctypes.byref(sym), ctypes.byref(mv)))
        return {"V": V.value, "E": E.value, "has_weights": bool(hw.value), "symmetric": bool(sym.value),
                "max_outdeg_vertex": mv.value}

    @property
    def V(self):
        return self.info()["V"]

    @property
    def E(self):
        return self.info()["E"]

    def export(self):
        inf = self.info()
        off = np.empty(inf["V"] + 1, np.uint64)
        dst = np.empty(max(inf["E"], 1), np.uint32)
        w = np.empty(max(inf["E"], 1), np.uint32) if inf["has_weights"] else None
        check(lib().cg_graph_export(self._h, _arr_ptr(off), _arr_ptr(dst), _arr_ptr(w)))
        return off, dst[:inf["E"]], (None if w is None else w[:inf["E"]])

    def build_transpose(self):
        check(lib().cg_graph_build_transpose(self._h))

    def partition(self, algorithm):
        info = _lib.PartitionInfo()
        check(lib().cg_partition(self._h, int(algorithm), ctypes.byref(info)))
        return {"segment_num": info.segment_num, "bucket_start": list(info.bucket_start),
                "can_hold_all": bool(info.can_hold_all), "device_bytes": info.device_bytes}

    def close(self):
        if self._h and self._h.value:
            lib().cg_graph_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """One traversal engine (COORDINATION by default) over a Graph."""

    def __init__(self, graph, algorithm, engine_type=0, use_device_num=1, options=None, **opt):
        self.graph = graph
        self.algorithm = int(algorithm)
        o = options if options is not None else _lib.default_options(**opt)
        self._opt = o
        h = ctypes.c_void_p()
        check(lib().cg_engine_create(graph._h, self.algorithm, int(engine_type), int(use_device_num),
                                     ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.V = graph.V

    def run(self, root=0, copy_out=True):
        """Returns (values or None, stats dict)."""
        st = _lib.RunStats()
        out = None
        if copy_out:
            out = np.empty(self.V, np.float32 if self.algorithm == 3 else np.uint32)
        check(lib().cg_engine_run(self._h, int(root), _arr_ptr(out), ctypes.byref(st)))
        return out, st.as_dict()

    def result(self):
        out = np.empty(self.V, np.float32 if self.algorithm == 3 else np.uint32)
        check(lib().cg_engine_result(self._h, _arr_ptr(out)))
        return out

    def close(self):
        if self._h and self._h.value:
            lib().cg_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
