"""ctypes binding of oracle/liboracle.so -- the CPU checker used by the tests.

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg load it, and only to check or time against.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

SWEEP_CUT_DTYPE = np.dtype([("clusterId", np.int32), ("sparsity", np.float32), ("offset", np.uint32)])
NODE_DATA_DTYPE = np.dtype([("nix", np.uint32), ("prefixVolume", np.uint32),
                            ("prefixEdgeDiff", np.int32), ("offsetInCluster", np.uint32)])
LABELED_NODE_DTYPE = np.dtype([("nix", np.uint32), ("clusterId", np.int32)])
CDATA_DTYPE = np.dtype([("rwSum", np.float32), ("maxPotential", np.float32), ("minPotential", np.float32),
                        ("totalElements", np.uint32)])

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle")
        L = C.CDLL(LIB)
        V = C.c_void_p
        L.ox_create.restype = V
        L.ox_create.argtypes = [C.c_uint32, C.c_uint32, _u32p, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                C.c_void_p]
        L.ox_destroy.argtypes = [V]
        for f in ("ox_walk", "ox_sweep", "ox_cut"):
            getattr(L, f).argtypes = [V]
        L.ox_run.argtypes = [V, C.c_int]
        L.ox_final_partition.argtypes = [V, _i32p]
        for f in ("ox_num_cut_edges", "ox_num_nodes", "ox_num_active_nodes", "ox_num_active_clusters",
                  "ox_num_clusters_with_cut"):
            getattr(L, f).argtypes = [V]
            getattr(L, f).restype = C.c_uint32
        L.ox_smallest_label.argtypes = [V]
        L.ox_error_flags.argtypes = [V]
        for f in ("ox_get_values", "ox_get_labels", "ox_get_internal_degrees", "ox_get_neighbors",
                  "ox_get_active", "ox_get_ndata", "ox_get_sweepcuts", "ox_get_cluster_labels",
                  "ox_get_cluster_data"):
            getattr(L, f).argtypes = [V, C.c_void_p]
        L.ox_init_vector.argtypes = [C.c_uint32, _u32p, C.c_uint32, _f64p]
        L.ox_init_value.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.ox_init_value.restype = C.c_float
        L.ox_erfcinv.argtypes = [C.c_double]
        L.ox_erfcinv.restype = C.c_double
        L.ox_hash.argtypes = [C.c_uint32]
        L.ox_hash.restype = C.c_uint32
        L.ox_minstd_first.argtypes = [C.c_uint32]
        L.ox_minstd_first.restype = C.c_uint32
        L.ox_pack_key.argtypes = [C.c_int32, C.c_float]
        L.ox_pack_key.restype = C.c_uint64
        L.ox_fixsum.argtypes = [_f64p, C.c_uint32, C.POINTER(C.c_int)]
        L.ox_fixsum.restype = C.c_double
        L.ox_cluster_sum.argtypes = [_f32p, C.c_uint32]
        L.ox_cluster_sum.restype = C.c_float
        L.ox_finalize.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, _u32p, _u32p]
        L.ox_finalize.restype = C.c_uint32
        L.ox_verify_cut_edges.argtypes = [C.c_uint32, _u32p, _u32p, _i32p]
        L.ox_verify_cut_edges.restype = C.c_uint32
        L.ox_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


class Oracle:
    """Stateful CPU restatement of CudaDeviceManager's decomposition path."""

    def __init__(self, graph, seed=6582, sc=0.3, rw=1e-4, init_values=None):
        L = lib()
        self.n, self.m = graph.n, graph.m
        self._edges = np.ascontiguousarray(graph.edges, np.uint32)
        iv = None
        if init_values is not None:
            self._iv = np.ascontiguousarray(init_values, np.float64)
            iv = self._iv.ctypes.data
        self.h = L.ox_create(graph.n, graph.m, np.ascontiguousarray(graph.ranges, np.uint32),
                             self._edges.ctypes.data if graph.m else None, seed, sc, rw, iv)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ox_destroy(self.h)
            self.h = None

    def walk(self):
        lib().ox_walk(self.h)

    def sweep(self):
        lib().ox_sweep(self.h)

    def cut(self):
        lib().ox_cut(self.h)

    def run(self, max_iter=200000) -> int:
        return lib().ox_run(self.h, max_iter)

    def final_partition(self):
        ids = np.zeros(max(self.n, 1), np.int32)
        k = lib().ox_final_partition(self.h, ids)
        return ids[: self.n], int(k)

    def num_cut_edges(self) -> int:
        return int(lib().ox_num_cut_edges(self.h))

    def counts(self):
        L = lib()
        return (int(L.ox_num_active_nodes(self.h)), int(L.ox_num_active_clusters(self.h)),
                int(L.ox_num_clusters_with_cut(self.h)))

    def errors(self) -> int:
        return int(lib().ox_error_flags(self.h))

    def smallest(self) -> int:
        return int(lib().ox_smallest_label(self.h))

    def _get(self, fn, count, dtype):
        out = np.zeros(max(count, 1), dtype)
        getattr(lib(), fn)(self.h, out.ctypes.data)
        return out[:count]

    def values(self):
        return self._get("ox_get_values", self.n, np.float64)

    def labels(self):
        return self._get("ox_get_labels", self.n, np.int32)

    def internal_degrees(self):
        return self._get("ox_get_internal_degrees", self.n, np.uint32)

    def neighbors(self):
        return self._get("ox_get_neighbors", 2 * self.m, np.uint32)

    def active(self):
        return self._get("ox_get_active", self.counts()[0], LABELED_NODE_DTYPE)

    def ndata(self):
        return self._get("ox_get_ndata", self.counts()[0], NODE_DATA_DTYPE)

    def sweepcuts(self):
        return self._get("ox_get_sweepcuts", self.counts()[1], SWEEP_CUT_DTYPE)

    def cluster_labels(self):
        return self._get("ox_get_cluster_labels", self.counts()[1], np.int32)


def set_threads(n: int):
    """OpenMP threads of the calling Python thread's later oracle calls."""
    lib().ox_set_threads(int(n))


def init_vector(degrees, seed):
    d = np.ascontiguousarray(degrees, np.uint32)
    out = np.zeros(max(len(d), 1), np.float64)
    lib().ox_init_vector(len(d), d, seed, out)
    return out[: len(d)]


def fixsum(terms):
    t = np.ascontiguousarray(terms, np.float64)
    ovf = C.c_int()
    v = lib().ox_fixsum(t, len(t), C.byref(ovf))
    return v, ovf.value


def cluster_sum(vals):
    v = np.ascontiguousarray(vals, np.float32)
    return lib().ox_cluster_sum(v, len(v))


def pack_key(label, v):
    return int(lib().ox_pack_key(int(label), float(v)))
