"""parex_amd -- MI355X-native expander decomposition (HIP, gfx950).

Python host mirror of the reference's GPU interface (``class CudaDeviceManager``,
cuda/interface.h:16-57 of RobinMnk/parex) over the engine's C ABI
(``include/parex_amd/pxd.h``), plus the host graph ingest library
(``include/parex_amd/pxgen.h``).  Both are loaded from shared libraries built
in-tree by ``parex_amd.build``; nothing here falls back to a CPU path: a
missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
INVALID_EDGE = 0xFFFFFFFF

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")

# lib/core/definitions.h:57-77 layouts
SWEEP_CUT_DTYPE = np.dtype([("clusterId", np.int32), ("sparsity", np.float32), ("offset", np.uint32)])
NODE_DATA_DTYPE = np.dtype([("nix", np.uint32), ("prefixVolume", np.uint32),
                            ("prefixEdgeDiff", np.int32), ("offsetInCluster", np.uint32)])
LABELED_NODE_DTYPE = np.dtype([("nix", np.uint32), ("clusterId", np.int32)])


class PxdError(RuntimeError):
    """Error reported by the engine's C ABI (status code + message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"pxd error {code}: {msg}")
        self.code = code


# --------------------------------------------------------------------------- libraries

class _PxgGraph(C.Structure):
    _fields_ = [("n", C.c_uint32), ("m", C.c_uint32),
                ("ranges", C.POINTER(C.c_uint32)), ("edges", C.POINTER(C.c_uint32))]


class _SbmSpec(C.Structure):
    _fields_ = [("n", C.c_uint32), ("k", C.c_uint32), ("p_in", C.c_double),
                ("p_out", C.c_double), ("seed", C.c_uint64)]


class _Params(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("sc_threshold", C.c_float), ("rw_threshold", C.c_float),
                ("max_iter", C.c_uint32), ("keep_original", C.c_uint32),
                ("record_node_data", C.c_uint32), ("device", C.c_int32)]


_libs = {}


def lib_path(name: str) -> str:
    return os.path.join(PKG_DIR, name)


def _load(name: str) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                           "or `python -m parex_amd.build`")
    lib = C.CDLL(path)
    _libs[name] = lib
    return lib


def pxgen() -> C.CDLL:
    lib = _load("libpxgen.so")
    if not getattr(lib, "_typed", False):
        G = C.POINTER(_PxgGraph)
        lib.pxg_sbm.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_uint64, G]
        lib.pxg_torus.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, G]
        lib.pxg_rmat.argtypes = [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_uint64, G]
        lib.pxg_sbm_specs.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double,
                                      C.c_double, C.c_double, C.c_double, C.c_uint64, C.POINTER(_SbmSpec)]
        lib.pxg_sbm_batch.argtypes = [C.c_uint32, C.POINTER(_SbmSpec), G]
        lib.pxg_read_mtx.argtypes = [C.c_char_p, G]
        lib.pxg_write.argtypes = [C.c_char_p, G]
        lib.pxg_read.argtypes = [C.c_char_p, G]
        lib.pxg_from_edges.argtypes = [C.c_uint32, C.c_uint64, _u32p, _u32p, G]
        lib.pxg_verify_partition.argtypes = [G, _i32p, C.c_int32, C.POINTER(C.c_uint64), C.c_void_p,
                                             C.c_void_p, C.c_void_p]
        lib.pxg_free.argtypes = [G]
        lib.pxg_last_error.restype = C.c_char_p
        lib._typed = True
    return lib


def pxd() -> C.CDLL:
    lib = _load("libpxd.so")
    if not getattr(lib, "_typed", False):
        P = C.c_void_p
        lib.pxd_default_params.argtypes = [C.POINTER(_Params)]
        lib.pxd_last_error.restype = C.c_char_p
        lib.pxd_create.argtypes = [C.POINTER(_Params), C.POINTER(C.c_void_p)]
        lib.pxd_destroy.argtypes = [P]
        lib.pxd_initialize.argtypes = [P, C.c_uint32, C.c_uint32, _u32p, _u32p]
        lib.pxd_initialize_edges.argtypes = [P, C.c_uint32, C.c_uint64, _u32p, _u32p]
        lib.pxd_initialize_batch.argtypes = [P, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_void_p),
                                             C.POINTER(C.c_void_p)]
        for f in ("pxd_iterate_random_walk", "pxd_compute_sweep_cuts", "pxd_cut_clusters", "pxd_reset"):
            getattr(lib, f).argtypes = [P]
        lib.pxd_set_thresholds.argtypes = [P, C.c_float, C.c_float]
        lib.pxd_expander_decomposition.argtypes = [P, C.POINTER(C.c_int32)]
        lib.pxd_graph_iterations.argtypes = [P, _i32p, C.c_uint32]
        lib.pxd_get_final_partition.argtypes = [P, _i32p, _i32p]
        lib.pxd_read_random_walk_values.argtypes = [P, _f32p]
        lib.pxd_read_random_walk_values_f64.argtypes = [P, _f64p]
        lib.pxd_set_random_walk_values_f64.argtypes = [P, _f64p]
        lib.pxd_download_degrees.argtypes = [P, _u32p]
        lib.pxd_download_labels.argtypes = [P, _i32p]
        lib.pxd_download_graph.argtypes = [P, _u32p, _u32p]
        lib.pxd_download_active_edge_map.argtypes = [P, _i32p]
        lib.pxd_download_partition.argtypes = [P, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        lib.pxd_read_sweep_cuts.argtypes = [P, _i32p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        lib.pxd_get_num_cut_edges.argtypes = [P, C.POINTER(C.c_uint32)]
        lib.pxd_download_active_nodes.argtypes = [P, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        lib.pxd_get_counts.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.pxd_last_run_ms.argtypes = [P, C.POINTER(C.c_double)]
        lib.pxd_enable_walk_timing.argtypes = [P, C.c_int]
        lib.pxd_walk_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        lib.pxd_walk_launch_log.argtypes = [P, C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_uint32)]
        lib.pxd_set_walk_variant.argtypes = [P, C.c_int]
        lib.pxd_set_small_graph_kernel.argtypes = [P, C.c_int]
        lib.pxd_walk_tuning.argtypes = [P, _i32p, C.c_uint32, C.POINTER(C.c_uint64)]
        lib.pxd_serve_create.argtypes = [C.POINTER(_Params), C.c_uint32, C.c_uint64, C.c_uint32,
                                         C.POINTER(C.c_void_p)]
        lib.pxd_serve_admit.argtypes = [P, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_uint64,
                                        C.POINTER(C.c_int32)]
        lib.pxd_serve_step.argtypes = [P, C.POINTER(C.c_uint32)]
        lib.pxd_serve_enable_edges.argtypes = [P, C.c_uint32, C.c_uint64]
        lib.pxd_serve_admit_edges.argtypes = [P, C.c_uint32, C.c_uint64, _u32p, _u32p, C.c_uint64,
                                              C.POINTER(C.c_int32)]
        lib.pxd_serve_next.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]
        lib.pxd_serve_take.argtypes = [P, _i32p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        lib.pxd_serve_wait.argtypes = [P]
        lib.pxd_serve_stats.argtypes = [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.pxd_csr_from_edges_noalloc.argtypes = [C.c_uint32, C.c_uint64, P, P, P, P, P, P]
        lib._typed = True
    return lib


# C-ABI entry points declared in include/parex_amd/pxd.h (checked by the tests)
PXD_SYMBOLS = [
    "pxd_default_params", "pxd_last_error", "pxd_create", "pxd_destroy", "pxd_initialize",
    "pxd_initialize_batch", "pxd_set_thresholds", "pxd_reset", "pxd_iterate_random_walk", "pxd_compute_sweep_cuts", "pxd_cut_clusters",
    "pxd_expander_decomposition", "pxd_graph_iterations", "pxd_get_final_partition",
    "pxd_read_random_walk_values", "pxd_read_random_walk_values_f64", "pxd_download_degrees",
    "pxd_download_labels", "pxd_download_graph", "pxd_download_active_edge_map", "pxd_download_partition",
    "pxd_read_sweep_cuts", "pxd_get_num_cut_edges", "pxd_download_active_nodes", "pxd_get_counts",
    "pxd_set_random_walk_values_f64", "pxd_last_run_ms", "pxd_enable_walk_timing", "pxd_walk_stats",
    "pxd_walk_launch_log", "pxd_set_small_graph_kernel",
    "pxd_set_walk_variant", "pxd_walk_tuning", "pxd_initialize_edges", "pxd_serve_enable_edges",
    "pxd_serve_admit_edges",
    "pxd_serve_create", "pxd_serve_admit", "pxd_serve_step", "pxd_serve_next", "pxd_serve_wait", "pxd_serve_take",
    "pxd_serve_stats", "pxd_csr_from_edges_noalloc",
]

NOALLOC_MAX_NODES = 1 << 20    # PXD_NOALLOC_MAX_NODES
NOALLOC_MAX_EDGES = 1 << 21    # PXD_NOALLOC_MAX_EDGES

PXD_FULL = 1       # pxd_serve_admit: no room right now
PXD_EMPTY = 1      # pxd_serve_next / pxd_serve_take: no finished result waiting


# --------------------------------------------------------------------------- graphs

@dataclass
class Graph:
    """CSR graph (lib/core/graph.h:145-158): ranges u32[n+1], edges u32[2m]."""

    n: int
    m: int
    ranges: np.ndarray
    edges: np.ndarray

    def degree(self) -> np.ndarray:
        return np.diff(self.ranges).astype(np.uint32)


def _take(g: _PxgGraph) -> Graph:
    lib = pxgen()
    n, m = int(g.n), int(g.m)
    ranges = np.ctypeslib.as_array(g.ranges, shape=(n + 1,)).copy() if g.ranges else np.zeros(1, np.uint32)
    edges = (np.ctypeslib.as_array(g.edges, shape=(2 * m,)).copy() if (g.edges and m)
             else np.zeros(0, np.uint32))
    lib.pxg_free(C.byref(g))
    return Graph(n, m, ranges.astype(np.uint32), edges.astype(np.uint32))


def _gen_check(rc: int):
    if rc != 0:
        raise RuntimeError(pxgen().pxg_last_error().decode())


def sbm(n: int, k: int, p_in: float, p_out: float, seed: int) -> Graph:
    g = _PxgGraph()
    _gen_check(pxgen().pxg_sbm(n, k, p_in, p_out, seed, C.byref(g)))
    return _take(g)


def torus(L: int, extra: int, seed: int) -> Graph:
    g = _PxgGraph()
    _gen_check(pxgen().pxg_torus(L, extra, seed, C.byref(g)))
    return _take(g)


def rmat(scale: int, ef: int, a: float, b: float, c: float, seed: int) -> Graph:
    g = _PxgGraph()
    _gen_check(pxgen().pxg_rmat(scale, ef, a, b, c, seed, C.byref(g)))
    return _take(g)


def sbm_specs(count, nmin, nmax, kmin, kmax, din, dout, seed):
    arr = (_SbmSpec * count)()
    _gen_check(pxgen().pxg_sbm_specs(count, nmin, nmax, kmin, kmax, din[0], din[1], dout[0], dout[1], seed, arr))
    return arr


def sbm_batch(specs) -> List[Graph]:
    count = len(specs)
    arr = (_SbmSpec * count)(*specs)
    out = (_PxgGraph * count)()
    _gen_check(pxgen().pxg_sbm_batch(count, arr, out))
    return [_take(out[i]) for i in range(count)]


def from_edges(n: int, fx: Sequence[int], tx: Sequence[int]) -> Graph:
    fx = np.ascontiguousarray(fx, dtype=np.uint32)
    tx = np.ascontiguousarray(tx, dtype=np.uint32)
    g = _PxgGraph()
    _gen_check(pxgen().pxg_from_edges(n, len(fx), fx, tx, C.byref(g)))
    return _take(g)


def read_mtx(path: str) -> Graph:
    g = _PxgGraph()
    _gen_check(pxgen().pxg_read_mtx(path.encode(), C.byref(g)))
    return _take(g)


def write_pxg(path: str, graph: Graph):
    g = _PxgGraph(graph.n, graph.m, graph.ranges.ctypes.data_as(C.POINTER(C.c_uint32)),
                  graph.edges.ctypes.data_as(C.POINTER(C.c_uint32)))
    _gen_check(pxgen().pxg_write(path.encode(), C.byref(g)))


def read_pxg(path: str) -> Graph:
    g = _PxgGraph()
    _gen_check(pxgen().pxg_read(path.encode(), C.byref(g)))
    return _take(g)


def verify_partition(graph: Graph, ids: np.ndarray, num_clusters: int):
    """Host re-verification (main.cpp:53-60): returns (cut_edges, vol, internal_vol, conductance)."""
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    g = _PxgGraph(graph.n, graph.m, graph.ranges.ctypes.data_as(C.POINTER(C.c_uint32)),
                  graph.edges.ctypes.data_as(C.POINTER(C.c_uint32)))
    cut = C.c_uint64()
    vol = np.zeros(max(num_clusters, 1), np.uint64)
    ivol = np.zeros(max(num_clusters, 1), np.uint64)
    cond = np.zeros(max(num_clusters, 1), np.float64)
    _gen_check(pxgen().pxg_verify_partition(C.byref(g), ids, num_clusters, C.byref(cut), vol.ctypes.data,
                                            ivol.ctypes.data, cond.ctypes.data))
    return int(cut.value), vol[:num_clusters], ivol[:num_clusters], cond[:num_clusters]


# --------------------------------------------------------------------------- engine

@dataclass
class Params:
    seed: int = 6582                 # lib/core/definitions.h:46
    sc_threshold: float = 0.3        # lib/core/definitions.h:50
    rw_threshold: float = 1e-4       # lib/core/definitions.h:49
    max_iter: int = 200000           # cuda/device.h:309
    keep_original: bool = True
    record_node_data: bool = True
    device: int = 0


class Server:
    """Continuous-batching decomposition service over one engine: graphs are
    admitted between iterations while others are in flight (pxd_serve_*),
    each result equal to the graph's solo decomposition."""

    def __init__(self, node_capacity: int, arc_capacity: int, graph_slots: int, params: Params | None = None):
        self._lib = pxd()
        self.params = params or Params()
        p = _Params(self.params.seed, self.params.sc_threshold, self.params.rw_threshold, self.params.max_iter,
                    0, int(self.params.record_node_data), self.params.device)
        h = C.c_void_p()
        self._check(self._lib.pxd_serve_create(C.byref(p), node_capacity, arc_capacity, graph_slots, C.byref(h)))
        self._h = h
        self._keep = []

    def _check(self, rc: int):
        if rc < 0:
            raise PxdError(rc, self._lib.pxd_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pxd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, graph: Graph, tag: int) -> bool:
        """True if admitted, False if it does not fit the free capacity now."""
        ranges = np.ascontiguousarray(graph.ranges, np.uint32)
        edges = np.ascontiguousarray(graph.edges, np.uint32)
        slot = C.c_int32()
        rc = self._check(self._lib.pxd_serve_admit(self._h, graph.n, graph.m, ranges, edges, tag, C.byref(slot)))
        if rc == 0:
            self._keep.append((ranges, edges))   # read by the device until the next step
        return rc == 0

    def set_walk_variant(self, variant: int):
        """Kernel choice of the walk and cut contribution, as
        DeviceManager.setWalkVariant."""
        self._check(self._lib.pxd_set_walk_variant(self._h, int(variant)))

    def walk_tuning(self):
        t = np.full(66, -1, np.int32)
        n = C.c_uint64()
        self._check(self._lib.pxd_walk_tuning(self._h, t, 66, C.byref(n)))
        return ({b: int(v) for b, v in enumerate(t[:33]) if v >= 0},
                {b: int(v) for b, v in enumerate(t[33:]) if v >= 0}, int(n.value))

    def enable_edges(self, max_nodes: int, max_edges: int):
        """Workspace for edge-list admissions of graphs up to these sizes."""
        self._check(self._lib.pxd_serve_enable_edges(self._h, int(max_nodes), int(max_edges)))

    def admit_edges(self, n: int, fx, tx, tag: int) -> bool:
        """Admit an undirected edge list (CSR built on the device); True if
        admitted, False if it does not fit the free capacity now."""
        fx = np.ascontiguousarray(fx, np.uint32)
        tx = np.ascontiguousarray(tx, np.uint32)
        if fx.shape != tx.shape:
            raise ValueError("fx and tx differ in length")
        slot = C.c_int32()
        rc = self._check(self._lib.pxd_serve_admit_edges(self._h, int(n), int(fx.size), fx, tx, tag, C.byref(slot)))
        if rc == 0:
            self._keep.append((fx, tx))     # read by the device until the next step
        return rc == 0

    def step(self) -> int:
        """One iteration of every graph in flight; returns how many finished."""
        f = C.c_uint32()
        self._check(self._lib.pxd_serve_step(self._h, C.byref(f)))
        self._keep = []
        return f.value

    def results(self, wait: bool = False):
        """Pop the ready results as (tag, ids, num_clusters, iterations); with
        wait, first wait until every finished graph's result is ready."""
        if wait:
            self._check(self._lib.pxd_serve_wait(self._h))
        out = []
        tag, n = C.c_uint64(), C.c_uint32()
        while self._check(self._lib.pxd_serve_next(self._h, C.byref(tag), C.byref(n))) == 0:
            ids = np.empty(n.value, np.int32)
            k, it = C.c_int32(), C.c_int32()
            self._check(self._lib.pxd_serve_take(self._h, ids, C.byref(k), C.byref(it)))
            out.append((tag.value, ids, k.value, it.value))
        return out

    def stats(self):
        g, nn, aa, na, w = C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.pxd_serve_stats(self._h, C.byref(g), C.byref(nn), C.byref(aa), C.byref(na), C.byref(w)))
        return {"live_graphs": g.value, "live_nodes": nn.value, "live_arcs": aa.value, "active_nodes": na.value,
                "waiting_results": w.value}


def csr_from_edges_noalloc(n: int, m: int, fx_ptr: int, tx_ptr: int, ranges_ptr: int, edges_ptr: int, bad_ptr: int,
                           stream: int = 0):
    """pxd_csr_from_edges_noalloc: the device CSR build of DeviceManager.initialize_edges
    from device edge arrays (fx, tx: u32[m]) into device buffers ranges u32[n+1] and
    edges u32[2m], with the library's static work area (no device allocation); queued
    on the HIP stream `stream` (0 = null stream).  *bad (device u32) becomes nonzero
    when an end is >= n.  Arguments are device addresses (ints)."""
    lib = pxd()
    rc = lib.pxd_csr_from_edges_noalloc(int(n), int(m), fx_ptr or None, tx_ptr or None, ranges_ptr or None,
                                        edges_ptr or None, bad_ptr or None, stream or None)
    if rc != 0:
        raise PxdError(rc, lib.pxd_last_error().decode())


class DeviceManager:
    """Mirror of ``CudaDeviceManager`` (cuda/interface.h:16-57) over the HIP engine.

    Method names follow the reference (camelCase) with snake_case aliases.
    Errors raise :class:`PxdError`; the reference reported none.
    """

    def __init__(self, params: Params | None = None):
        self._lib = pxd()
        self.params = params or Params()
        p = _Params(self.params.seed, self.params.sc_threshold, self.params.rw_threshold, self.params.max_iter,
                    int(self.params.keep_original), int(self.params.record_node_data), self.params.device)
        h = C.c_void_p()
        self._check(self._lib.pxd_create(C.byref(p), C.byref(h)))
        self._h = h
        self._n = 0
        self._m = 0
        self._graphs = 1
        self._node_counts = None

    def _check(self, rc: int):
        if rc != 0:
            raise PxdError(rc, self._lib.pxd_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pxd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # cuda/interface.h:24
    def initialize(self, graph: Graph):
        self._n, self._m, self._graphs = graph.n, graph.m, 1
        self._node_counts = np.array([graph.n], np.uint32)
        self._check(self._lib.pxd_initialize(self._h, graph.n, graph.m,
                                             np.ascontiguousarray(graph.ranges, np.uint32),
                                             np.ascontiguousarray(graph.edges, np.uint32)))

    def initialize_edges(self, n: int, fx, tx):
        """readGraph + Graph::finalize + initialize with the CSR built on the
        device from the undirected edge list (fx[i], tx[i])."""
        fx = np.ascontiguousarray(fx, np.uint32)
        tx = np.ascontiguousarray(tx, np.uint32)
        if fx.shape != tx.shape:
            raise ValueError("fx and tx differ in length")
        self._n, self._m, self._graphs = int(n), int(fx.size), 1
        self._node_counts = np.array([n], np.uint32)
        self._check(self._lib.pxd_initialize_edges(self._h, int(n), int(fx.size), fx, tx))

    def initialize_batch(self, graphs: Sequence[Graph]):
        G = len(graphs)
        nc = np.array([g.n for g in graphs], np.uint32)
        ec = np.array([g.m for g in graphs], np.uint32)
        rs = [np.ascontiguousarray(g.ranges, np.uint32) for g in graphs]
        es = [np.ascontiguousarray(g.edges, np.uint32) if g.m else np.zeros(1, np.uint32) for g in graphs]
        rp = (C.c_void_p * G)(*[r.ctypes.data for r in rs])
        ep = (C.c_void_p * G)(*[e.ctypes.data for e in es])
        self._n, self._m, self._graphs = int(nc.sum()), int(ec.sum()), G
        self._node_counts = nc
        self._check(self._lib.pxd_initialize_batch(self._h, G, nc, ec, rp, ep))

    def setThresholds(self, sc_threshold: float, rw_threshold: float):
        self.params.sc_threshold, self.params.rw_threshold = sc_threshold, rw_threshold
        self._check(self._lib.pxd_set_thresholds(self._h, sc_threshold, rw_threshold))

    def reset(self):
        self._check(self._lib.pxd_reset(self._h))

    def iterateRandomWalk(self):
        self._check(self._lib.pxd_iterate_random_walk(self._h))

    def computeSweepCuts(self):
        self._check(self._lib.pxd_compute_sweep_cuts(self._h))

    def cutClusters(self):
        self._check(self._lib.pxd_cut_clusters(self._h))

    def expanderDecomposition(self) -> int:
        it = C.c_int32()
        self._check(self._lib.pxd_expander_decomposition(self._h, C.byref(it)))
        return int(it.value)

    def graphIterations(self) -> np.ndarray:
        out = np.zeros(self._graphs, np.int32)
        self._check(self._lib.pxd_graph_iterations(self._h, out, self._graphs))
        return out

    def getFinalPartition(self):
        ids = np.zeros(max(self._n, 1), np.int32)
        nc = np.zeros(self._graphs, np.int32)
        self._check(self._lib.pxd_get_final_partition(self._h, ids, nc))
        ids = ids[: self._n]
        if self._graphs == 1:
            return ids, int(nc[0])
        return ids, nc

    def readRandomWalkValues(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), np.float32)
        self._check(self._lib.pxd_read_random_walk_values(self._h, out))
        return out[: self._n]

    def readRandomWalkValuesF64(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), np.float64)
        self._check(self._lib.pxd_read_random_walk_values_f64(self._h, out))
        return out[: self._n]

    def setRandomWalkValuesF64(self, values: np.ndarray):
        v = np.ascontiguousarray(values, np.float64)
        self._check(self._lib.pxd_set_random_walk_values_f64(self._h, v))

    def downloadDegrees(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), np.uint32)
        self._check(self._lib.pxd_download_degrees(self._h, out))
        return out[: self._n]

    def downloadLabels(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), np.int32)
        self._check(self._lib.pxd_download_labels(self._h, out))
        return out[: self._n]

    def downloadGraph(self) -> Graph:
        r = np.zeros(self._n + self._graphs, np.uint32)
        e = np.zeros(max(2 * self._m, 1), np.uint32)
        self._check(self._lib.pxd_download_graph(self._h, r, e))
        return Graph(self._n, self._m, r[: self._n + 1], e[: 2 * self._m])

    def downloadActiveEdgeMap(self) -> np.ndarray:
        out = np.zeros(max(2 * self._m, 1), np.int32)
        self._check(self._lib.pxd_download_active_edge_map(self._h, out))
        return out[: 2 * self._m]

    def downloadPartition(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), NODE_DATA_DTYPE)
        cnt = C.c_uint32()
        self._check(self._lib.pxd_download_partition(self._h, out.ctypes.data, max(self._n, 1), C.byref(cnt)))
        return out[: cnt.value]

    def readSweepCuts(self) -> Tuple[np.ndarray, np.ndarray]:
        ids = np.zeros(max(self._n, 1), np.int32)
        cuts = np.zeros(max(self._n, 1), SWEEP_CUT_DTYPE)
        cnt = C.c_uint32()
        self._check(self._lib.pxd_read_sweep_cuts(self._h, ids, cuts.ctypes.data, max(self._n, 1), C.byref(cnt)))
        return ids[: cnt.value], cuts[: cnt.value]

    def getNumCutEdges(self) -> int:
        out = C.c_uint32()
        self._check(self._lib.pxd_get_num_cut_edges(self._h, C.byref(out)))
        return int(out.value)

    def downloadActiveNodes(self) -> np.ndarray:
        out = np.zeros(max(self._n, 1), LABELED_NODE_DTYPE)
        cnt = C.c_uint32()
        self._check(self._lib.pxd_download_active_nodes(self._h, out.ctypes.data, max(self._n, 1), C.byref(cnt)))
        return out[: cnt.value]

    def counts(self) -> Tuple[int, int, int]:
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        self._check(self._lib.pxd_get_counts(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), int(b.value), int(c.value)

    def getNumClusters(self) -> int:
        """Returns 0 like the reference stub (cuda/device.h:58-61)."""
        return 0

    def updateLabels(self, node_labels, active_clusters):
        """No-op like the reference (cuda/devGraph.h:191-206)."""
        return None

    def lastRunMs(self) -> float:
        v = C.c_double()
        self._check(self._lib.pxd_last_run_ms(self._h, C.byref(v)))
        return float(v.value)

    def enableWalkTiming(self, on: bool = True):
        self._check(self._lib.pxd_enable_walk_timing(self._h, int(on)))

    def walkStats(self):
        a, b, c = C.c_uint64(), C.c_double(), C.c_double()
        self._check(self._lib.pxd_walk_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return int(a.value), float(b.value), float(c.value)

    def walkLaunchLog(self):
        """(ms, algorithmic bytes) of every timed walk launch, in launch order."""
        n = C.c_uint32()
        self._check(self._lib.pxd_walk_launch_log(self._h, None, None, 0, C.byref(n)))
        ms, by = np.zeros(n.value), np.zeros(n.value)
        self._check(self._lib.pxd_walk_launch_log(self._h, ms.ctypes.data_as(C.c_void_p),
                                                   by.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
        return ms, by

    def setSmallGraphKernel(self, enable):
        """The one-workgroup run of a small single graph (at most 4096 nodes)
        in expanderDecomposition: True / 1 (default) for graphs of at most
        2^16 arcs, 2 for every graph of at most 2^19 arcs, False / 0 never."""
        self._check(self._lib.pxd_set_small_graph_kernel(self._h, int(enable)))

    def setWalkVariant(self, variant: int):
        """Pin the kernels of the walk and the cut contribution (0..5: 2..64
        lanes per row; 6, 7, 8: LDS window with 1, 2, 4 lanes per row; 9: LDS
        window, 2 lanes, 1024-entry chunks; 10: 7 with clamped walk slots; 11:
        LDS window over node-range chunks of 8192 indices; 12: the same over
        4096-index chunks with half the window; 13: LDS window, 2 lanes,
        2048-entry chunks, half the window; 14 / 15: 11 / 12 with one lane per
        row), -1 =
        online tuning (default), -2 = the degree rule.  Every variant computes
        the same bits."""
        self._check(self._lib.pxd_set_walk_variant(self._h, int(variant)))

    def walkTuning(self):
        """({log2 list-length class: walk variant}, {class: cut-contribution
        variant}, number of tunings run)."""
        t = np.full(66, -1, np.int32)
        n = C.c_uint64()
        self._check(self._lib.pxd_walk_tuning(self._h, t, 66, C.byref(n)))
        walk = {b: int(v) for b, v in enumerate(t[:33]) if v >= 0}
        contrib = {b: int(v) for b, v in enumerate(t[33:]) if v >= 0}
        return walk, contrib, int(n.value)

    # snake_case aliases
    iterate_random_walk = iterateRandomWalk
    compute_sweep_cuts = computeSweepCuts
    cut_clusters = cutClusters
    expander_decomposition = expanderDecomposition
    get_final_partition = getFinalPartition
    read_random_walk_values = readRandomWalkValues
    download_degrees = downloadDegrees
    download_labels = downloadLabels
    download_graph = downloadGraph
    download_active_edge_map = downloadActiveEdgeMap
    download_partition = downloadPartition
    read_sweep_cuts = readSweepCuts
    get_num_cut_edges = getNumCutEdges
    get_num_clusters = getNumClusters
    update_labels = updateLabels


# --------------------------------------------------------------------------- hierarchy

def pxh() -> C.CDLL:
    lib = _load("libpxh.so")
    if not getattr(lib, "_typed", False):
        P = C.c_void_p
        lib.pxh_last_error.restype = C.c_char_p
        lib.pxh_create.argtypes = [C.c_uint32, C.c_uint32, _u32p, _u32p, C.POINTER(C.c_void_p)]
        lib.pxh_destroy.argtypes = [P]
        lib.pxh_build.argtypes = [P, C.POINTER(_Params), C.c_float, C.c_float, C.c_uint32]
        lib.pxh_add_level.argtypes = [P, C.c_uint32, _u32p, _u32p]
        lib.pxh_height.argtypes = [P, C.POINTER(C.c_uint32)]
        lib.pxh_level_size.argtypes = [P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint32)]
        lib.pxh_level.argtypes = [P, C.c_uint32] + [_u32p] * 7
        lib.pxh_normalized_cut.argtypes = [P, C.c_uint32, _i32p, C.POINTER(C.c_float)]
        lib.pxh_subtract_volume.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        lib.pxh_add_cut.argtypes = [P, C.c_uint32, C.c_uint32, C.c_uint32]
        lib._typed = True
    return lib


# C-ABI entry points declared in include/parex_amd/pxh.h (checked by the tests)
PXH_SYMBOLS = [
    "pxh_last_error", "pxh_create", "pxh_destroy", "pxh_build", "pxh_add_level", "pxh_height",
    "pxh_level_size", "pxh_level", "pxh_normalized_cut", "pxh_subtract_volume", "pxh_add_cut",
]


@dataclass
class Level:
    """One level of the hierarchy (lib/algorithms/expander_hierarchy.h:17-23):
    the contracted multigraph, the subtree volumes and cut sizes of its nodes,
    the cluster of each node below, and the partition of the graph below."""

    graph: Graph
    volumes: np.ndarray
    cut_sizes: np.ndarray
    parents: np.ndarray
    member_off: np.ndarray
    members: np.ndarray


class ExpanderHierarchy:
    """Mirror of ``ExpanderHierarchy`` (lib/algorithms/expander_hierarchy.h:13-160)
    whose decompositions run on the HIP engine (libpxh.so over libpxd.so)."""

    def __init__(self, graph: Graph):
        self._lib = pxh()
        self._graph = graph
        h = C.c_void_p()
        self._check(self._lib.pxh_create(graph.n, graph.m, np.ascontiguousarray(graph.ranges, np.uint32),
                                         np.ascontiguousarray(graph.edges, np.uint32), C.byref(h)))
        self._h = h

    def _check(self, rc: int):
        if rc != 0:
            raise PxdError(rc, self._lib.pxh_last_error().decode())

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pxh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # expander_hierarchy.h:107-140 (Config{} defaults: targetSparsity 0.2, randomWalkThreshold 1e-5)
    def build(self, params: Params | None = None, target_sparsity: float = 0.2, rw_threshold: float = 1e-5,
              max_retries: int = 64):
        p = params or Params()
        cp = _Params(p.seed, p.sc_threshold, p.rw_threshold, p.max_iter, 0, 0, p.device)
        self._check(self._lib.pxh_build(self._h, C.byref(cp), target_sparsity, rw_threshold, max_retries))

    # expander_hierarchy.h:29-88 for a given partition of the top graph
    def add_level(self, member_off: np.ndarray, members: np.ndarray):
        off = np.ascontiguousarray(member_off, np.uint32)
        mem = np.ascontiguousarray(members, np.uint32)
        h = self.height()
        top = self.lvl(h - 1).graph.n if h else self._graph.n
        if off.ndim != 1 or len(off) < 1 or mem.ndim != 1 or len(mem) != top:
            raise ValueError(f"add_level: members must list each of the top level's {top} nodes once "
                             f"and member_off hold num_clusters + 1 offsets")
        self._check(self._lib.pxh_add_level(self._h, len(off) - 1, off, mem))

    def getGraph(self) -> Graph:
        return self._graph

    def height(self) -> int:
        v = C.c_uint32()
        self._check(self._lib.pxh_height(self._h, C.byref(v)))
        return int(v.value)

    def lvl(self, level: int) -> Level:
        k, arcs, nb = C.c_uint32(), C.c_uint64(), C.c_uint32()
        self._check(self._lib.pxh_level_size(self._h, level, C.byref(k), C.byref(arcs), C.byref(nb)))
        k, arcs, nb = int(k.value), int(arcs.value), int(nb.value)
        a = [np.zeros(s, np.uint32) for s in (k + 1, max(arcs, 1), k, k, nb, k + 1, nb)]
        self._check(self._lib.pxh_level(self._h, level, *a))
        ranges, edges, vol, cut, parents, off, members = a
        return Level(Graph(k, arcs // 2, ranges, edges[:arcs]), vol, cut, parents, off, members)

    # expander_hierarchy.h:154-159
    def subtractVolume(self, level: int, nix: int, vol: int):
        self._check(self._lib.pxh_subtract_volume(self._h, level, nix, vol))

    def addCut(self, level: int, nix: int, cut: int):
        self._check(self._lib.pxh_add_cut(self._h, level, nix, cut))

    get_graph = getGraph
    subtract_volume = subtractVolume
    add_cut = addCut


class NormalizedCut:
    """Mirror of ``NormalizedCut`` (lib/algorithms/normalized_cut.h:12-145)."""

    def __init__(self, hierarchy: ExpanderHierarchy):
        self.hierarchy = hierarchy

    # normalized_cut.h:108-144: part of each input node, and the objective
    # (lib/core/partition.h:418-427) of that partition
    def compute(self, k: int):
        h = self.hierarchy
        labels = np.zeros(h.getGraph().n, np.int32)
        val = C.c_float()
        h._check(h._lib.pxh_normalized_cut(h._h, k, labels, C.byref(val)))
        return labels, float(val.value)
