"""Runs the CPU emulation of the engine (tests/emu/_build/libpxd_emu*.so,
through its C ABI) end to end on small graphs and compares with the oracle
seeded with the emulated initial vector.  Test infrastructure: started by
tests/test_emu.py in a child process (with the sanitizer runtime preloaded
for the AddressSanitizer build).  Prints one line per graph; exit status 1
on the first difference."""
import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]

import numpy as np  # noqa: E402

import parex_amd as px  # noqa: E402
from oracle_ctypes import Oracle  # noqa: E402


class Params(C.Structure):
    _fields_ = [("seed", C.c_uint32), ("sc_threshold", C.c_float), ("rw_threshold", C.c_float),
                ("max_iter", C.c_uint32), ("keep_original", C.c_uint32), ("record_node_data", C.c_uint32),
                ("device", C.c_int32)]


GRAPHS = {
    # 100 positions: waves 2 and 3 of the single workgroup of every
    # position-range kernel lie wholly past the end of the list
    "path100": lambda: px.from_edges(100, np.arange(99), np.arange(1, 100)),
    "sbm300": lambda: px.sbm(300, 3, 0.06, 0.003, 100),
    "torus12": lambda: px.torus(12, 3, 1),
    # 2500 positions: the CC regroup sorts one segment of more than a tile
    # (the large-segment plan built on the device, several tiles)
    "sbm2500": lambda: px.sbm(2500, 2, 0.01, 0.0004, 7),
    # triangles, a path and isolated nodes; a graph without arcs
    "islands": lambda: px.from_edges(40, [0, 1, 2, 5, 6, 7, 20, 21, 22, 22], [1, 2, 0, 6, 7, 5, 21, 22, 23, 24]),
    "noarcs": lambda: px.from_edges(30, [], []),
    # one row of 599 arcs on a ring: rows above the one-thread limit of the
    # one-workgroup run
    "hub600": lambda: px.from_edges(*_hub_edges(600)),
    # every row above the one-thread limit: the 16-lane groups of the
    # one-workgroup run, several rows per wave
    "clique48": lambda: px.from_edges(48, *np.triu_indices(48, 1)),
    "sbm_dense": lambda: px.sbm(400, 4, 0.4, 0.02, 5),
}


def _hub_edges(n):
    fx = [0] * (n - 1) + list(range(1, n))
    tx = list(range(1, n)) + [i % (n - 1) + 1 for i in range(2, n + 1)]
    return n, np.array(fx, np.uint32), np.array(tx, np.uint32)


def _loops_dups_edges():
    rng = np.random.default_rng(5)
    fx = rng.integers(0, 300, 2000).astype(np.uint32)
    tx = rng.integers(0, 300, 2000).astype(np.uint32)
    fx[:10] = tx[:10]
    fx[10:20], tx[10:20] = fx[20:30], tx[20:30]
    return 300, fx, tx


# device CSR builds (pxd_initialize_edges) against the host build
# (lib/core/graph.h:147-184): rows of a wave, and one row of 2999 arcs above
# a workgroup's tile (the long-row digit passes)
EDGE_LISTS = {"ingest_loops_dups": _loops_dups_edges, "ingest_hub3000": lambda: _hub_edges(3000)}


def main(lib_path, names):
    L = C.CDLL(lib_path)
    L.pxd_last_error.restype = C.c_char_p

    def chk(rc):
        if rc:
            raise RuntimeError(L.pxd_last_error().decode())

    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    for name in names:
        if name in EDGE_LISTS:
            n, fx, tx = EDGE_LISTS[name]()
            g = px.from_edges(n, fx, tx)
            p = Params()
            L.pxd_default_params(C.byref(p))
            h = C.c_void_p()
            chk(L.pxd_create(C.byref(p), C.byref(h)))
            chk(L.pxd_initialize_edges(h, n, C.c_uint64(len(fx)), ptr(fx), ptr(tx)))
            r, e = np.zeros(n + 1, np.uint32), np.zeros(2 * len(fx), np.uint32)
            chk(L.pxd_download_graph(h, ptr(r), ptr(e)))
            L.pxd_destroy(h)
            same = np.array_equal(r, g.ranges) and np.array_equal(e, g.edges)
            print(f"{name}: n={n} m={len(fx)} CSR {'equal' if same else 'DIFFERENT'}", flush=True)
            if not same:
                return 1
            continue
        # "graph@v": the per-arc kernels pinned to variant v (pxd_set_walk_variant);
        # "graph/phases": the per-phase kernels instead of the one-workgroup run
        # (forced otherwise, whatever the graph's arc count)
        name, _, mode = name.partition("/")
        gname, _, var = name.partition("@")
        g = GRAPHS[gname]()
        p = Params()
        L.pxd_default_params(C.byref(p))
        h = C.c_void_p()
        chk(L.pxd_create(C.byref(p), C.byref(h)))
        if var:
            chk(L.pxd_set_walk_variant(h, int(var)))
        chk(L.pxd_set_small_graph_kernel(h, 0 if mode == "phases" else 2))
        r, e = np.ascontiguousarray(g.ranges, np.uint32), np.ascontiguousarray(g.edges, np.uint32)
        chk(L.pxd_initialize(h, g.n, g.m, ptr(r), ptr(e)))
        x0 = np.zeros(g.n)
        chk(L.pxd_read_random_walk_values_f64(h, ptr(x0)))
        it = C.c_int32()
        chk(L.pxd_expander_decomposition(h, C.byref(it)))
        ids, k = np.zeros(g.n, np.int32), C.c_int32()
        chk(L.pxd_get_final_partition(h, ptr(ids), C.byref(k)))
        x = np.zeros(g.n)
        chk(L.pxd_read_random_walk_values_f64(h, ptr(x)))
        L.pxd_destroy(h)
        ox = Oracle(g, init_values=x0)
        oit = ox.run()
        oids, ok = ox.final_partition()
        same = it.value == oit and k.value == ok and np.array_equal(ids, oids) and \
            np.array_equal(x.view(np.uint64), ox.values().view(np.uint64))
        print(f"{name}: iterations {it.value} (oracle {oit}), clusters {k.value} (oracle {ok}), "
              f"{'equal' if same else 'DIFFERENT'}", flush=True)
        if not same:
            return 1
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1], sys.argv[2:] or list(GRAPHS) + list(EDGE_LISTS)))
