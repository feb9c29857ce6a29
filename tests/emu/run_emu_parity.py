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
}


def main(lib_path, names):
    L = C.CDLL(lib_path)
    L.pxd_last_error.restype = C.c_char_p

    def chk(rc):
        if rc:
            raise RuntimeError(L.pxd_last_error().decode())

    ptr = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    for name in names:
        g = GRAPHS[name]()
        p = Params()
        L.pxd_default_params(C.byref(p))
        h = C.c_void_p()
        chk(L.pxd_create(C.byref(p), C.byref(h)))
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
    sys.exit(main(sys.argv[1], sys.argv[2:] or list(GRAPHS)))
