"""Host-side checks that need no GPU: the C ABI exports what its headers
declare, graph ingest follows the reference's parse and CSR rules
(lib/utils/graph_io.h:49-84, lib/core/graph.h:147-184), the generators build
what they promise, the host verifier (main.cpp:53-60) and the CLI's argument
handling (main.cpp:15-27).  Nothing here launches a kernel."""
import ctypes as C
import itertools
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import parex_amd as px

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include", "parex_amd")


def declared(header, prefix):
    """Function names declared in a C header (one declaration may span lines)."""
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(%s\w+)\s*\(" % prefix, text)))


@pytest.mark.parametrize("header,lib,prefix", [("pxd.h", "libpxd.so", "pxd_"), ("pxgen.h", "libpxgen.so", "pxg_")])
def test_library_exports_every_declared_symbol(header, lib, prefix):
    names = declared(header, prefix)
    assert len(names) >= (29 if prefix == "pxd_" else 12)
    dll = C.CDLL(px.lib_path(lib))            # loading runs no kernel
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_python_binding_covers_the_c_abi():
    assert sorted(px.PXD_SYMBOLS) == declared("pxd.h", "pxd_")


@pytest.mark.skipif(shutil.which("nm") is None, reason="nm not available")
def test_facade_exports_the_reference_class():
    """libparex_cuda.so defines the members of CudaDeviceManager (cuda/interface.h:16-57)."""
    out = subprocess.run(["nm", "-DC", "--defined-only", px.lib_path("libparex_cuda.so")], capture_output=True,
                         text=True, check=True).stdout
    for member in ["CudaDeviceManager()", "~CudaDeviceManager()", "initialize(Graph const&)", "downloadGraph()",
                   "downloadDegrees()", "downloadPartition()", "iterateRandomWalk()", "readRandomWalkValues()",
                   "computeSweepCuts()", "readSweepCuts()", "cutClusters()", "getNumCutEdges()",
                   "expanderDecomposition()", "getFinalPartition()", "downloadActiveEdgeMap()", "downloadLabels()",
                   "getNumClusters()", "updateLabels("]:
        assert "CudaDeviceManager::" + member in out, member


def test_missing_library_fails_loudly():
    with pytest.raises(RuntimeError, match="not built"):
        px._load("libdoes_not_exist.so")


# ------------------------------------------------------------------ ingest

def reference_csr(n, pairs):
    """The reference's finalize: both directions of every edge, sorted by
    (from, to), duplicates and loops kept (a loop gives two arcs)."""
    arcs = sorted([(a, b) for a, b in pairs] + [(b, a) for a, b in pairs])
    ranges = np.zeros(n + 1, np.uint32)
    for a, _ in arcs:
        ranges[a + 1] += 1
    return np.cumsum(ranges).astype(np.uint32), np.array([b for _, b in arcs], np.uint32)


def test_read_mtx_follows_reference_rules(tmp_path):
    pairs = [(2, 0), (0, 1), (1, 2), (2, 0), (3, 3), (1, 4)]        # duplicate, loop; nodes 5, 6 isolated
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n% comment\n" + f"7 7 {len(pairs)}\n"
                 + "".join(f"{a + 1} {b + 1}\n" for a, b in pairs))
    g = px.read_mtx(str(p))
    ranges, edges = reference_csr(7, pairs)
    assert (g.n, g.m) == (7, len(pairs))
    np.testing.assert_array_equal(g.ranges, ranges)      # trailing isolated nodes: empty rows
    np.testing.assert_array_equal(g.edges, edges)


def test_from_edges_equals_mtx_path(tmp_path):
    rng = np.random.default_rng(3)
    n, m = 50, 200
    fx, tx = rng.integers(0, n, m), rng.integers(0, n, m)
    g = px.from_edges(n, fx, tx)
    ranges, edges = reference_csr(n, list(zip(fx.tolist(), tx.tolist())))
    np.testing.assert_array_equal(g.ranges, ranges)
    np.testing.assert_array_equal(g.edges, edges)


def test_pxg_round_trip(tmp_path):
    g = px.sbm(300, 3, 0.1, 0.01, 5)
    p = str(tmp_path / "g.pxg")
    px.write_pxg(p, g)
    h = px.read_pxg(p)
    assert (h.n, h.m) == (g.n, g.m)
    np.testing.assert_array_equal(h.ranges, g.ranges)
    np.testing.assert_array_equal(h.edges, g.edges)


def arcs_of(g):
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.ranges.astype(np.int64)))
    return src, g.edges.astype(np.int64)


def assert_simple_symmetric(g):
    s, t = arcs_of(g)
    assert len(s) == 2 * g.m
    assert not np.any(s == t), "self-loop"
    key = s * g.n + t
    assert len(np.unique(key)) == len(key), "duplicate arc"
    np.testing.assert_array_equal(np.sort(key), np.sort(t * g.n + s))     # every arc has its reverse
    for u in range(0, g.n, max(1, g.n // 50)):                           # rows sorted ascending
        row = g.edges[g.ranges[u]:g.ranges[u + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)


def test_torus_closed_form():
    L = 8
    g = px.torus(L, 0, 1)
    assert (g.n, g.m) == (L * L, 2 * L * L)
    for r, c in itertools.product(range(L), range(L)):
        u = r * L + c
        want = sorted({((r + 1) % L) * L + c, ((r - 1) % L) * L + c, r * L + (c + 1) % L, r * L + (c - 1) % L})
        assert sorted(g.edges[g.ranges[u]:g.ranges[u + 1]].tolist()) == want
    h = px.torus(16, 40, 2)
    assert h.m == 2 * 256 + 40
    assert_simple_symmetric(h)


def test_sbm_blocks_and_density():
    n, k = 2000, 4
    g = px.sbm(n, k, 0.05, 0.0, 9)
    s, t = arcs_of(g)
    assert np.all(s * k // n == t * k // n), "p_out = 0 must keep every edge inside its block"
    assert_simple_symmetric(g)
    pairs = k * (n // k) * (n // k - 1) / 2
    assert abs(g.m - 0.05 * pairs) < 5 * np.sqrt(0.05 * pairs)         # binomial, 5 sigma
    h = px.sbm(n, k, 0.0, 0.01, 9)
    s, t = arcs_of(h)
    assert np.all(s * k // n != t * k // n), "p_in = 0 leaves only inter-block edges"


def test_rmat_simple_symmetric():
    g = px.rmat(10, 8, 0.57, 0.19, 0.19, 7)
    assert g.n == 1024 and 0 < g.m <= 8 * 1024
    assert_simple_symmetric(g)
    assert np.array_equal(g.edges, px.rmat(10, 8, 0.57, 0.19, 0.19, 7).edges)      # seeded


def test_sbm_specs_and_batch():
    specs = px.sbm_specs(40, 1024, 262144, 4, 64, (8.0, 32.0), (0.5, 3.0), 11)
    again = px.sbm_specs(40, 1024, 262144, 4, 64, (8.0, 32.0), (0.5, 3.0), 11)
    for a, b in zip(specs, again):
        assert (a.n, a.k, a.p_in, a.p_out, a.seed) == (b.n, b.k, b.p_in, b.p_out, b.seed)
        assert 1024 <= a.n <= 262144 and 4 <= a.k <= 64
        din = a.p_in * (a.n // a.k - 1)
        assert 8.0 - 1e-9 <= din <= 32.0 + 1e-9
    batch = px.sbm_batch(specs[:3])
    for sp, g in zip(specs[:3], batch):
        solo = px.sbm(sp.n, sp.k, sp.p_in, sp.p_out, sp.seed)
        np.testing.assert_array_equal(g.edges, solo.edges)


# ------------------------------------------------------------------ verifier

def test_verify_partition_brute_force():
    rng = np.random.default_rng(4)
    g = px.sbm(120, 3, 0.2, 0.02, 4)
    ids = rng.integers(0, 5, g.n).astype(np.int32)
    cut, vol, ivol, cond = px.verify_partition(g, ids, 5)
    s, t = arcs_of(g)
    assert cut == int(np.sum(ids[s] != ids[t])) // 2
    total = len(s)
    for c in range(5):
        v = int(np.sum(ids[s] == c))
        iv = int(np.sum((ids[s] == c) & (ids[t] == c)))
        assert (vol[c], ivol[c]) == (v, iv)
        den = min(v, total - v)
        assert cond[c] == pytest.approx((v - iv) / den if den else 0.0, rel=0, abs=0)


def test_verify_partition_rejects_bad_ids():
    g = px.torus(4, 0, 1)
    with pytest.raises(Exception):
        px.verify_partition(g, np.full(g.n, 3, np.int32), 2)


# ------------------------------------------------------------------ CLI

CLI = os.path.join(ROOT, "parex_amd", "bin", "parex")


def test_cli_without_arguments_exits_with_message():
    r = subprocess.run([CLI], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "missing arguments" in r.stdout


def test_cli_unreadable_graph_fails_before_the_device(tmp_path):
    r = subprocess.run([CLI, str(tmp_path / "none.mtx"), "1", "0.3", "1e-4"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1
    assert r.stdout.startswith("sparsity: 0.300000\nthreshold: 0.000100\n")
