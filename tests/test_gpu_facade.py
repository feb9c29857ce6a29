"""GPU tests of the C++ boundary and of the reference's extreme paths.

* tests/cpp/cuda_step_test drives CudaDeviceManager's step API in the pattern
  of the reference's tests/cudaTest.cpp:33-232 and checks every step against
  the oracle (it links the oracle; exit status = failed tests);
* the reference's own main.cpp, built against the facade (oracle/Makefile,
  target dropin), prints the oracle's numbers;
* the iteration cap (cuda/device.h:309, 343-345) through the CLI, and the
  threshold extremes (sc_threshold = 0, a huge rw_threshold), on the engine;
* the deactivation-label collision of the oracle's hand-worked case (Q1) on
  the engine, and reset() after a failed initialize.
"""
import os
import subprocess

import numpy as np
import pytest

import parex_amd as px
from oracle_ctypes import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STEP_TEST = os.path.join(ROOT, "tests", "cpp", "_build", "cuda_step_test")
CLI = os.path.join(ROOT, "parex_amd", "bin", "parex")


def write_mtx(path, g):
    """Matrix Market pattern file with each undirected edge once, 1-based
    (the format lib/utils/graph_io.h:49-84 reads)."""
    src = np.repeat(np.arange(g.n, dtype=np.int64), np.diff(g.ranges))
    keep = src < g.edges
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern symmetric\n")
        f.write(f"{g.n} {g.n} {int(keep.sum())}\n")
        np.savetxt(f, np.stack([src[keep] + 1, g.edges[keep].astype(np.int64) + 1], axis=1), fmt="%d")


STEP_GRAPHS = {
    "sbm_c1": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
    "torus32": lambda: px.torus(32, 10, 1),
    "rmat10": lambda: px.rmat(10, 8, 0.57, 0.19, 0.19, 7),
    "sbm_dense": lambda: px.sbm(800, 4, 0.17, 0.01, 3),
}


def _no_tail_isolated(g):
    """main.cpp's loader corrupts ranges after a trailing isolated node (Q9)."""
    return g.n == 0 or g.ranges[-2] < g.ranges[-1]


@pytest.mark.parametrize("name", list(STEP_GRAPHS))
def test_step_api_program_matches_oracle(tmp_path, name):
    if not os.path.exists(STEP_TEST):
        pytest.fail(f"{STEP_TEST} is not built (make -C tests/cpp)")
    g = STEP_GRAPHS[name]()
    path = str(tmp_path / f"{name}.mtx")
    write_mtx(path, g)
    print(f"cuda_step_test {name}: n={g.n} m={g.m}", flush=True)
    r = subprocess.run([STEP_TEST, path, "6582", "0.3", "1e-4"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-6000:] + r.stderr[-2000:]
    assert "[==========] 16 tests, 0 failed" in r.stdout


@pytest.mark.parametrize("binary", ["parex_main_refhdr", "parex_main_ourhdr"])
@pytest.mark.parametrize("args", [("6582", "0.3", "1e-4"), ("11", "0.2", "1e-5")])
def test_reference_main_on_the_engine_matches_oracle(tmp_path, binary, args):
    """The reference's CLI source (main.cpp), unchanged, linked to the facade:
    its printed iterations, clusters and cut edges are the oracle's."""
    exe = os.path.join(ROOT, "oracle", "_ref", binary)
    if not os.path.exists(exe):
        pytest.skip("reference CLI not built here (needs the reference tree at build time)")
    g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    assert _no_tail_isolated(g)
    path = str(tmp_path / "c1.mtx")
    write_mtx(path, g)
    r = subprocess.run([exe, path, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    out = dict(l.split(": ", 1) for l in r.stdout.splitlines() if ": " in l)
    seed, sc, rw = int(args[0]), float(args[1]), float(args[2])
    dm = px.DeviceManager(px.Params(seed=seed, sc_threshold=sc, rw_threshold=rw))
    dm.initialize(g)
    ox = Oracle(g, seed=seed, sc=sc, rw=rw, init_values=dm.readRandomWalkValuesF64())
    it = ox.run()
    ids, k = ox.final_partition()
    cut, _, _, _ = px.verify_partition(g, ids, k)
    assert int(out["iterations"]) == it
    assert (int(out["numClusters"]), int(out["cutEdges"])) == (k, cut)
    assert (int(out["nodes"]), int(out["edges"])) == (g.n, g.m)


@pytest.mark.parametrize("cap", [1, 3, 12])
def test_cli_iteration_cap_prints_abort(tmp_path, cap):
    """cuda/device.h:309, 343-345: when the loop runs out of iterations the
    reference prints "Not terminating. ABORT." and iterations = cap + 1, and the
    partition is whatever the labels are then (PAREX_MAX_ITER lowers the cap)."""
    g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    path = str(tmp_path / "c1.pxg")
    px.write_pxg(path, g)
    env = dict(os.environ, PAREX_MAX_ITER=str(cap))
    r = subprocess.run([CLI, path, "6582", "0.3", "1e-4"], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stderr
    assert "Not terminating. ABORT." in r.stdout
    out = dict(l.split(": ", 1) for l in r.stdout.splitlines() if ": " in l)
    dm = px.DeviceManager()
    dm.initialize(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert ox.run(cap) == cap + 1 == int(out["iterations"])
    ids, k = ox.final_partition()
    cut, _, _, _ = px.verify_partition(g, ids, k)
    assert (int(out["numClusters"]), int(out["cutEdges"])) == (k, cut)
    # the uncapped run does not print the line
    r = subprocess.run([CLI, path, "6582", "0.3", "1e-4"], capture_output=True, text=True, timeout=120)
    assert "Not terminating" not in r.stdout


def test_engine_max_iter_equals_oracle():
    g = px.torus(32, 10, 1)
    for cap in (1, 2, 5):
        dm = px.DeviceManager(px.Params(max_iter=cap))
        dm.initialize(g)
        ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
        assert dm.expanderDecomposition() == ox.run(cap) == cap + 1
        ids, k = dm.getFinalPartition()
        oids, ok = ox.final_partition()
        assert k == ok and np.array_equal(ids, oids)


EXTREME_GRAPHS = {
    "sbm_c1": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
    "torus32": lambda: px.torus(32, 10, 1),
    "rmat12": lambda: px.rmat(12, 8, 0.57, 0.19, 0.19, 7),
    "islands": lambda: px.from_edges(40, [0, 1, 2, 5, 6, 7, 20, 21, 22, 22], [1, 2, 0, 6, 7, 5, 21, 22, 23, 24]),
}


@pytest.mark.parametrize("name", list(EXTREME_GRAPHS))
@pytest.mark.parametrize("sc,rw", [(0.0, 1e-4), (0.3, 1e30), (0.0, 1e30), (1.0, 1e-4)])
def test_threshold_extremes_match_oracle(name, sc, rw):
    """sc_threshold = 0: no sweep cut is ever taken (sparsity < 0 never holds);
    a huge rw_threshold deactivates every cluster in iteration 1, so the
    printed iterations are 2; sc_threshold = 1 cuts almost every cluster."""
    g = EXTREME_GRAPHS[name]()
    dm = px.DeviceManager(px.Params(sc_threshold=sc, rw_threshold=rw, record_node_data=False))
    dm.initialize(g)
    ox = Oracle(g, sc=sc, rw=rw, init_values=dm.readRandomWalkValuesF64())
    it = dm.expanderDecomposition()
    assert it == ox.run()
    if rw > 1e20:
        assert it == 2
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)
    assert ox.errors() == 0


def test_label_collision_case_on_the_engine():
    """The hand-worked Q1 case of the oracle's tests, on the engine: labels -1
    of iteration 1 and of iteration 3 collide and merge in the final ids."""
    g = px.from_edges(7, [1, 1, 2, 4, 4, 5], [2, 3, 3, 5, 6, 6])
    x0 = np.array([0.0, -10 - 7.5e-4, -10.0, -10 + 7.5e-4, 10 - 2e-3, 10.0, 10 + 2e-3])
    want = [[-1, 0, 0, 0, 1, 1, 1], [-1, 0, 0, 0, 1, 1, 1], [-1, -1, -1, -1, 0, 0, 0], [-1, -1, -1, -1, -2, -2, -2]]
    dm = px.DeviceManager()
    dm.initialize(g)
    dm.setRandomWalkValuesF64(x0)
    for it in range(4):
        dm.iterateRandomWalk()
        dm.computeSweepCuts()
        dm.cutClusters()
        assert list(dm.downloadLabels()) == want[it], it
    assert dm.counts()[1] == 0
    ids, k = dm.getFinalPartition()
    assert k == 2 and list(ids) == [1, 1, 1, 1, 0, 0, 0]


def test_reset_after_failed_initialize_is_refused():
    """A failed initialize leaves no valid input copy: reset() must refuse
    before any copy or launch, and a valid initialize afterwards works."""
    g = px.sbm(600, 5, 0.08, 0.004, 9)
    bad = g.edges.copy()
    bad[5] = g.n + 1
    dm = px.DeviceManager()
    dm.initialize(px.sbm(300, 3, 0.06, 0.003, 100))
    with pytest.raises(px.PxdError):
        dm.initialize(px.Graph(g.n, g.m, g.ranges.copy(), bad))
    with pytest.raises(px.PxdError, match="not initialized"):
        dm.reset()
    with pytest.raises(px.PxdError):
        dm.initialize_edges(10, [0, 1], [1, 10])
    with pytest.raises(px.PxdError, match="not initialized"):
        dm.reset()
    dm.initialize(g)
    x0 = dm.readRandomWalkValuesF64()
    ox = Oracle(g, init_values=x0)
    assert dm.expanderDecomposition() == ox.run()
    oids, ok = ox.final_partition()
    dm.reset()                       # valid again: back to the initial state
    assert np.array_equal(dm.readRandomWalkValuesF64().view(np.uint64), x0.view(np.uint64))
    assert dm.expanderDecomposition() == Oracle(g, init_values=x0).run()
    ids, k = dm.getFinalPartition()
    assert k == ok and np.array_equal(ids, oids)
