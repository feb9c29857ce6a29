"""The one-workgroup run of small single graphs (k_decompose_small) on the GPU.

pxd_expander_decomposition of one graph of at most 4096 nodes and 2^19 arcs,
straight after initialize, runs every iteration inside one workgroup when
forced (pxd_set_small_graph_kernel(2); the default takes it up to 2^16 arcs).
Every case below is decomposed three ways from the same initial vector: that
run, the per-phase kernels (pxd_set_small_graph_kernel(0)) and the oracle.  The
iteration count, the final partition, the labels, the internal degrees and
the final walk values must be identical.  A capped run (max_iter) must leave
the engine in the state the per-phase kernels leave it in: step calls after
it give the same sweeps, labels and values.
"""
import numpy as np
import pytest

import parex_amd as px
from oracle_ctypes import Oracle

pytestmark = pytest.mark.gpu


def _hub(n):
    fx = [0] * (n - 1) + list(range(1, n))
    tx = list(range(1, n)) + [i % (n - 1) + 1 for i in range(2, n + 1)]
    return px.from_edges(n, fx, tx)


def cases():
    return {
        "sbm_c1": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
        "sbm_4096_dense": lambda: px.sbm(4096, 8, 30 / 511, 2 / 3584, 3),    # the largest one-workgroup graph
        "sbm_4097": lambda: px.sbm(4097, 8, 14 / 511, 1 / 3584, 3),         # one node over: per-phase kernels
        "torus48": lambda: px.torus(48, 20, 5),
        "rmat12": lambda: px.rmat(12, 8, 0.57, 0.19, 0.19, 11),
        "path2000": lambda: px.from_edges(2000, np.arange(1999), np.arange(1, 2000)),
        "hub3000": lambda: _hub(3000),
        "clique64": lambda: px.from_edges(64, *np.triu_indices(64, 1)),
        "no_arcs": lambda: px.from_edges(50, [], []),
        "islands": lambda: px.from_edges(40, [0, 1, 2, 5, 6, 7, 20, 21, 22, 22], [1, 2, 0, 6, 7, 5, 21, 22, 23, 24]),
        "dense_deg229": lambda: px.sbm(500, 2, 0.9, 0.02, 7),
    }


def _run(g, small, **kw):
    dm = px.DeviceManager(px.Params(**kw))
    dm.setSmallGraphKernel(2 if small else 0)
    dm.initialize(g)
    x0 = dm.readRandomWalkValuesF64()
    it = dm.expanderDecomposition()
    ids, k = dm.getFinalPartition()
    st = {"it": it, "ids": ids, "k": k, "labels": dm.downloadLabels(), "dint": dm.downloadDegrees(),
          "x": dm.readRandomWalkValuesF64().view(np.uint64), "x0": x0, "cut": dm.getNumCutEdges()}
    return dm, st


@pytest.mark.parametrize("name", list(cases()))
def test_one_workgroup_run_equals_phases_and_oracle(name):
    g = cases()[name]()
    dm1, a = _run(g, True)
    dm2, b = _run(g, False)
    dm1.close()
    dm2.close()
    assert np.array_equal(a["x0"].view(np.uint64), b["x0"].view(np.uint64))
    for key in ("it", "k", "cut"):
        assert a[key] == b[key], f"{name}: {key} {a[key]} != {b[key]}"
    for key in ("ids", "labels", "dint", "x"):
        assert np.array_equal(a[key], b[key]), f"{name}: {key} differs"
    ox = Oracle(g, init_values=a["x0"])
    assert a["it"] == ox.run()
    oids, ok = ox.final_partition()
    assert a["k"] == ok and np.array_equal(a["ids"], oids)
    assert np.array_equal(a["x"], ox.values().view(np.uint64))
    assert ox.errors() == 0


@pytest.mark.parametrize("seed", range(12))
def test_one_workgroup_run_random_small_graphs(seed):
    """Random small graphs (planted partitions, RMAT, sparse random graphs with
    isolated nodes), both paths and the oracle."""
    rng = np.random.default_rng(500 + seed)
    kind = seed % 3
    if kind == 0:
        n = int(rng.integers(100, 4097))
        k = int(rng.integers(2, 30))
        bs = max(n // k, 2)
        g = px.sbm(n, k, min(rng.uniform(3, 25) / (bs - 1), 1.0), rng.uniform(0.2, 3) / max(n - bs, 1),
                   int(rng.integers(1 << 30)))
    elif kind == 1:
        g = px.rmat(int(rng.integers(7, 13)), int(rng.integers(2, 12)), 0.57, 0.19, 0.19, int(rng.integers(1 << 30)))
    else:
        n = int(rng.integers(30, 4096))
        m = int(rng.integers(0, 3 * n))
        fx, tx = rng.integers(0, n, m), rng.integers(0, n, m)
        keep = fx != tx
        pairs = np.unique(np.sort(np.stack([fx[keep], tx[keep]]), axis=0), axis=1)
        g = px.from_edges(n, pairs[0], pairs[1])
    dm1, a = _run(g, True, record_node_data=False)
    dm2, b = _run(g, False, record_node_data=False)
    dm1.close()
    dm2.close()
    assert a["it"] == b["it"] and a["k"] == b["k"]
    assert np.array_equal(a["ids"], b["ids"]) and np.array_equal(a["x"], b["x"]), f"seed {seed}: n={g.n} m={g.m}"
    ox = Oracle(g, init_values=a["x0"])
    assert a["it"] == ox.run()
    oids, ok = ox.final_partition()
    assert a["k"] == ok and np.array_equal(a["ids"], oids)


@pytest.mark.parametrize("cap", [1, 3, 7])
def test_capped_run_hands_over_the_engine_state(cap):
    """max_iter stops the one-workgroup run with clusters still active; the
    engine's state then equals the per-phase kernels' state: two more
    iterations by step calls give the same sweep cuts, labels, degrees and
    values, and the final partitions agree."""
    g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    out = []
    for small in (True, False):
        dm = px.DeviceManager(px.Params(max_iter=cap))
        dm.setSmallGraphKernel(2 if small else 0)
        dm.initialize(g)
        it = dm.expanderDecomposition()
        rec = [it, dm.downloadLabels(), dm.readRandomWalkValuesF64().view(np.uint64)]
        for _ in range(2):
            dm.iterateRandomWalk()
            dm.computeSweepCuts()
            lab, cuts = dm.readSweepCuts()
            rec += [lab, cuts, dm.readRandomWalkValuesF64().view(np.uint64)]
            dm.cutClusters()
            rec += [dm.downloadLabels(), dm.downloadDegrees(), dm.downloadActiveNodes()]
        rec.append(dm.getFinalPartition()[0])
        dm.close()
        out.append(rec)
    a, b = out
    assert a[0] == b[0]
    for i, (u, v) in enumerate(zip(a[1:], b[1:])):
        assert np.array_equal(np.asarray(u), np.asarray(v)), f"cap {cap}: record {i + 1} differs"


def test_automatic_choice_bounds_the_arcs():
    """The default mode (1) chooses by arc count (a 4096-node SBM of 52,918
    arcs below 2^16, a denser one of the same size above); whichever path it
    takes, the results equal forcing the one-workgroup run (2) and the
    per-phase kernels (0).  A mode outside 0..2 is an error."""
    sparse = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    dense = px.sbm(4096, 8, 30 / 511, 2 / 3584, 3)
    assert 2 * sparse.m <= 1 << 16 < 2 * dense.m
    for g in (sparse, dense):
        res = []
        for mode in (1, 2, 0):
            dm = px.DeviceManager(px.Params())
            dm.setSmallGraphKernel(mode)
            dm.initialize(g)
            res.append((dm.expanderDecomposition(), dm.getFinalPartition()[0]))
            dm.close()
        assert all(r[0] == res[0][0] and np.array_equal(r[1], res[0][1]) for r in res)
    with pytest.raises(px.PxdError):
        px.DeviceManager(px.Params()).setSmallGraphKernel(3)
