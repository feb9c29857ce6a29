"""GPU parity: the HIP engine against the oracle, step by step and end to end.

The oracle is seeded with the device's initial walk vector (the reference's own
lockstep pattern, tests/cudaTest.cpp:50-69); from there every quantity must be
bit-identical: FP64 walk state, sort order, NodeData prefixes, sweep cuts,
labels, internal degrees, iteration counts and the final partition.
"""
import numpy as np
import pytest

import parex_amd as px
from oracle_ctypes import Oracle, init_vector

pytestmark = pytest.mark.gpu


def graph_cases():
    return {
        "sbm_c1": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
        "torus32": lambda: px.torus(32, 10, 1),
        "rmat10": lambda: px.rmat(10, 8, 0.57, 0.19, 0.19, 7),
        "path300": lambda: px.from_edges(300, np.arange(299), np.arange(1, 300)),
        "islands": lambda: px.from_edges(40, [0, 1, 2, 5, 6, 7, 20, 21, 22, 22],
                                         [1, 2, 0, 6, 7, 5, 21, 22, 23, 24]),
        "sbm_small_seed9": lambda: px.sbm(600, 5, 0.08, 0.004, 9),
        # mean degree 39 / 82 / 229: 8, 16 and 64 lanes per node
        "sbm_dense_deg39": lambda: px.sbm(800, 4, 0.17, 0.01, 3),
        "sbm_dense_deg82": lambda: px.sbm(1000, 4, 0.3, 0.01, 5),
        "sbm_dense_deg229": lambda: px.sbm(500, 2, 0.9, 0.02, 7),
        # one row of 2999 arcs on a ring: long rows under 4 lanes per node
        "hub3000": lambda: hub(3000),
    }


def hub(n):
    """Node 0 joined to every other node, which also form a ring."""
    fx = [0] * (n - 1) + list(range(1, n))
    tx = list(range(1, n)) + [i % (n - 1) + 1 for i in range(2, n + 1)]
    return px.from_edges(n, fx, tx)


def _engine(g, **kw):
    dm = px.DeviceManager(px.Params(**kw))
    dm.initialize(g)
    return dm


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_init_vector_matches_oracle_within_ulps():
    g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    dm = _engine(g)
    dev = dm.readRandomWalkValuesF64().astype(np.float32)
    ref = init_vector(g.degree(), 6582).astype(np.float32)
    # HIP erfcinvf vs the oracle's FP64 Newton erfcinv rounded to float
    ulp = np.abs(dev.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    assert ulp.max() <= 4, f"max ulp distance {ulp.max()}"
    # moments: N(0, sqrt(deg)) -> x/sqrt(deg) ~ N(0,1)
    z = dev[g.degree() > 0] / np.sqrt(g.degree()[g.degree() > 0])
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1.0) < 0.1


@pytest.mark.parametrize("name", list(graph_cases()))
def test_walk_lockstep(name):
    g = graph_cases()[name]()
    dm = _engine(g)
    x0 = dm.readRandomWalkValuesF64()
    ox = Oracle(g, init_values=x0)
    for step in range(16):
        dm.iterateRandomWalk()
        ox.walk()
        a, b = dm.readRandomWalkValuesF64(), ox.values()
        bad = np.nonzero(_bits(a) != _bits(b))[0]
        assert bad.size == 0, f"{name} step {step}: {bad.size} values differ, first {bad[:5]} {a[bad[:3]]} {b[bad[:3]]}"


WALK_VARIANT_GRAPHS = {
    "sbm_c1": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
    "hub3000": lambda: hub(3000),
    # neighbours spread over the whole index range: gathers outside every window
    "rmat14": lambda: px.rmat(14, 16, 0.57, 0.19, 0.19, 7),
    # 30,000 nodes in blocks of 3,000: windows that cover a block, chunks that span two
    "sbm30k": lambda: px.sbm(30000, 10, 8 / 2999, 1 / 27000, 3),
    # above every window size: keys and roots outside the 4-byte windows too
    "sbm100k": lambda: px.sbm(100000, 40, 12 / 2499, 1 / 97500, 5),
    "rmat17": lambda: px.rmat(17, 8, 0.57, 0.19, 0.19, 7),
}


@pytest.mark.parametrize("variant", range(16))
def test_walk_variants_lockstep(variant):
    """Every kernel variant of the walk and the cut contribution -- 2..64
    lanes per row, the LDS-window kernels -- gives the oracle's bits (walk
    values, NodeData prefixes, sweep cuts, labels), through cuts, CC sampling
    (the window form when a window variant is pinned) and arc disabling."""
    for name, make in WALK_VARIANT_GRAPHS.items():
        g = make()
        dm = _engine(g)
        dm.setWalkVariant(variant)
        ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
        for it in range(1, 9):
            if ox.counts()[1] == 0:
                break
            dm.iterateRandomWalk(); ox.walk()
            a, b = dm.readRandomWalkValuesF64(), ox.values()
            bad = np.nonzero(_bits(a) != _bits(b))[0]
            assert bad.size == 0, f"{name} variant {variant} iter {it}: {bad.size} values differ, first {bad[:5]}"
            dm.computeSweepCuts(); ox.sweep()
            _compare_sweep(dm, ox, f"{name} variant {variant} iter {it} sweep")
            dm.cutClusters(); ox.cut()
            assert np.array_equal(dm.downloadLabels(), ox.labels()), f"{name} variant {variant} iter {it}: labels"
        dm.close()


def test_walk_tuning_runs_and_keeps_results():
    """Online tuning on a list of 2^17 nodes: a table entry is chosen, and the
    decomposition equals the one under the degree rule."""
    g = px.sbm(150000, 30, 14 / 4999, 1 / 145000, 17)
    dm = _engine(g, record_node_data=False)
    it = dm.expanderDecomposition()
    table, ctable, tunings = dm.walkTuning()
    assert tunings >= 2 and 17 in table and 0 <= table[17] <= 15 and 17 in ctable and 0 <= ctable[17] <= 15
    ids, k = dm.getFinalPartition()
    dm.reset()
    dm.setWalkVariant(-2)
    assert dm.expanderDecomposition() == it
    ids2, k2 = dm.getFinalPartition()
    assert k2 == k and np.array_equal(ids2, ids)


def _compare_sweep(dm, ox, tag):
    na, nc, ncut = dm.counts()
    ona, onc, oncut = ox.counts()
    assert (na, nc, ncut) == (ona, onc, oncut), f"{tag}: counts {na, nc, ncut} vs oracle {ona, onc, oncut}"
    act, oact = dm.downloadActiveNodes(), ox.active()
    assert np.array_equal(act["nix"], oact["nix"]), f"{tag}: sorted order differs at {np.nonzero(act['nix'] != oact['nix'])[0][:5]}"
    assert np.array_equal(act["clusterId"], oact["clusterId"]), f"{tag}: active labels differ"
    nd, ond = dm.downloadPartition(), ox.ndata()
    for f in ("nix", "prefixVolume", "prefixEdgeDiff", "offsetInCluster"):
        assert np.array_equal(nd[f], ond[f]), f"{tag}: NodeData.{f} differs at {np.nonzero(nd[f] != ond[f])[0][:5]}"
    ids, cuts = dm.readSweepCuts()
    ocuts = ox.sweepcuts()
    assert np.array_equal(ids, ocuts["clusterId"]), f"{tag}: sweep-cut labels differ"
    assert np.array_equal(cuts["sparsity"].view(np.uint32), ocuts["sparsity"].view(np.uint32)), \
        f"{tag}: sparsity differs at {np.nonzero(cuts['sparsity'] != ocuts['sparsity'])[0][:5]}"
    assert np.array_equal(cuts["offset"], ocuts["offset"]), f"{tag}: cut offsets differ"


def _compare_cut(dm, ox, tag):
    assert np.array_equal(dm.downloadLabels(), ox.labels()), \
        f"{tag}: labels differ at {np.nonzero(dm.downloadLabels() != ox.labels())[0][:8]}"
    a, b = dm.readRandomWalkValuesF64(), ox.values()
    assert np.array_equal(_bits(a), _bits(b)), f"{tag}: values differ at {np.nonzero(_bits(a) != _bits(b))[0][:5]}"
    assert np.array_equal(dm.downloadDegrees(), ox.internal_degrees()), f"{tag}: internal degrees differ"
    assert dm.counts()[1] == ox.counts()[1], f"{tag}: active clusters {dm.counts()[1]} vs {ox.counts()[1]}"
    oact = ox.active()
    oact = oact[oact["clusterId"] >= 0]
    act = dm.downloadActiveNodes()
    assert np.array_equal(act["nix"], oact["nix"]) and np.array_equal(act["clusterId"], oact["clusterId"]), \
        f"{tag}: active list after cut differs"


@pytest.mark.parametrize("name", list(graph_cases()))
def test_iteration_lockstep(name):
    g = graph_cases()[name]()
    dm = _engine(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    for it in range(1, 61):
        if ox.counts()[1] == 0:
            break
        dm.iterateRandomWalk(); ox.walk()
        dm.computeSweepCuts(); ox.sweep()
        _compare_sweep(dm, ox, f"{name} iter {it} sweep")
        dm.cutClusters(); ox.cut()
        _compare_cut(dm, ox, f"{name} iter {it} cut")
    assert ox.errors() == 0
    assert dm.getNumCutEdges() == ox.num_cut_edges()
    ag = dm.downloadActiveEdgeMap()
    assert np.array_equal(ag, (ox.neighbors() != px.INVALID_EDGE).astype(np.int32))
    assert np.array_equal(dm.downloadGraph().edges, ox.neighbors())


@pytest.mark.parametrize("name", list(graph_cases()))
def test_end_to_end(name):
    g = graph_cases()[name]()
    dm = _engine(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    it = dm.expanderDecomposition()
    oit = ox.run()
    assert it == oit
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)
    cut, vol, ivol, cond = px.verify_partition(g, ids, k)
    assert vol.sum() == 2 * g.m and ivol.sum() + 2 * cut == 2 * g.m


def test_end_to_end_sbm_50k():
    """50,000 nodes, 20 blocks: clusters far above one workgroup's sort tile."""
    g = px.sbm(50000, 20, 16 / 2499, 1 / 47500, 21)
    dm = _engine(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert dm.expanderDecomposition() == ox.run()
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)


def _end_to_end_exact(g):
    dm = _engine(g, keep_original=False, record_node_data=False)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert dm.expanderDecomposition() == ox.run()
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)
    return ids, k


def test_torus_1024_full_size_matches_oracle():
    """The 1024 x 1024 torus with 10,486 random long-range edges (seed 1) at full size."""
    g = px.torus(1024, 10486, 1)
    ids, k = _end_to_end_exact(g)
    cut, vol, ivol, cond = px.verify_partition(g, ids, k)
    assert vol.sum() == 2 * g.m and ivol.sum() + 2 * cut == 2 * g.m


def test_rmat16_matches_oracle():
    """RMAT scale 16, edge factor 16 (.57, .19, .19): hub rows and many isolated nodes."""
    _end_to_end_exact(px.rmat(16, 16, 0.57, 0.19, 0.19, 7))


def test_rmat22_full_size_properties():
    """RMAT scale 22, edge factor 16, seed 7 (64M edges) at full size, where the
    oracle takes minutes: a valid partition re-verified on the host, the same
    result on a second run, and the same result inside a batch."""
    g = px.rmat(22, 16, 0.57, 0.19, 0.19, 7)
    dm = _engine(g, keep_original=True, record_node_data=False)
    it = dm.expanderDecomposition()
    ids, k = dm.getFinalPartition()
    assert ids.min() == 0 and ids.max() == k - 1 and len(np.unique(ids)) == k
    cut, vol, ivol, cond = px.verify_partition(g, ids, k)
    assert vol.sum() == 2 * g.m and ivol.sum() + 2 * cut == 2 * g.m
    dm.reset()
    assert dm.expanderDecomposition() == it
    ids2, k2 = dm.getFinalPartition()
    assert k2 == k and np.array_equal(ids2, ids)
    dm.close()
    small = px.sbm(600, 5, 0.08, 0.004, 9)
    bm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    bm.initialize_batch([small, g])
    bm.expanderDecomposition()
    bids, bks = bm.getFinalPartition()
    assert bm.graphIterations()[1] == it and bks[1] == k and np.array_equal(bids[small.n:], ids)


def test_batch_mixed_density_equals_solo():
    """A batch picks its lanes per node from the batch's mean degree, a solo
    run from its own: the exact walk sum makes both give the same bits."""
    graphs = [px.sbm(500, 2, 0.9, 0.02, 7), px.sbm(600, 5, 0.08, 0.004, 9), hub(500),
              px.sbm(1000, 4, 0.3, 0.01, 5)]
    dm = px.DeviceManager()
    dm.initialize_batch(graphs)
    dm.expanderDecomposition()
    ids, ks = dm.getFinalPartition()
    its = dm.graphIterations()
    off = 0
    for i, g in enumerate(graphs):
        solo = _engine(g)
        assert its[i] == solo.expanderDecomposition()
        sids, sk = solo.getFinalPartition()
        assert ks[i] == sk and np.array_equal(ids[off:off + g.n], sids), f"graph {i}: partition differs"
        off += g.n


def test_cli_matches_oracle(tmp_path):
    """bin/parex prints the reference's lines (main.cpp:15-64) for the engine's result."""
    import os
    import subprocess
    g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
    path = str(tmp_path / "c1.pxg")
    px.write_pxg(path, g)
    cli = os.path.join(os.path.dirname(px.__file__), "bin", "parex")
    r = subprocess.run([cli, path, "6582", "0.3", "1e-4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = dict(l.split(": ", 1) for l in r.stdout.splitlines() if ": " in l)
    assert out["sparsity"] == "0.300000" and out["threshold"] == "0.000100"
    assert "Begin Expander Decomposition" in r.stdout
    dm = _engine(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    it = ox.run()
    oids, ok = ox.final_partition()
    cut, _, _, _ = px.verify_partition(g, oids, ok)
    assert int(out["iterations"]) == it
    assert (int(out["numClusters"]), int(out["cutEdges"])) == (ok, cut)
    assert (int(out["nodes"]), int(out["edges"])) == (g.n, g.m)


def test_batch_equals_solo():
    graphs = [px.sbm(300 + 37 * i, 3 + i % 4, 0.06, 0.003, 100 + i) for i in range(7)]
    graphs.insert(3, px.from_edges(50, np.arange(49), np.arange(1, 50)))
    dm = px.DeviceManager()
    dm.initialize_batch(graphs)
    dm.expanderDecomposition()
    ids, ks = dm.getFinalPartition()
    its = dm.graphIterations()
    off = 0
    for i, g in enumerate(graphs):
        solo = _engine(g)
        sit = solo.expanderDecomposition()
        sids, sk = solo.getFinalPartition()
        assert its[i] == sit, f"graph {i}: iterations {its[i]} vs solo {sit}"
        assert ks[i] == sk and np.array_equal(ids[off:off + g.n], sids), f"graph {i}: partition differs"
        off += g.n


def test_initialize_checks_arcs_on_the_device():
    """Arcs are uploaded as they are and checked (and rebased) on the device:
    an arc past n is refused with the graph and arc named, and the engine
    takes a valid input afterwards."""
    g = px.sbm(600, 5, 0.08, 0.004, 9)
    edges = g.edges.copy()
    edges[17] = g.n + 3
    dm = px.DeviceManager()
    with pytest.raises(px.PxdError, match="graph 1: arc 17 points to"):
        dm.initialize_batch([px.sbm(300, 3, 0.06, 0.003, 100), px.Graph(g.n, g.m, g.ranges.copy(), edges)])
    dm.initialize(g)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert dm.expanderDecomposition() == ox.run()
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)


def _edge_list(g):
    """Each undirected edge once (u < v for the generators' simple graphs)."""
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.ranges))
    keep = src < g.edges
    return src[keep], g.edges[keep]


@pytest.mark.parametrize("name", ["sbm_c1", "rmat14", "hub", "loops_dups", "empty", "isolated_tail"])
def test_initialize_edges_builds_the_host_csr(name):
    """The device CSR build gives the host build's ranges and arcs
    (lib/core/graph.h:147-184: both arcs of every edge, rows in target
    order, loops and duplicates kept), and the same decomposition."""
    if name == "sbm_c1":
        g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
        fx, tx = _edge_list(g)
    elif name == "rmat14":
        g = px.rmat(14, 16, 0.57, 0.19, 0.19, 7)
        fx, tx = _edge_list(g)
    elif name == "hub":
        g = hub(3000)
        fx, tx = _edge_list(g)
    elif name == "loops_dups":
        rng = np.random.default_rng(5)
        fx = rng.integers(0, 300, 2000).astype(np.uint32)
        tx = rng.integers(0, 300, 2000).astype(np.uint32)
        fx[:10] = tx[:10]                       # loops
        fx[10:20], tx[10:20] = fx[20:30], tx[20:30]   # duplicates
        g = px.from_edges(300, fx, tx)
    elif name == "empty":
        fx = tx = np.zeros(0, np.uint32)
        g = px.from_edges(50, fx, tx)
    else:
        fx, tx = np.arange(99, dtype=np.uint32), np.arange(1, 100, dtype=np.uint32)
        g = px.from_edges(130, fx, tx)       # nodes 100..129 have no arcs
    dm = px.DeviceManager()
    dm.initialize_edges(g.n, fx, tx)
    d = dm.downloadGraph()
    assert np.array_equal(d.ranges, g.ranges) and np.array_equal(d.edges, g.edges)
    it = dm.expanderDecomposition()
    ids, k = dm.getFinalPartition()
    ref = _engine(g)
    assert ref.expanderDecomposition() == it
    rids, rk = ref.getFinalPartition()
    assert rk == k and np.array_equal(rids, ids)


def test_initialize_edges_refuses_an_end_past_n():
    dm = px.DeviceManager()
    with pytest.raises(px.PxdError, match="edge 3 .* has an end >= n"):
        dm.initialize_edges(10, [0, 1, 2, 3], [1, 2, 3, 10])
    dm.initialize_edges(10, [0, 1, 2, 3], [1, 2, 3, 9])
    assert dm.expanderDecomposition() >= 1


def _random_graph(rng):
    kind = rng.integers(0, 6)
    if kind == 0:
        n = int(rng.integers(200, 20000))
        k = int(rng.integers(2, 40))
        din, dout = rng.uniform(3, 30), rng.uniform(0.2, 3)
        bs = max(n // k, 2)
        return px.sbm(n, k, min(din / (bs - 1), 1.0), dout / max(n - bs, 1), int(rng.integers(1 << 30)))
    if kind == 1:
        return px.rmat(int(rng.integers(8, 14)), int(rng.integers(2, 17)), 0.57, 0.19, 0.19, int(rng.integers(1 << 30)))
    if kind == 2:
        L = int(rng.integers(8, 120))
        return px.torus(L, int(rng.integers(0, L)), int(rng.integers(1 << 30)))
    if kind == 3:                                   # random sparse graph, isolated nodes included
        n = int(rng.integers(50, 5000))
        m = int(rng.integers(0, 3 * n))
        fx, tx = rng.integers(0, n, m), rng.integers(0, n, m)
        keep = fx != tx
        pairs = np.unique(np.sort(np.stack([fx[keep], tx[keep]]), axis=0), axis=1)
        return px.from_edges(n, pairs[0], pairs[1])
    if kind == 4:                                   # star plus path: one long row, long chains
        n = int(rng.integers(300, 3000))
        fx = list(range(1, n // 2)) + list(range(n // 2, n - 1))
        tx = [0] * (n // 2 - 1) + list(range(n // 2 + 1, n))
        return px.from_edges(n, fx, tx)
    n = int(rng.integers(4, 40))                    # a clique among isolated nodes
    iu = np.triu_indices(n, 1)
    return px.from_edges(n + 10, iu[0], iu[1])


@pytest.mark.parametrize("seed", range(16))
def test_random_graphs_match_oracle(seed):
    """Random graph families (planted partitions, RMAT, tori, sparse random
    graphs with isolated nodes, a star joined to a path, a clique), end to end:
    iterations, cluster ids and their count equal the oracle's."""
    rng = np.random.default_rng(1000 + seed)
    g = _random_graph(rng)
    dm = _engine(g, record_node_data=False)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert dm.expanderDecomposition() == ox.run()
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids), f"seed {seed}: n={g.n} m={g.m}"
    assert ox.errors() == 0


def test_tuned_engine_matches_oracle_on_a_large_sbm():
    """200,000 nodes: lists long enough for the online tuning of both per-arc
    passes (every length class from 2^16 up); the result equals the oracle's."""
    g = px.sbm(200000, 40, 16 / 4999, 1.5 / 195000, 29)
    dm = _engine(g, keep_original=False, record_node_data=False)
    ox = Oracle(g, init_values=dm.readRandomWalkValuesF64())
    assert dm.expanderDecomposition() == ox.run()
    table, ctable, tunings = dm.walkTuning()
    assert tunings >= 2 and 17 in table
    ids, k = dm.getFinalPartition()
    oids, ok = ox.final_partition()
    assert k == ok and np.array_equal(ids, oids)


def test_reinitialized_engine_equals_fresh_engines():
    """One engine initialized again and again (its buffers are kept and handed
    back to the next input): every result equals a fresh engine's, through a
    larger input after smaller ones, a batch and an edge-list upload."""
    seq = [px.sbm(30000, 12, 20 / 2499, 1 / 27500, 31), px.torus(64, 40, 3), px.rmat(14, 8, 0.57, 0.19, 0.19, 5),
           px.sbm(60000, 20, 16 / 2999, 1 / 57000, 32)]
    dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    for g in seq:
        dm.initialize(g)
        it = dm.expanderDecomposition()
        ids, k = dm.getFinalPartition()
        fresh = _engine(g, keep_original=False, record_node_data=False)
        assert fresh.expanderDecomposition() == it
        fids, fk = fresh.getFinalPartition()
        assert fk == k and np.array_equal(fids, ids)
        fresh.close()
    dm.initialize_batch(seq[:3])
    dm.expanderDecomposition()
    bids, bk = dm.getFinalPartition()
    off = 0
    for i, g in enumerate(seq[:3]):
        fresh = _engine(g, keep_original=False, record_node_data=False)
        fresh.expanderDecomposition()
        fids, fk = fresh.getFinalPartition()
        assert fk == bk[i] and np.array_equal(fids, bids[off:off + g.n])
        off += g.n
        fresh.close()
    g = seq[1]
    src = np.repeat(np.arange(g.n), np.diff(g.ranges))
    keep = src < g.edges
    dm.initialize_edges(g.n, src[keep], g.edges[keep])
    dm.expanderDecomposition()
    eids, ek = dm.getFinalPartition()
    fresh = _engine(g, keep_original=False, record_node_data=False)
    fresh.expanderDecomposition()
    fids, fk = fresh.getFinalPartition()
    assert ek == fk and np.array_equal(eids, fids)
    dm.close()
