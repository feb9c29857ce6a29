"""GPU: continuous batching.  Graphs join a running engine between iterations
and leave when they finish; each result must be bit-identical to the graph's
solo decomposition (cluster ids, cluster count, iteration count)."""
import numpy as np
import pytest

import parex_amd as px

pytestmark = pytest.mark.gpu


def hub(n):
    fx = [0] * (n - 1) + list(range(1, n))
    tx = list(range(1, n)) + [i % (n - 1) + 1 for i in range(2, n + 1)]
    return px.from_edges(n, fx, tx)


def mix():
    return [
        px.sbm(600, 5, 0.08, 0.004, 9),
        px.torus(32, 10, 1),
        px.rmat(10, 8, 0.57, 0.19, 0.19, 7),
        hub(500),
        px.sbm(800, 4, 0.17, 0.01, 3),
        px.from_edges(300, np.arange(299), np.arange(1, 300)),
        px.from_edges(40, [0, 1, 2, 5, 6, 7, 20, 21, 22, 22], [1, 2, 0, 6, 7, 5, 21, 22, 23, 24]),
        px.from_edges(7, [], []),                       # isolated nodes only
        px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
    ]


def solo(g):
    dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    dm.initialize(g)
    dm.expanderDecomposition()
    ids, k = dm.getFinalPartition()
    it = int(dm.graphIterations()[0])
    dm.close()
    return ids, k, it


def drain(srv, pending, graphs, got, max_steps=100000):
    """Admit from `pending` (FIFO) whenever room allows, step until all done."""
    refused = 0
    steps = 0
    while pending or srv.stats()["live_graphs"]:
        while pending:
            i = pending[0]
            if not srv.admit(graphs[i], i):
                refused += 1
                break
            pending.pop(0)
        srv.step()
        steps += 1
        for tag, ids, k, it in srv.results():
            got[tag] = (ids, k, it)
        assert steps < max_steps
    for tag, ids, k, it in srv.results(wait=True):
        got[tag] = (ids, k, it)
    return refused


def check(graphs, got):
    assert sorted(got) == list(range(len(graphs)))
    for i, g in enumerate(graphs):
        ids, k, it = got[i]
        sids, sk, sit = solo(g)
        assert it == sit, f"graph {i}: iterations {it} vs solo {sit}"
        assert k == sk and np.array_equal(ids, sids), f"graph {i}: partition differs from solo"


def test_serve_staggered_arrivals_equal_solo():
    graphs = mix()
    srv = px.Server(node_capacity=1 << 16, arc_capacity=1 << 20, graph_slots=16)
    got = {}
    for i in range(4):                                  # the first wave
        assert srv.admit(graphs[i], i)
    for _ in range(3):                                  # in flight before the rest arrive
        srv.step()
        for tag, ids, k, it in srv.results():
            got[tag] = (ids, k, it)
    for i in range(4, len(graphs)):
        assert srv.admit(graphs[i], i)
        srv.step()
        for tag, ids, k, it in srv.results():
            got[tag] = (ids, k, it)
    drain(srv, [], graphs, got)
    srv.close()
    check(graphs, got)


def test_serve_capacity_pressure_reuses_ranges():
    """Capacities below the workload: admissions wait for finished graphs and
    reuse their node ranges, arc ranges and slots."""
    graphs = mix() + [px.sbm(700, 3, 0.05, 0.004, 10 + i) for i in range(6)]
    total_nodes = sum(g.n for g in graphs)
    srv = px.Server(node_capacity=5000, arc_capacity=300000, graph_slots=3)
    got = {}
    refused = drain(srv, list(range(len(graphs))), graphs, got)
    assert refused > 0 and total_nodes > 5000
    st = srv.stats()
    assert st["live_graphs"] == 0 and st["live_nodes"] == 0 and st["live_arcs"] == 0
    srv.close()
    check(graphs, got)


def test_serve_rejects_batch_calls_and_oversized_graphs():
    srv = px.Server(node_capacity=100, arc_capacity=1000, graph_slots=2)
    with pytest.raises(px.PxdError):
        srv.admit(px.sbm(600, 5, 0.08, 0.004, 9), 0)      # larger than the capacity
    assert srv.admit(px.from_edges(0, [], []), 5)         # empty graph: result at once
    res = srv.results(wait=True)
    assert len(res) == 1 and res[0][0] == 5 and res[0][2] == 0 and res[0][3] == 1
    srv.close()


def test_serve_refuses_invalid_graph_and_carries_on():
    """Arcs outside the graph or decreasing ranges are caught on the device
    before the graph joins: the result says so (num_clusters = -1), its room
    is returned and the server goes on serving."""
    good = px.sbm(300, 3, 0.1, 0.01, 4)
    edges = good.edges.copy()
    edges[17] = good.n                                   # one arc past the last node
    bad = px.Graph(good.n, good.m, good.ranges.copy(), edges)
    ranges = good.ranges.copy()
    ranges[5], ranges[6] = ranges[6], ranges[5]          # decreasing ranges
    bad2 = px.Graph(good.n, good.m, ranges, good.edges.copy())
    srv = px.Server(node_capacity=400, arc_capacity=good.edges.size, graph_slots=1)
    for tag, g in ((7, bad), (8, bad2)):
        assert srv.admit(g, tag)
        srv.step()
        res = srv.results(wait=True)
        assert [(r[0], r[2]) for r in res] == [(tag, -1)]
        assert srv.stats()["live_graphs"] == 0
    with pytest.raises(px.PxdError):                        # host-side check: ranges[n] != 2m
        srv.admit(px.Graph(good.n, good.m + 1, good.ranges.copy(), good.edges.copy()), 9)
    got = {}
    drain(srv, [0], [good], got)
    srv.close()
    check([good], got)


def _edges(g):
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.ranges))
    keep = src < g.edges
    return src[keep], g.edges[keep]


def test_serve_edge_lists_equal_solo_and_csr_admissions():
    """Edge-list admissions (CSR built on the device, no host sync per
    admission) mixed with CSR admissions, joining while others run: every
    result equals the graph's solo run; an edge past n fails that graph only."""
    graphs = mix()
    ref = [solo(g) for g in graphs]
    srv = px.Server(node_capacity=20000, arc_capacity=400000, graph_slots=16,
                    params=px.Params(record_node_data=False))
    srv.enable_edges(5000, 60000)
    got = {}
    bad_fx, bad_tx = np.array([0, 1, 2], np.uint32), np.array([1, 2, 9], np.uint32)
    for rnd in range(2):
        for i, g in enumerate(graphs):
            tag = rnd * 100 + i
            if (i + rnd) % 2 == 0:
                assert srv.admit_edges(g.n, *_edges(g), tag)
            else:
                assert srv.admit(g, tag)
            if i == 3 and rnd == 0:
                assert srv.admit_edges(5, bad_fx, bad_tx, 999)
            srv.step()
            for t, ids, k, it in srv.results():
                got[t] = (ids, k, it)
    while srv.stats()["live_graphs"]:
        srv.step()
    for t, ids, k, it in srv.results(wait=True):
        got[t] = (ids, k, it)
    assert got[999][1] == -1
    for rnd in range(2):
        for i in range(len(graphs)):
            ids, k, it = got[rnd * 100 + i]
            rids, rk, rit = ref[i]
            assert k == rk and it == rit and np.array_equal(ids, rids), f"graph {i} round {rnd} differs from solo"
    with pytest.raises(px.PxdError, match="edge-admission limits"):
        srv.admit_edges(6000, [0], [1], 5)
    srv.close()


@pytest.mark.parametrize("variant", [1, 7, 9, 10, 11, 12, 13, 14, 15])
def test_serve_pinned_kernel_variants_equal_solo(variant):
    """A serving engine's node list is a concatenation of the admitted graphs'
    lists, not one list in index order: every pinned kernel variant still
    gives each graph its solo result (variants 11, 12, 14 and 15, which
    need the list in index order, run as variants 7 and 6 there)."""
    graphs = mix()
    ref = [solo(g) for g in graphs]
    srv = px.Server(node_capacity=30000, arc_capacity=400000, graph_slots=16,
                    params=px.Params(record_node_data=False))
    srv.set_walk_variant(variant)
    got = {}
    # later graphs take lower node ranges freed by earlier ones, so the list runs out of index order
    drain(srv, list(range(len(graphs)))[::-1], graphs, got)
    drain(srv, list(range(len(graphs))), graphs, got)
    for i in range(len(graphs)):
        ids, k, it = got[i]
        rids, rk, rit = ref[i]
        assert k == rk and it == rit and np.array_equal(ids, rids), f"variant {variant}: graph {i} differs from solo"
    srv.close()
