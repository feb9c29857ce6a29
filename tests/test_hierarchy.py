"""Expander hierarchy and normalized cut (libpxh.so, include/parex_amd/pxh.h).

CPU tests pin the contraction (lib/algorithms/expander_hierarchy.h:29-88) and
NormalizedCut::compute (lib/algorithms/normalized_cut.h:108-144) against
tests/golden/hierarchy_ref.npz, written by the reference's own code
(tests/golden/make_hierarchy_golden.py): the reference's level partitions are
fed in and every level array, every part's size, volume and internal volume
and the float objective must come out identical.  GPU tests check that
pxh_build's level partitions are the engine's decompositions, level by level,
as the oracle computes them.
"""
import ctypes as C
import os

import numpy as np
import pytest

import parex_amd as px
from oracle_ctypes import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hierarchy_ref.npz")
GRAPHS = ("csbm500", "csbm2000", "torus16", "path50")
KS = (2, 4, 8, 16)


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _graph(gold, name):
    r, e = gold[f"{name}_ranges"], gold[f"{name}_edges"]
    return px.Graph(len(r) - 1, len(e) // 2, r, e)


def _ref_hierarchy(gold, name):
    h = px.ExpanderHierarchy(_graph(gold, name))
    for li in range(int(gold[f"{name}_height"][0])):
        h.add_level(gold[f"{name}_L{li}_member_off"], gold[f"{name}_L{li}_members"])
    return h


def _part_stats(g, labels, k):
    deg = g.degree().astype(np.int64)
    src = np.repeat(np.arange(g.n), deg)
    size = np.bincount(labels, minlength=k)
    vol = np.bincount(labels, weights=deg, minlength=k).astype(np.int64)
    inner = np.bincount(labels[src], weights=(labels[src] == labels[g.edges]), minlength=k).astype(np.int64)
    return size, vol, inner


def test_libpxh_exports_every_declared_symbol():
    lib = C.CDLL(px.lib_path("libpxh.so"))
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "parex_amd", "pxh.h")).read()
    for s in px.PXH_SYMBOLS:
        assert s + "(" in header
        getattr(lib, s)


@pytest.mark.parametrize("name", GRAPHS)
def test_contraction_matches_reference_levels(gold, name):
    h = _ref_hierarchy(gold, name)
    assert h.height() == int(gold[f"{name}_height"][0])
    for li in range(h.height()):
        L = h.lvl(li)
        p = f"{name}_L{li}_"
        for f in ("volumes", "cut_sizes", "parents", "member_off", "members"):
            assert np.array_equal(getattr(L, f), gold[p + f]), f
        assert np.array_equal(L.graph.ranges, gold[p + "ranges"])
        ref_edges = gold[p + "edges"]
        # same arcs per row; rows come out ascending (the engine's CSR convention)
        for c in range(L.graph.n):
            a, b = L.graph.ranges[c], L.graph.ranges[c + 1]
            assert np.array_equal(L.graph.edges[a:b], np.sort(ref_edges[a:b]))


@pytest.mark.parametrize("name", GRAPHS)
@pytest.mark.parametrize("k", KS)
def test_normalized_cut_matches_reference(gold, name, k):
    g = _graph(gold, name)
    h = _ref_hierarchy(gold, name)
    labels, value = px.NormalizedCut(h).compute(k)
    size, vol, inner = _part_stats(g, labels, k)
    p = f"{name}_k{k}_"
    assert np.array_equal(size, gold[p + "nc_size"])
    assert np.array_equal(vol, gold[p + "nc_volume"])
    assert np.array_equal(inner, gold[p + "nc_internal"])
    assert np.float32(value) == gold[p + "nc_value"][0]
    # the reference's compute changes the hierarchy; this one does not
    again, value2 = px.NormalizedCut(h).compute(k)
    assert np.array_equal(again, labels) and value2 == value


def test_contraction_of_a_small_graph_by_hand():
    """Two triangles joined by one edge, and a pendant node on the second:
    clusters {0,1,2} and {3,4,5,6} contract to two nodes with one arc each."""
    g = px.from_edges(7, [0, 1, 0, 2, 3, 4, 3, 5], [1, 2, 2, 3, 4, 5, 5, 6])
    h = px.ExpanderHierarchy(g)
    h.add_level(np.array([0, 3, 7]), np.array([0, 1, 2, 3, 4, 5, 6]))
    L = h.lvl(0)
    assert L.graph.n == 2 and L.graph.m == 1
    assert L.graph.ranges.tolist() == [0, 1, 2] and L.graph.edges.tolist() == [1, 0]
    assert L.volumes.tolist() == [7, 9] and L.cut_sizes.tolist() == [1, 1]
    assert L.parents.tolist() == [0, 0, 0, 1, 1, 1, 1]
    h.add_level(np.array([0, 2]), np.array([1, 0]))
    top = h.lvl(1)
    assert top.graph.n == 1 and top.graph.m == 0 and top.volumes.tolist() == [16]
    labels, value = px.NormalizedCut(h).compute(2)
    # cutting either triangle side costs 1/7 + 1/9 against the root's 0 / 16
    assert sorted(np.bincount(labels).tolist()) == [3, 4]
    assert np.float32(value) == np.float32(np.float32(1) / np.float32(7)) + np.float32(np.float32(1) / np.float32(9))


def test_subtract_volume_and_add_cut_change_the_next_cut():
    """expander_hierarchy.h:154-159: the level arrays change, and the next
    normalized cut sees the change (it copies the arrays when it starts)."""
    g = px.from_edges(7, [0, 1, 0, 2, 3, 4, 3, 5], [1, 2, 2, 3, 4, 5, 5, 6])
    h = px.ExpanderHierarchy(g)
    h.add_level(np.array([0, 3, 7]), np.array([0, 1, 2, 3, 4, 5, 6]))
    h.add_level(np.array([0, 2]), np.array([0, 1]))
    # both clusters cost 1/7 + 1/9 (float sums in either order): the first wins
    labels, before = px.NormalizedCut(h).compute(2)
    assert labels.tolist() == [0, 0, 0, 1, 1, 1, 1]
    h.subtractVolume(0, 1, 2)
    h.addCut(0, 0, 5)
    L = h.lvl(0)
    assert L.volumes.tolist() == [7, 7] and L.cut_sizes.tolist() == [6, 1]
    with pytest.raises(px.PxdError):
        h.addCut(0, 2, 1)
    # now cluster 0 costs 6/7 + 6/9 and cluster 1 costs 1/7 + 1/9: cluster 1
    # becomes the first root in (level, node) order and takes label 0
    labels, after = px.NormalizedCut(h).compute(2)
    assert labels.tolist() == [1, 1, 1, 0, 0, 0, 0]
    assert after == before   # the objective is taken on the input graph


def test_add_level_rejects_a_non_partition():
    g = px.from_edges(4, [0, 1, 2], [1, 2, 3])
    h = px.ExpanderHierarchy(g)
    with pytest.raises(px.PxdError):
        h.add_level(np.array([0, 2, 4]), np.array([0, 1, 1, 3]))
    # the binding checks that members lists every node once before the call
    with pytest.raises(ValueError):
        h.add_level(np.array([0, 2, 3]), np.array([0, 1, 2]))
    # offsets that overshoot the node count and come back: refused before any
    # member is read (the library checks every offset first)
    with pytest.raises(px.PxdError, match="not ascending"):
        h.add_level(np.array([0, 7, 4]), np.array([0, 1, 2, 3]))
    lib, hh = h._lib, h._h
    off = np.array([0, 9, 4], np.uint32)
    assert lib.pxh_add_level(hh, 2, off, np.array([0, 1, 2, 3], np.uint32)) != 0
    assert h.height() == 0


def test_normalized_cut_needs_one_top_node():
    g = px.from_edges(4, [0, 2], [1, 3])
    h = px.ExpanderHierarchy(g)
    h.add_level(np.array([0, 2, 4]), np.array([0, 1, 2, 3]))
    with pytest.raises(px.PxdError):
        px.NormalizedCut(h).compute(2)


# --------------------------------------------------------------------------- GPU

def _oracle_levels(g, sc=0.2, rw=1e-5, max_retries=64):
    """pxh_build restated over the oracle's decomposition
    (lib/algorithms/expander_hierarchy.h:113-132): partitions per level.
    Needs the GPU for each level's initial walk vector."""
    parts = []
    n, ranges, edges = g.n, g.ranges, g.edges
    sc = np.float32(sc)
    retries = 0
    while n > 1 and ranges[-1] > 0:
        cur = px.Graph(n, int(ranges[-1]) // 2, ranges, edges)
        # the oracle starts from the engine's initial walk vector (its
        # erfcinvf differs from the oracle's by a few ulps)
        dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
        dm.initialize(cur)
        iv = dm.readRandomWalkValuesF64()
        dm.close()
        ox = Oracle(cur, sc=float(sc), rw=rw, init_values=iv)
        ox.run()
        ids, k = ox.final_partition()
        src = np.repeat(np.arange(n), np.diff(ranges))
        cut = int((ids[src] != ids[edges]).sum()) // 2
        if cut > 0.95 * cur.m:
            retries += 1
            assert retries <= max_retries
            sc = np.float32(sc * np.float32(0.8))
            continue
        retries = 0
        parts.append((ids.astype(np.uint32), k))
        # contract: rows ascending, one arc per cut arc
        cs = ids[src]
        keep = cs != ids[edges]
        order = np.lexsort((ids[edges][keep], cs[keep]))
        new_edges = ids[edges][keep][order].astype(np.uint32)
        new_ranges = np.zeros(k + 1, np.uint32)
        np.cumsum(np.bincount(cs[keep], minlength=k), out=new_ranges[1:])
        n, ranges, edges = k, new_ranges, new_edges
    return parts


HIER_GPU_GRAPHS = {
    "csbm2000": lambda: _graph_from_gold("csbm2000"),
    "torus32": lambda: px.torus(32, 20, 4),
    "sbm4096": lambda: px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
    "rmat12": lambda: px.rmat(12, 8, 0.57, 0.19, 0.19, 3),
}


def _graph_from_gold(name):
    with np.load(GOLD, allow_pickle=False) as z:
        r, e = z[f"{name}_ranges"], z[f"{name}_edges"]
    return px.Graph(len(r) - 1, len(e) // 2, r, e)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(HIER_GPU_GRAPHS))
def test_gpu_hierarchy_levels_are_the_oracle_decompositions(name):
    g = HIER_GPU_GRAPHS[name]()
    h = px.ExpanderHierarchy(g)
    h.build()
    parts = _oracle_levels(g)
    assert h.height() == len(parts)
    below = g.n
    for li, (ids, k) in enumerate(parts):
        L = h.lvl(li)
        assert L.graph.n == k and np.array_equal(L.parents, ids), li
        assert len(L.parents) == below
        below = k
    top = h.lvl(h.height() - 1).graph
    assert top.n <= 1 or top.m == 0
    # level volumes add up to the input volume at every level
    for li in range(h.height()):
        assert int(h.lvl(li).volumes.astype(np.int64).sum()) == 2 * g.m
    if top.n != 1:
        # a disconnected input ends with several top nodes and no edges; the
        # reference's cut needs one top node
        with pytest.raises(px.PxdError):
            px.NormalizedCut(h).compute(2)
    else:
        for k in (2, 8):
            labels, value = px.NormalizedCut(h).compute(k)
            size, vol, inner = _part_stats(g, labels, k)
            assert size.sum() == g.n and labels.min() >= 0 and labels.max() < k
            want = np.float32(0)
            for p in range(k):
                if vol[p]:
                    want = np.float32(want + np.float32(vol[p] - inner[p]) / np.float32(vol[p]))
            assert np.float32(value) == want


@pytest.mark.gpu
def test_gpu_hierarchy_of_the_torus_workload_matches_oracle():
    """The hierarchy bench workload at full size (torus 1024^2 + 0.5 %, 2.1 M
    edges, six levels): every level equals the oracle's decomposition."""
    g = px.torus(1024, 10486, 1)
    h = px.ExpanderHierarchy(g)
    h.build()
    parts = _oracle_levels(g)
    assert h.height() == len(parts)
    for li, (ids, k) in enumerate(parts):
        L = h.lvl(li)
        assert L.graph.n == k and np.array_equal(L.parents, ids), li
    assert h.lvl(h.height() - 1).graph.n == 1
    labels, value = px.NormalizedCut(h).compute(16)
    size, vol, inner = _part_stats(g, labels, 16)
    assert size.sum() == g.n and vol.sum() == 2 * g.m
