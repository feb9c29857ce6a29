"""pxd_csr_from_edges_noalloc -- the device CSR build of pxd_initialize_edges
with the library's static work area (no allocation) -- through the C ABI:
its CSR equals the one pxd_initialize_edges builds (read back with
pxd_download_graph), on calls in a row of decreasing size that share the
work area, with loops, duplicate edges, a row above one workgroup's tile and
trailing isolated nodes; an end past n sets the caller's flag; a graph above
the stated maximum is refused before any launch."""
import numpy as np
import pytest
import torch

import parex_amd as px

pytestmark = pytest.mark.gpu


def _edge_cases():
    rng = np.random.default_rng(21)
    g = px.rmat(17, 8, 0.57, 0.19, 0.19, 3)            # hub rows above 2048 arcs
    src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.ranges))
    keep = src < g.edges
    big = (g.n, src[keep], g.edges[keep])
    n = 5000                                           # loops, duplicates, isolated tail
    fx = rng.integers(0, 4000, 30000).astype(np.uint32)
    tx = rng.integers(0, 4000, 30000).astype(np.uint32)
    fx[:50] = tx[:50]
    fx[50:100], tx[50:100] = fx[100:150], tx[100:150]
    small = (n, fx, tx)
    return big, small


def _noalloc_csr(n, fx, tx):
    dev = torch.device("cuda:0")
    dfx = torch.from_numpy(fx.view(np.int32)).to(dev)
    dtx = torch.from_numpy(tx.view(np.int32)).to(dev)
    rg = torch.full((n + 1,), -1, dtype=torch.int32, device=dev)
    ed = torch.full((max(2 * len(fx), 1),), -1, dtype=torch.int32, device=dev)
    bad = torch.full((1,), 7, dtype=torch.int32, device=dev)
    st = torch.cuda.current_stream(dev)
    px.csr_from_edges_noalloc(n, len(fx), dfx.data_ptr(), dtx.data_ptr(), rg.data_ptr(), ed.data_ptr(), bad.data_ptr(),
                              st.cuda_stream)
    torch.cuda.synchronize(dev)
    return (rg.cpu().numpy().view(np.uint32), ed.cpu().numpy().view(np.uint32)[:2 * len(fx)],
            int(bad.cpu().numpy()[0]))


def _engine_csr(n, fx, tx):
    dm = px.DeviceManager()
    dm.initialize_edges(n, fx, tx)
    g = dm.downloadGraph()
    dm.close()
    return g.ranges, g.edges


def test_noalloc_csr_equals_initialize_edges_on_calls_in_a_row():
    big, small = _edge_cases()
    assert big[0] <= px.NOALLOC_MAX_NODES and len(big[1]) <= px.NOALLOC_MAX_EDGES
    print(f"noalloc: big n={big[0]} m={len(big[1])}, small n={small[0]} m={len(small[1])}", flush=True)
    for n, fx, tx in (big, small, big):
        r, e, bad = _noalloc_csr(n, fx, tx)
        wr, we = _engine_csr(n, fx, tx)
        assert bad == 0
        assert np.array_equal(r, wr) and np.array_equal(e, we), f"n={n}: CSR differs from initialize_edges"
        # and the host build (lib/core/graph.h:147-184)
        g = px.from_edges(n, fx, tx)
        assert np.array_equal(r, g.ranges) and np.array_equal(e, g.edges)


def test_noalloc_flags_an_end_past_n_and_refuses_oversize():
    fx = np.array([0, 1, 2, 3], np.uint32)
    tx = np.array([1, 2, 3, 10], np.uint32)
    _, _, bad = _noalloc_csr(10, fx, tx)
    assert bad != 0
    _, _, bad = _noalloc_csr(11, fx, tx)
    assert bad == 0
    with pytest.raises(px.PxdError, match="at most"):
        px.csr_from_edges_noalloc(px.NOALLOC_MAX_NODES + 1, 0, 0, 0, 1, 0, 1)
    with pytest.raises(px.PxdError, match="at most"):
        px.csr_from_edges_noalloc(10, px.NOALLOC_MAX_EDGES + 1, 1, 1, 1, 1, 1)
    with pytest.raises(px.PxdError, match="null"):
        px.csr_from_edges_noalloc(10, 4, 0, 0, 0, 0, 0)
