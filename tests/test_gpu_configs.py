"""The benchmark's batched and served workloads under -m gpu, against the oracle.

* The 512-graph planted-partition batch of the default benchmark (n
  log-uniform in [1K, 256K], k U{4..64}, intra degree U(8, 32), inter degree
  U(0.5, 3), seed 11) at full size in one engine: every graph's result is
  re-verified on the host (dense ids, a partition whose internal volume plus
  twice the cut is 2m), and every 32nd graph (16 graphs across the size
  range) equals the oracle's decomposition bit for bit, the oracle seeded
  with that graph's slice of the batch's initial vector.
* The serving run: the first 200 arrivals of the seed-13 trace (the
  benchmark's request pool and generators: SBM 10K-200K nodes, tori 64^2 -
  1024^2, RMAT scale 14-21), admitted four per step into a running Server so
  that graphs join while others are in flight; every served result (ids,
  cluster count, iterations) equals the oracle's decomposition of that graph.

The oracle runs for several graphs at once on Python threads (its ctypes
calls release the interpreter lock), each with a few OpenMP threads.
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import parex_amd as px
from oracle_ctypes import Oracle, set_threads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(msg):
    print(msg, flush=True)


def oracle_runs(jobs, threads_per_run=4):
    """jobs: (key, graph, initial vector); largest first over a few Python
    threads.  Returns key -> (iterations, ids, num_clusters, error flags)."""
    cpus = min(os.cpu_count() or 4, 16)
    workers = max(1, cpus // threads_per_run)

    def one(job):
        key, g, x0 = job
        set_threads(threads_per_run)
        ox = Oracle(g, init_values=x0)
        it = ox.run()
        ids, k = ox.final_partition()
        err = ox.errors()
        del ox
        return key, (it, ids, k, err)

    out = {}
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        for key, res in ex.map(one, sorted(jobs, key=lambda j: -j[1].m)):
            out[key] = res
    log(f"oracle: {len(jobs)} graphs, {sum(j[1].m for j in jobs)} edges in {time.perf_counter() - t0:.1f} s "
        f"({workers} x {threads_per_run} threads)")
    return out


@pytest.mark.timeout(900)
def test_sbm512_batch_full_size():
    t0 = time.perf_counter()
    specs = px.sbm_specs(512, 1024, 262144, 4, 64, (8.0, 32.0), (0.5, 3.0), 11)
    graphs = px.sbm_batch(specs)
    offs = np.concatenate([[0], np.cumsum([g.n for g in graphs])])
    log(f"512-SBM batch: {offs[-1]} nodes, {sum(g.m for g in graphs)} edges ({time.perf_counter() - t0:.1f} s)")
    dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    dm.initialize_batch(graphs)
    x0 = dm.readRandomWalkValuesF64()
    dm.expanderDecomposition()
    ids, ks = dm.getFinalPartition()
    its = dm.graphIterations()
    dm.close()
    log(f"decomposed: iterations up to {its.max()}, {int(np.sum(ks))} clusters")
    for i, g in enumerate(graphs):
        gi, k = ids[offs[i]:offs[i + 1]], int(ks[i])
        assert gi.min() == 0 and gi.max() == k - 1 and len(np.unique(gi)) == k, f"graph {i}: ids not dense"
        cut, vol, ivol, cond = px.verify_partition(g, gi, k)
        assert vol.sum() == 2 * g.m and ivol.sum() + 2 * cut == 2 * g.m, f"graph {i}: not a partition"
        assert its[i] >= 2
    picked = list(range(0, 512, 32))
    log(f"oracle graphs: n = {[graphs[i].n for i in picked]}")
    ref = oracle_runs([(i, graphs[i], x0[offs[i]:offs[i + 1]].copy()) for i in picked])
    for i in picked:
        it, oids, ok, err = ref[i]
        assert err == 0, f"graph {i}: oracle error flags {err}"
        assert its[i] == it, f"graph {i} (n={graphs[i].n}): iterations {its[i]} vs oracle {it}"
        assert ks[i] == ok and np.array_equal(ids[offs[i]:offs[i + 1]], oids), \
            f"graph {i} (n={graphs[i].n}): partition differs from the oracle"


@pytest.mark.timeout(900)
def test_serving_trace_prefix_matches_oracle():
    sys.path.insert(0, ROOT)
    import bench
    t0 = time.perf_counter()
    pool = bench.serve_pool()
    _, picks, _ = bench.serve_trace(pool, 200.0, 60.0, 13)
    picks = picks[:200]
    req = [pool[k][j] for k, j in picks]
    log(f"serving pool: {sum(len(v) for v in pool.values())} graphs ({time.perf_counter() - t0:.1f} s); "
        f"first 200 arrivals: {len(set(picks))} distinct graphs, {sum(g.m for g in req)} edges")
    srv = px.Server(node_capacity=48 << 20, arc_capacity=(1 << 32) - 2, graph_slots=4096,
                    params=px.Params(record_node_data=False))
    got = {}
    i, steps, most_live = 0, 0, 0
    while i < len(req) or srv.stats()["live_graphs"]:
        admitted = 0
        while i < len(req) and admitted < 4 and srv.admit(req[i], i):
            i += 1
            admitted += 1
        live = srv.stats()["live_graphs"]
        most_live = max(most_live, live)
        if live:
            srv.step()
            steps += 1
        for tag, ids, k, it in srv.results():
            got[tag] = (ids, k, it)
    for tag, ids, k, it in srv.results(wait=True):
        got[tag] = (ids, k, it)
    srv.close()
    log(f"served {len(got)} requests in {steps} steps, up to {most_live} graphs in flight")
    assert len(got) == len(req) and most_live > 4
    distinct = sorted(set(picks))
    jobs = []
    for key in distinct:
        g = pool[key[0]][key[1]]
        d = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
        d.initialize(g)
        jobs.append((key, g, d.readRandomWalkValuesF64()))
        d.close()
    ref = oracle_runs(jobs)
    for tag, key in enumerate(picks):
        ids, k, it = got[tag]
        oit, oids, ok, err = ref[key]
        assert err == 0, f"{key}: oracle error flags {err}"
        assert k == ok and it == oit and np.array_equal(ids, oids), \
            f"request {tag} ({key}, n={req[tag].n}): served result differs from the oracle " \
            f"(clusters {k} vs {ok}, iterations {it} vs {oit})"
