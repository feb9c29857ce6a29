"""Benchmark: expander decomposition on one MI355X, batched vs plain, and served.

A step is one complete decomposition of the workload's graph set (every graph
run to its final partition), timed as the reference's CLI times it
(main.cpp:42-48): upload of the host CSR (pinned host buffers) + initial
state + decomposition.  `value` is the rate of a stream of `--steps` such
batches on `--engines` engines (4), each with its own HIP stream and host
thread, so that one batch's upload and short late-iteration kernels overlap
the others' work (`--no-stream`: one engine, upload then decomposition,
which is `value_serial`); `value_resident` is the decomposition alone with the
input already in HBM.

Default workload (`--workload sbm512`): 512 independent planted-partition
graphs, n log-uniform in [1K, 256K], k ~ U{4..64}, intra degree ~ U(8, 32),
inter degree ~ U(0.5, 3), seed 11; parameters seed 6582, sweep threshold
0.3, walk threshold 1e-4.  The default run also serves a bounded prefix
(--serve-duration seconds, default 20) of the serving trace: Poisson arrivals
at 200 requests/s, 60 % SBM 10K-200K nodes, 30 % tori 64^2-1024^2, 10 % RMAT
scale 14-21 (seed 13), reporting p50/p99 request latency and completed edges/s.

Printed (one JSON line on stdout): batched edges/s as `value`, the plain
one-graph-at-a-time path, their ratio, the walk kernel's HBM roofline, a CPU
baseline (the reference's single-threaded CPU decomposition, oracle/_ref, on a
bounded sample of the same graphs) and the serving leg.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from collections import deque

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import parex_amd as px  # noqa: E402

METRIC = "edges decomposed/sec, batched vs plain (1 GPU) + p50/p99 request latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name, graphs_limit=None):
    if name == "sbm512":
        specs = px.sbm_specs(512, 1024, 262144, 4, 64, (8.0, 32.0), (0.5, 3.0), 11)
        if graphs_limit:
            specs = specs[:graphs_limit]
        gs = px.sbm_batch(specs)
        desc = {"workload": "512 planted-partition SBMs, n log-uniform [1K,256K], k U{4..64}, "
                            "d_in U(8,32), d_out U(0.5,3), block-diagonal batch",
                "graphs": len(gs), "seed": 11}
        return gs, desc
    if name == "torus1024":
        g = px.torus(1024, 10486, 1)
        return [g], {"workload": "torus 1024x1024 + 0.5% random long-range edges", "graphs": 1, "seed": 1}
    if name == "sbm4096":
        g = px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582)
        return [g], {"workload": "SBM n=4096, 16 blocks, p_in=12/255, p_out=1/3840", "graphs": 1, "seed": 6582}
    if name == "rmat22":
        g = px.rmat(22, 16, 0.57, 0.19, 0.19, 7)
        return [g], {"workload": "RMAT scale 22, edge factor 16 (.57,.19,.19,.05), symmetrised", "graphs": 1,
                     "seed": 7}
    raise SystemExit(f"unknown workload {name}")


def pinned_u32(a, keep):
    """A copy of a u32 array in pinned host memory (uploads are plain DMA)."""
    import torch
    t = torch.empty(max(a.size, 1), dtype=torch.int32, pin_memory=True)
    keep.append(t)
    v = t.numpy().view(np.uint32)[:a.size]
    v[:] = a
    return v


def pinned_graphs(graphs):
    """The graphs' CSR arrays in pinned host memory, as a client's receive
    buffers would hold them."""
    keep = []
    return [px.Graph(g.n, g.m, pinned_u32(g.ranges, keep), pinned_u32(g.edges, keep)) for g in graphs], keep


# --------------------------------------------------------------------------- serving

def serve_pool(seed=13):
    """Request graphs of the serving mix: SBM with 10K-200K nodes (n log-uniform,
    other parameters as the 512-graph batch), tori with side 64-1024, RMAT with
    scale 14-21 and edge factor 16.  Requests draw from a pool of distinct graphs
    (generating 12,000 fresh inputs at the offered rate is not possible on the host)."""
    rng = np.random.default_rng(seed)
    specs = px.sbm_specs(48, 10000, 200000, 4, 64, (8.0, 32.0), (0.5, 3.0), seed)
    sbm = px.sbm_batch(specs)
    tor = [px.torus(int(L), int(round(0.005 * 2 * L * L)), seed + i)
           for i, L in enumerate(rng.integers(64, 1025, 16))]
    rm = [px.rmat(int(sc), 16, 0.57, 0.19, 0.19, seed + i) for i, sc in enumerate(range(14, 22))]
    return {"sbm": sbm, "torus": tor, "rmat": rm}


def serve_trace(pool, rate=200.0, dur=60.0, seed=13):
    """Arrivals of the serving run: Poisson at `rate` per second for `dur`
    seconds; 60/30/10 % SBM/torus/RMAT, each request a uniform pick among
    the pool's graphs of its kind.  Returns the arrival times, the (kind,
    index) picks and the generator (its later draws pick the probes)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, int(rate * dur * 1.5) + 100)
    arr = np.cumsum(gaps)
    arr = arr[arr < dur]
    kinds = rng.choice(["sbm", "torus", "rmat"], size=len(arr), p=[0.6, 0.3, 0.1])
    picks = [(str(k), int(rng.integers(len(pool[k])))) for k in kinds]
    return arr, picks, rng


def pinned_pool(pool):
    keep, out = [], {}
    for kind, graphs in pool.items():
        out[kind] = [px.Graph(g.n, g.m, pinned_u32(g.ranges, keep), pinned_u32(g.edges, keep)) for g in graphs]
    return out, keep


def edge_lists(pool):
    """Each pool graph as an undirected edge list (every edge once) in pinned
    host memory: requests of clients that send edges, not a CSR."""
    keep, out = [], {}
    for kind, graphs in pool.items():
        out[kind] = []
        for g in graphs:
            src = np.repeat(np.arange(g.n, dtype=np.uint32), np.diff(g.ranges))
            sel = src < g.edges
            out[kind].append((g.n, pinned_u32(src[sel], keep), pinned_u32(g.edges[sel], keep)))
    return out, keep


def run_serving(dur, walk_variant=None, admit_kind="csr", check=True):
    """Serve the first `dur` seconds of the trace in real time.  Latency =
    time the client holds the result - arrival.  Returns a dict."""
    t = time.perf_counter()
    pool, _pinned = pinned_pool(serve_pool())
    epool, _epinned = edge_lists(pool) if admit_kind == "edges" else (None, None)
    log(f"serving pool: {sum(len(v) for v in pool.values())} graphs ({time.perf_counter() - t:.1f} s)")
    arr, picks, rng = serve_trace(pool, 200.0, dur, 13)
    req = [pool[k][i] for k, i in picks]
    offered = sum(g.m for g in req) / dur
    log(f"{len(req)} requests over {dur:.0f} s, offered {offered:.3e} edges/s")
    srv = px.Server(node_capacity=48 << 20, arc_capacity=(1 << 32) - 2, graph_slots=4096,
                    params=px.Params(record_node_data=False))
    if walk_variant is not None:
        srv.set_walk_variant(walk_variant)
    if epool is not None:
        # requests as edge lists: the CSR of each is built on the device at admission
        srv.enable_edges(max(g.n for v in pool.values() for g in v), max(g.m for v in pool.values() for g in v))
        eidx = {}
        for kind, graphs in pool.items():
            for g, e in zip(graphs, epool[kind]):
                eidx[id(g)] = e

        def admit(j, tag):
            n, fx, tx = eidx[id(req[j])]
            return srv.admit_edges(n, fx, tx, tag)
    else:
        def admit(j, tag):
            return srv.admit(req[j], tag)
    # untimed warm-up: one graph of each kind
    for i, k in enumerate(["sbm", "torus", "rmat"]):
        srv.admit(pool[k][0], 10 ** 9 + i)
    while srv.stats()["live_graphs"]:
        srv.step()
    srv.results(wait=True)
    queue, lat, done_edges = deque(), np.zeros(len(req)), 0
    finished, i, steps, inflight_sum = 0, 0, 0, 0
    t0 = time.perf_counter()
    last_log = t0
    while finished < len(req):
        now = time.perf_counter() - t0
        while i < len(req) and arr[i] <= now:
            queue.append(i)
            i += 1
        while queue and admit(queue[0], queue[0]):
            queue.popleft()
        live = srv.stats()["live_graphs"]
        if live:
            srv.step()
            steps += 1
            inflight_sum += live
            for tag, ids, k, it in srv.results():
                lat[tag] = time.perf_counter() - t0 - arr[tag]     # when the client has the result
                done_edges += req[tag].m
                finished += 1
        elif i < len(req):
            time.sleep(max(0.0, arr[i] - (time.perf_counter() - t0)))
        elif not queue:               # nothing in flight: collect the last results
            for tag, ids, k, it in srv.results(wait=True):
                lat[tag] = time.perf_counter() - t0 - arr[tag]
                done_edges += req[tag].m
                finished += 1
        if time.perf_counter() - last_log > 10:
            last_log = time.perf_counter()
            log(f"  t={last_log - t0:.0f} s: {finished}/{len(req)} done, {live} in flight, {len(queue)} queued")
    wall = time.perf_counter() - t0
    tuning = srv.walk_tuning()
    srv.close()
    same = None
    if check:
        # sampled requests alone and inside another mix give the same result (untimed)
        same = True
        probe = rng.choice(len(req), size=min(12, len(req)), replace=False)
        srv2 = px.Server(node_capacity=48 << 20, arc_capacity=(1 << 32) - 2, graph_slots=64,
                         params=px.Params(record_node_data=False))
        ref = {}
        for j in probe:
            srv2.admit(req[j], int(j))
            while srv2.stats()["live_graphs"]:
                srv2.step()
            for tag, ids, k, it in srv2.results(wait=True):
                ref[tag] = (ids, k, it)
        srv2.close()
        srv3 = px.Server(node_capacity=48 << 20, arc_capacity=(1 << 32) - 2, graph_slots=256,
                         params=px.Params(record_node_data=False))
        mixset = list(probe) + list(rng.choice(len(req), size=min(24, len(req)), replace=False))
        for j, tagj in enumerate(mixset):
            srv3.admit(req[tagj], (j << 32) | int(tagj))
            srv3.step()
        while srv3.stats()["live_graphs"]:
            srv3.step()
        for tag, ids, k, it in srv3.results(wait=True):
            tj = tag & 0xFFFFFFFF
            if tj in ref and (tag >> 32) < len(probe):
                rid, rk, rit = ref[tj]
                same = same and rk == k and rit == it and np.array_equal(rid, ids)
        srv3.close()
    p50, p99 = np.percentile(lat * 1e3, [50, 99])
    return {"edges_per_s": done_edges / wall, "p50_ms": float(p50), "p99_ms": float(p99), "requests": len(req),
            "duration_s": dur, "offered_edges_per_s": offered, "wall_s": wall, "steps": steps,
            "ms_per_step": wall * 1e3 / max(steps, 1), "mean_in_flight": inflight_sum / max(steps, 1),
            "results_equal_solo": same, "admission": "edge lists, CSR built on the device"
            if admit_kind == "edges" else "CSR", "walk_variant": walk_variant,
            "tuning": dict(zip(("walk", "contrib", "tunings"), tuning))}


SERVE_DESC = ("serving: Poisson 200 req/s, 60% SBM 10K-200K nodes, 30% torus 64^2-1024^2, 10% RMAT scale 14-21 "
              "ef 16 (seed 13); requests drawn from a pool of 72 graphs")


def serve_workload(a):
    s = run_serving(a.duration, a.walk_variant, a.admit)
    line = {
        "metric": METRIC, "value": s["edges_per_s"], "unit": "edges/s", "n_gpus": 1, "steps": s["steps"],
        "warmup": 3, "ms_per_step": s["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": SERVE_DESC, "admission": s["admission"], "requests": s["requests"],
                   "duration_s": s["duration_s"], "seed": 13, "offered_edges_per_s": s["offered_edges_per_s"]},
        "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"], "mean_in_flight": s["mean_in_flight"], "wall_s": s["wall_s"],
        "results_equal_solo": s["results_equal_solo"], "walk_variant": s["walk_variant"], "tuning": s["tuning"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


# --------------------------------------------------------------------------- measurement helpers

def profiled_traffic(workload):
    """HBM bytes per k_walk launch from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE passes, copy-calibrated)."""
    path = os.path.join(ROOT, "profiles", f"k_walk_{workload}_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def cpu_baseline(graphs, budget_s):
    """Reference single-threaded CPU decomposition (oracle/_ref) on a bounded
    prefix of the workload; falls back to the oracle's restatement."""
    ref = os.path.join(ROOT, "oracle", "_ref", "parex_cpu_ref")
    with tempfile.TemporaryDirectory() as td:
        paths, edges = [], 0
        for i, g in enumerate(graphs):
            if edges > 40e6:
                break
            p = os.path.join(td, f"g{i}.pxg")
            px.write_pxg(p, g)
            paths.append(p)
            edges += g.m
        if os.path.exists(ref):
            cmd = [ref, "decompose", str(budget_s), "6582", "0.3", "1e-4", *paths]
            if shutil.which("taskset"):
                cmd = ["taskset", "-c", "0", *cmd]
            out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=budget_s * 4 + 60).stdout
            # the reference's CPU code prints its own warnings on stdout first
            r = json.loads([l for l in out.splitlines() if l.startswith("{")][-1])
            return {"value": r["edges"] / r["seconds"], "unit": "edges/s", "cores": 1, "kind": "reference",
                    "sample": f"first {r['graphs']} graphs of the workload ({r['edges']} edges), "
                              f"lib/algorithms/expander_decomposition.h with Config{{0.3,1e-4,1,1}}, "
                              f"{r['seconds']:.2f} s on 1 core"}
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_ctypes import Oracle, set_threads
    set_threads(1)
    t0, e, k = time.perf_counter(), 0, 0
    for g in graphs:
        if time.perf_counter() - t0 > budget_s:
            break
        Oracle(g).run()
        e += g.m
        k += 1
    dt = time.perf_counter() - t0
    return {"value": e / dt, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": f"first {k} graphs ({e} edges) through the oracle, {dt:.2f} s on 1 core"}


def hierarchy_workload(a, k=16):
    """Expander hierarchy (every level decomposed by the engine, contracted on
    the host; lib/algorithms/expander_hierarchy.h:107-140) and a k-way
    normalized cut (lib/algorithms/normalized_cut.h:108-144) of one connected
    graph.  value = input edges / (build + cut) seconds."""
    g = px.torus(1024, 10486, 1)
    cfg = {"workload": f"expander hierarchy + {k}-way normalized cut of torus 1024x1024 + 0.5% random "
                       "long-range edges", "graphs": 1, "seed": 1, "nodes": g.n, "edges": g.m,
           "params": {"seed": 6582, "target_sparsity": 0.2, "rw_threshold": 1e-5, "k": k}}

    def once():
        h = px.ExpanderHierarchy(g)
        t0 = time.perf_counter()
        h.build()
        t1 = time.perf_counter()
        labels, val = px.NormalizedCut(h).compute(k)
        t2 = time.perf_counter()
        return h, t1 - t0, t2 - t1, val

    for w in range(a.warmup):
        _, b, c, _ = once()
        log(f"warmup {w}: build {b * 1e3:.1f} ms, cut {c * 1e3:.1f} ms")
    bs, cs = [], []
    for s in range(a.steps):
        h, b, c, val = once()
        bs.append(b)
        cs.append(c)
        log(f"step {s}: build {b * 1e3:.1f} ms, cut {c * 1e3:.1f} ms, objective {val:.6f}")
    levels = [h.lvl(i).graph.n for i in range(h.height())]
    sec = float(np.mean(bs) + np.mean(cs))
    cpu = None
    ref = os.path.join(ROOT, "oracle", "_ref", "parex_cpu_ref")
    if not a.no_cpu and os.path.exists(ref):
        with tempfile.TemporaryDirectory() as td:
            p = os.path.join(td, "g.pxg")
            px.write_pxg(p, g)
            cmd = [ref, "hierarchy", p, "6582", str(k)]
            if shutil.which("taskset"):
                cmd = ["taskset", "-c", "0", *cmd]
            out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600).stdout
            r = json.loads([line for line in out.splitlines() if line.startswith("{")][-1])
            cpu = {"value": g.m / (r["build_s"] + r["nc_s"]), "unit": "edges/s", "cores": 1, "kind": "reference",
                   "sample": f"the whole workload: ExpanderHierarchy::build<false> (CPU decompositions, "
                             f"{len(r['levels'])} levels) {r['build_s']:.2f} s + NormalizedCut::compute({k}) "
                             f"{r['nc_s']:.2f} s on 1 core"}
    line = {
        "metric": "input edges/s through the expander hierarchy build and a normalized cut",
        "value": g.m / sec, "unit": "edges/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": sec * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic", "config": cfg,
        "build_ms": float(np.mean(bs) * 1e3), "cut_ms": float(np.mean(cs) * 1e3),
        "levels": levels, "objective": val, "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


# --------------------------------------------------------------------------- batched and plain

def batched_leg(pinned, warmup, steps):
    """One engine re-initialized every step from the pinned host CSR (the
    engine keeps its device buffers and, for inputs of the same sizes, its
    tuning tables); each step times upload + initial state + decomposition."""
    dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    for w in range(warmup):
        t = time.perf_counter()
        dm.initialize_batch(pinned)
        it = dm.expanderDecomposition()
        log(f"warmup {w}: {(time.perf_counter() - t) * 1e3:.1f} ms with upload, iterations {it}")
    step_ms, init_ms, dec_ms = [], [], []
    for k in range(steps):
        t0 = time.perf_counter()
        dm.initialize_batch(pinned)
        t1 = time.perf_counter()
        it = dm.expanderDecomposition()
        t2 = time.perf_counter()
        step_ms.append((t2 - t0) * 1e3)
        init_ms.append((t1 - t0) * 1e3)
        dec_ms.append((t2 - t1) * 1e3)
        log(f"step {k}: {step_ms[-1]:.1f} ms = upload+init {init_ms[-1]:.1f} + decomposition {dec_ms[-1]:.1f} "
            f"({dm.lastRunMs():.1f} ms device), iterations {it}")
    ids, ks = dm.getFinalPartition()
    ks = np.atleast_1d(ks)            # one count per graph (a scalar for one graph)
    its = dm.graphIterations()
    # the walk kernel's roofline: one more decomposition with every walk launch timed
    dm.initialize_batch(pinned)
    dm.enableWalkTiming(True)
    dm.expanderDecomposition()
    walk = dm.walkStats()
    walk_log = dm.walkLaunchLog()
    dm.enableWalkTiming(False)
    tuning = dm.walkTuning()
    dm.close()
    return {"step_ms": float(np.mean(step_ms)), "init_ms": float(np.mean(init_ms)), "dec_ms": float(np.mean(dec_ms)),
            "ids": ids, "ks": ks, "its": its, "walk": walk, "walk_log": walk_log, "tuning": tuning}


def streamed_leg(pinned, warmup, steps, batch, engines=4):
    """A stream of `steps` batches over `engines` engines, each engine on its
    own HIP stream and host thread, batch k on engine k % engines: each engine
    uploads its batch (pinned host buffers, initial state) and decomposes it,
    while the other engines' uploads and decompositions run beside it, so the
    host-link copies and the short late-iteration kernels of one batch overlap
    another's work.  The timed region runs from the first upload to the last
    batch's final partition; every batch is the whole graph set.  Also checks
    every engine's partition equals `batch`'s."""
    import threading
    E = max(1, min(engines, steps))
    dms = [px.DeviceManager(px.Params(keep_original=False, record_node_data=False)) for _ in range(E)]
    for dm in dms:                    # each engine runs its own tuning launches here
        for _ in range(max(warmup, 1)):
            dm.initialize_batch(pinned)
            dm.expanderDecomposition()
    err, its = [], [0] * steps

    def run(e):
        try:
            for k in range(e, steps, E):
                dms[e].initialize_batch(pinned)
                its[k] = dms[e].expanderDecomposition()
        except Exception as ex:       # reported on the main thread
            err.append(ex)

    t0 = time.perf_counter()
    ths = [threading.Thread(target=run, args=(e,)) for e in range(E)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    t1 = time.perf_counter()
    if err:
        raise err[0]
    same = True
    for dm in dms:
        ids, ks = dm.getFinalPartition()
        same = same and np.array_equal(ids, batch["ids"]) and np.array_equal(np.atleast_1d(ks), batch["ks"]) \
            and np.array_equal(dm.graphIterations(), batch["its"])
    for dm in dms:
        dm.close()
    ms = (t1 - t0) * 1e3 / steps
    log(f"streamed: {steps} batches on {E} engines in {(t1 - t0) * 1e3:.1f} ms = {ms:.1f} ms a batch, "
        f"iterations {max(its)}; equal to the single-engine steps: {same}")
    return {"step_ms": ms, "same": same, "engines": E}


def plain_leg(pinned, batch):
    """One graph at a time on one engine (re-initialized per graph, degree
    rule kernels, no tuning launches), each timed like a step: upload +
    initial state + decomposition.  Also checks every partition equals the
    batch's."""
    dm = px.DeviceManager(px.Params(keep_original=False, record_node_data=False))
    dm.setWalkVariant(-2)
    tot, same, off = 0.0, True, 0
    for i, g in enumerate(pinned):
        t = time.perf_counter()
        dm.initialize(g)
        it1 = dm.expanderDecomposition()
        tot += time.perf_counter() - t
        ids1, k1 = dm.getFinalPartition()
        if not (np.array_equal(ids1, batch["ids"][off:off + g.n]) and k1 == batch["ks"][i] and it1 == batch["its"][i]):
            same = False
        off += g.n
    dm.close()
    return tot, same


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sbm512",
                    choices=["sbm512", "torus1024", "rmat22", "sbm4096", "serve", "hierarchy"])
    ap.add_argument("--graphs", type=int, default=None, help="use only the first G graphs of the set")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--engines", type=int, default=4, help="batch workloads: engines of the streamed leg")
    ap.add_argument("--no-stream", action="store_true",
                    help="batch workloads: value from the single-engine steps (upload, then decomposition)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="batch workloads: skip the serving leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--serve-duration", type=float, default=20.0,
                    help="batch workloads: seconds of the serving trace in the serving leg")
    ap.add_argument("--duration", type=float, default=60.0, help="serving workload: seconds of arrivals")
    ap.add_argument("--walk-variant", type=int, default=None,
                    help="serving workload: pin the per-arc kernel variant (default: online tuning)")
    ap.add_argument("--admit", choices=["csr", "edges"], default="csr",
                    help="serving workload: requests as CSR or as edge lists (CSR built on the device)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        return 2
    if a.workload == "serve":
        return serve_workload(a)
    if a.workload == "hierarchy":
        return hierarchy_workload(a)

    t = time.perf_counter()
    graphs, cfg = workload(a.workload, a.graphs)
    n_tot = sum(g.n for g in graphs)
    m_tot = sum(g.m for g in graphs)
    cfg.update({"nodes": n_tot, "edges": m_tot, "params": {"seed": 6582, "sc_threshold": 0.3, "rw_threshold": 1e-4},
                "timed": "upload from pinned host memory + initial state + decomposition, per step"})
    pinned, _keep = pinned_graphs(graphs)
    log(f"workload {a.workload}: {len(graphs)} graphs, {n_tot} nodes, {m_tot} edges ({time.perf_counter() - t:.1f} s)")

    b = batched_leg(pinned, a.warmup, a.steps)
    serial_value = m_tot / (b["step_ms"] / 1e3)
    s = None if a.no_stream else streamed_leg(pinned, a.warmup, a.steps, b, a.engines)
    if s is not None and not s["same"]:
        log("streamed batches differ from the single-engine steps")
        return 1
    step_ms = s["step_ms"] if s is not None else b["step_ms"]
    value = m_tot / (step_ms / 1e3)
    if s is not None:
        cfg["timed"] = (f"a stream of `steps` batches: upload from pinned host memory + initial state + "
                        f"decomposition of every batch, on {s['engines']} engines with a HIP stream and a host "
                        f"thread each, their batches running side by side; from the first upload to the last "
                        f"partition")
    launches, walk_ms, walk_bytes = b["walk"]
    tuning, ctuning, tunings = b["tuning"]
    achieved = walk_bytes / (walk_ms / 1e3) / 1e9 if walk_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(a.workload) if a.graphs is None else None,
                "kernel": "k_walk", "launches": launches,
                "avg_launch_us": round(walk_ms * 1e3 / max(launches, 1), 2),
                "alg_bytes_per_launch": round(walk_bytes / max(launches, 1)),
                # walk kernel chosen per log2(active-list length) class by the
                # online tuner: 0..5 = 2..64 lanes per row, 6/7/8 = LDS window
                # with 1/2/4 lanes per row, 9 = LDS window, 2 lanes, 1024-entry
                # chunks, 10 = 7 with clamped walk slots, 11 / 12 = LDS window
                # over node-range chunks of 8192 / 4096 indices, 13 = half
                # window over 2048-entry chunks, 14 / 15 = 11 / 12 with one lane
                # per row
                "walk_variant_by_log2_nodes": tuning, "contrib_variant_by_log2_nodes": ctuning,
                "tunings": tunings}
    lms, lby = b["walk_log"]
    if len(lms):
        big = lby >= 0.5 * lby.max()        # launches over at least half the largest one's bytes
        roofline["launches_at_half_max_bytes"] = {
            "count": int(big.sum()), "bytes_share": round(float(lby[big].sum() / lby.sum()), 4),
            "frac": round(float(lby[big].sum() / (lms[big].sum() / 1e3) / 1e9 / HBM_PEAK_GBS), 4)}
        roofline["launch_us"] = [round(float(v) * 1e3, 1) for v in lms]
        roofline["launch_mb"] = [round(float(v) / 1e6, 1) for v in lby]

    plain, same = None, None
    if not a.no_plain:
        tot, same = plain_leg(pinned, b)
        plain = m_tot / tot
        log(f"plain: {tot * 1e3:.1f} ms total, {plain:.3e} edges/s; batched == plain: {same}")

    cpu = None if a.no_cpu else cpu_baseline(graphs, a.cpu_budget)

    serve = None
    if not a.no_serve:
        serve = run_serving(a.serve_duration)
        log(f"serving: {serve['edges_per_s']:.3e} edges/s, p50 {serve['p50_ms']:.0f} ms, p99 {serve['p99_ms']:.0f} ms")

    line = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic", "config": cfg,
        "value_serial": serial_value, "serial_ms_per_step": b["step_ms"],
        "upload_init_ms": round(b["init_ms"], 2), "decomposition_ms": round(b["dec_ms"], 2),
        "value_resident": m_tot / (b["dec_ms"] / 1e3),
        "plain_value": plain, "batched_over_plain": (serial_value / plain) if plain else None,   # both upload, then run
        "batched_equals_plain": same,
        "iterations_max": int(b["its"].max()) if len(b["its"]) else 0,
        "p50_ms": serve["p50_ms"] if serve else None, "p99_ms": serve["p99_ms"] if serve else None,
        "served_edges_per_s": serve["edges_per_s"] if serve else None,
        "serving": ({k: v for k, v in serve.items() if k not in ("p50_ms", "p99_ms", "edges_per_s")} |
                    {"workload": SERVE_DESC}) if serve else None,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
