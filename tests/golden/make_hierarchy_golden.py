"""Writes tests/golden/hierarchy_ref.npz from the reference's own CPU code.

oracle/_ref/parex_cpu_ref (oracle/Makefile) compiles the reference's
header-only sources.  Its "hierarchy" mode runs
  ExpanderHierarchy::build<false>   lib/algorithms/expander_hierarchy.h:107-140
  NormalizedCut::compute(k)         lib/algorithms/normalized_cut.h:108-144
  compute_normalized_cut            lib/core/partition.h:418-427
on small connected graphs (the reference's build does not end on a graph with
more than one component).  Recorded per graph: the input CSR; per level the
partition of the graph below (member_off, members, in the reference's member
order), the contracted multigraph (ranges, edges in the reference's arc order),
volumes, cut sizes and parents; per k the size, volume and internal volume of
each of the k parts in label order and the objective (float32).

Run once in the container that has the reference tree:
    python tests/golden/make_hierarchy_golden.py
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import parex_amd as px  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "parex_cpu_ref")
SEED = 6582
KS = (2, 4, 8, 16)


def connected_sbm(n, k, p_in, p_out, seed):
    """A planted-partition graph plus the path 0-1-...-(n-1), so it is connected."""
    g = px.sbm(n, k, p_in, p_out, seed)
    src = np.repeat(np.arange(g.n), np.diff(g.ranges))
    keep = src < g.edges
    return px.from_edges(n, np.concatenate([src[keep], np.arange(n - 1)]),
                         np.concatenate([g.edges[keep], np.arange(1, n)]))


def golden_graphs():
    return {
        "csbm500": connected_sbm(500, 4, 0.05, 0.003, 3),
        "csbm2000": connected_sbm(2000, 8, 0.02, 0.0005, 5),
        "torus16": px.torus(16, 5, 1),
        "path50": px.from_edges(50, np.arange(49), np.arange(1, 50)),
    }


def run_ref(path, k):
    r = subprocess.run([REF, "hierarchy", path, str(SEED), str(k)], check=True, capture_output=True, text=True,
                       timeout=300)
    return json.loads([line for line in r.stdout.splitlines() if line.startswith("{")][-1])


def main():
    if not os.path.exists(REF):
        raise SystemExit(f"{REF} missing: run make -C oracle (needs the reference tree)")
    out = {}
    with tempfile.TemporaryDirectory() as td:
        for name, g in golden_graphs().items():
            gp = os.path.join(td, name + ".pxg")
            px.write_pxg(gp, g)
            out[f"{name}_ranges"] = g.ranges.astype(np.uint32)
            out[f"{name}_edges"] = g.edges.astype(np.uint32)
            for k in KS:
                d = run_ref(gp, k)
                if k == KS[0]:
                    out[f"{name}_height"] = np.array([len(d["levels"])], np.uint32)
                    for li, lvl in enumerate(d["levels"]):
                        for f in ("member_off", "members", "ranges", "edges", "volumes", "cut_sizes", "parents"):
                            out[f"{name}_L{li}_{f}"] = np.array(lvl[f], np.uint32)
                for f in ("nc_size", "nc_volume", "nc_internal"):
                    out[f"{name}_k{k}_{f}"] = np.array(d[f], np.uint32)
                out[f"{name}_k{k}_nc_value"] = np.array([d["nc_value"]], np.float32)
            print(name, g.n, g.m, "levels", int(out[f"{name}_height"][0]), flush=True)
    np.savez_compressed(os.path.join(HERE, "hierarchy_ref.npz"), **out)


if __name__ == "__main__":
    main()
