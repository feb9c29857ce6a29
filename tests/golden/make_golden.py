"""Writes tests/golden/ref_cpu.npz from the reference's own CPU code.

oracle/_ref/parex_cpu_ref is compiled (oracle/Makefile) from the reference's
header-only CPU sources.  This script feeds it small deterministic graphs and
records what those sources compute:

  walk   RandomWalk::iterate (lib/algorithms/random_walk.h:51-70), k steps
         from the oracle's initial vector (float32 in, float32 out)
  sweep  Partition::sweepCut (lib/core/partition.h:309-362) on the values the
         oracle holds after k walk steps: best sparsity and offset
  verify splitByIndices<false,false> (lib/core/partition.h:233-295) cut-edge
         counts of seeded random partitions

Run once in the container that has the reference tree:
    python tests/golden/make_golden.py
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
sys.path.insert(0, os.path.dirname(HERE))

import parex_amd as px  # noqa: E402
from oracle_ctypes import Oracle, init_vector  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "parex_cpu_ref")


def golden_graphs():
    return {
        "sbm_c1": px.sbm(4096, 16, 12 / 255, 1 / 3840, 6582),
        "torus16": px.torus(16, 5, 1),
        "path50": px.from_edges(50, np.arange(49), np.arange(1, 50)),
        "sbm500": px.sbm(500, 4, 0.05, 0.003, 3),
    }


WALK_STEPS = (1, 4, 16)
SWEEP_STEPS = (1, 3, 10)


def main():
    if not os.path.exists(REF):
        raise SystemExit(f"{REF} missing: run make -C oracle (needs the reference tree)")
    out = {}
    meta = {}
    with tempfile.TemporaryDirectory() as td:
        for name, g in golden_graphs().items():
            gp = os.path.join(td, name + ".pxg")
            px.write_pxg(gp, g)
            x0 = init_vector(g.degree(), 6582)
            xin = os.path.join(td, "x0.f32")
            x0.astype(np.float32).tofile(xin)
            for k in WALK_STEPS:
                xo = os.path.join(td, "xo.f32")
                subprocess.run([REF, "walk", gp, xin, str(k), xo], check=True)
                out[f"{name}_walk{k}"] = np.fromfile(xo, np.float32)
            for k in SWEEP_STEPS:
                ox = Oracle(g, init_values=x0)
                for _ in range(k):
                    ox.walk()
                v = ox.values().astype(np.float32)
                vp = os.path.join(td, "v.f32")
                v.tofile(vp)
                r = json.loads(subprocess.run([REF, "sweep", gp, vp], check=True, capture_output=True,
                                              text=True).stdout)
                out[f"{name}_sweep{k}"] = np.array([r["sparsity"]], np.float32)
                out[f"{name}_sweep{k}_offset"] = np.array([r["offset"]], np.uint32)
            rng = np.random.default_rng(11)
            cuts = []
            for parts in (1, 2, 5, 17):
                ids = rng.integers(0, parts, g.n).astype(np.int32)
                ip = os.path.join(td, "ids.i32")
                ids.tofile(ip)
                r = json.loads(subprocess.run([REF, "verify", gp, ip, str(parts)], check=True, capture_output=True,
                                              text=True).stdout)
                cuts.append([parts, r["numClusters"], r["cutEdges"]])
            out[f"{name}_verify"] = np.array(cuts, np.int64)
            meta[name] = {"n": g.n, "m": g.m}
    np.savez_compressed(os.path.join(HERE, "ref_cpu.npz"), **out)
    with open(os.path.join(HERE, "ref_cpu.json"), "w") as f:
        json.dump({"graphs": meta, "walk_steps": WALK_STEPS, "sweep_steps": SWEEP_STEPS,
                   "verify_seed": 11, "init_seed": 6582}, f, indent=1)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
