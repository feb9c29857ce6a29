"""The engine's own kernel and host sources, compiled for the CPU by the
emulator (tests/emu: one thread per lane, barriers for workgroup and wave
exchanges) with AddressSanitizer, decompose small graphs end to end and
equal the oracle (2500 nodes: segments above a sort tile, sorted by the
tile plan the device builds from its own counts), and the device CSR build of an edge list (loops,
duplicates, a row longer than a workgroup's tile) equals the host build.
Every device array is a host allocation there, so an
access out of bounds -- such as a wave wholly past the end of the position
list using the past-the-end cluster sentinel as an atomic index, the fault
behind commit 4b7b85d -- stops the run with a report.  No GPU needed."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "emu"))


def _asan_runtime():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.timeout(1800)
def test_emulated_engine_with_address_sanitizer_matches_oracle():
    asan = _asan_runtime()
    if asan is None:
        pytest.skip("no AddressSanitizer runtime for gcc")
    import build_emu
    lib = build_emu.build(asan=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0")
    # "/phases" forces the per-phase kernels (the 100-node path: waves wholly
    # past the end of the position list, the regression of 4b7b85d); the edge
    # lists run the device CSR build.  Each emulated decomposition takes about
    # two minutes under AddressSanitizer, so the default suite runs these; the
    # one-workgroup runs (path100, sbm300) and the large-segment sort plan
    # (sbm2500/phases) run with PXD_EMU_CASES=all (the GPU tests cover them too).
    cases = ["path100/phases", "ingest_loops_dups", "ingest_hub3000"]
    if os.environ.get("PXD_EMU_CASES") == "all":
        cases = ["path100", "path100/phases", "sbm300", "sbm2500/phases", "ingest_loops_dups",
                 "ingest_hub3000"]
    r = subprocess.run([sys.executable, os.path.join(HERE, "emu", "run_emu_parity.py"), lib, *cases],
                       capture_output=True, text=True, timeout=1700, env=env)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-6000:]
    assert r.stdout.count("equal") == len(cases) and "ERROR: AddressSanitizer" not in r.stderr
