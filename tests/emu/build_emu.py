"""Builds the CPU emulation of the HIP engine (test infrastructure only).

The engine source is compiled unchanged except that every kernel launch
`k<<<grid, block, shmem, stream>>>(args)` is rewritten into a call of the
emulator's launcher, that the scans' and sorts' size thresholds are
lowered so that small graphs also take the multi-tile paths, and that the
device-only optimizer barrier is dropped.  Output: tests/emu/_build/libpxd_emu.so (and
libpxd_emu_asan.so with AddressSanitizer when asan=True).
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "parex_amd", "csrc")
OUT = os.path.join(HERE, "_build")


def _split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    parts.append(cur.strip())
    return parts


def transform(src: str) -> str:
    out, i = [], 0
    pat = re.compile(r"([A-Za-z_][A-Za-z_0-9]*(?:<[A-Za-z_0-9, ]+>)?)<<<")
    while True:
        m = pat.search(src, i)
        if not m:
            out.append(src[i:])
            break
        out.append(src[i:m.start()])
        name = m.group(1)
        j = src.index(">>>", m.end())
        cfg = _split_top(src[m.end():j])
        k = j + 3
        assert src[k] == "(", src[k:k + 20]
        depth, p = 0, k
        while True:
            if src[p] == "(":
                depth += 1
            elif src[p] == ")":
                depth -= 1
                if depth == 0:
                    break
            p += 1
        args = src[k + 1:p]
        out.append(f"emu::launch_impl((uint32_t)({cfg[0]}), (uint32_t)({cfg[1]}), [&] {{ {name}({args}); }}, \"{name}\")")
        i = p + 1
    return "".join(out)


def build(asan=False):
    os.makedirs(OUT, exist_ok=True)
    src = open(os.path.join(CSRC, "pxd_engine.hip")).read()
    gen = os.path.join(OUT, "pxd_engine_emu.cpp")
    open(gen, "w").write(transform(src))
    # headers with launches of their own: transformed copies next to the
    # generated source, which its quoted includes find first
    for h in ("pxd_ingest.h", "pxd_small.h"):
        open(os.path.join(OUT, h), "w").write(transform(open(os.path.join(CSRC, h)).read()))
    # the kernel header with small look-back scan and selection tiles (1024
    # elements), no one-workgroup scans above 1024 elements and one-workgroup
    # sorts only up to 1024 items, so that the small graphs of the parity runs
    # already take the single-pass scan and selection over several tiles and
    # the size-class sort
    kh = open(os.path.join(CSRC, "pxd_kernels.h")).read()
    for old, new in (("constexpr uint32_t SMALL_SCAN_N = 1024 * SMALL_SCAN_ITEMS;", "constexpr uint32_t SMALL_SCAN_N = 0;"),
                     ("constexpr uint32_t LB_SCAN_ITEMS = 32;", "constexpr uint32_t LB_SCAN_ITEMS = 4;"),
                     ("constexpr uint32_t SORT_SMALL_USE = SORT_SMALL_N;", "constexpr uint32_t SORT_SMALL_USE = 1024;"),
                     ("constexpr uint32_t SE_ROWS = 16, SE_ROWS_SHORT = 8;",
                      "constexpr uint32_t SE_ROWS = 4, SE_ROWS_SHORT = 2;"),
                     ("constexpr uint32_t SE_LONG = 1u << 23;", "constexpr uint32_t SE_LONG = 1500;")):
        assert kh.count(old) == 1, old
        kh = kh.replace(old, new)
    open(os.path.join(OUT, "pxd_kernels.h"), "w").write(kh)
    # the device helpers without the VGPR constraint of the optimizer barrier,
    # which the host compiler does not know (the emulation needs no barrier)
    dh = open(os.path.join(CSRC, "pxd_device.h")).read()
    old = '__asm__ volatile("" : "+v"(v));'
    assert dh.count(old) >= 1, old
    open(os.path.join(OUT, "pxd_device.h"), "w").write(dh.replace(old, "(void)v;"))
    lib = os.path.join(OUT, "libpxd_emu_asan.so" if asan else "libpxd_emu.so")
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address"] if asan else ["-O2"]
    cmd = ["g++", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", *flags, "-pthread",
           "-I" + os.path.join(HERE, "include"), "-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
           os.path.join(HERE, "emu_hip.cpp"), gen, "-o", lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-6000:])
    return lib


if __name__ == "__main__":
    print(build(asan="--asan" in sys.argv))
