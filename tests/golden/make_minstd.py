"""Writes tests/golden/minstd.json: first output of libstdc++'s std::minstd_rand
(a = 48271, m = 2^31-1, the engine thrust::default_random_engine aliases)
for a set of seeds, as an independent pin of the oracle's generator."""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = [0, 1, 2, 6582, 12345, 2147483646, 2147483647, 2147483648, 4294967295, 0x45d9f3b]

SRC = r"""
#include <random>
#include <cstdio>
#include <cstdlib>
int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        std::minstd_rand r((std::minstd_rand::result_type)std::strtoull(argv[i], nullptr, 10));
        std::printf("%u\n", (unsigned)r());
    }
}
"""

with tempfile.TemporaryDirectory() as td:
    src = os.path.join(td, "m.cpp")
    open(src, "w").write(SRC)
    exe = os.path.join(td, "m")
    subprocess.run(["g++", "-O2", src, "-o", exe], check=True)
    out = subprocess.run([exe, *map(str, SEEDS)], check=True, capture_output=True, text=True).stdout.split()
json.dump({"seeds": SEEDS, "first": [int(v) for v in out]}, open(os.path.join(HERE, "minstd.json"), "w"), indent=1)
print(out)
