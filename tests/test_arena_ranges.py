"""The server arena's free-range bookkeeping (csrc/ranges.hpp), run on the CPU
against a unit bitmap: first fit must pick the lowest offset with enough
consecutive free 256-byte units, and after every operation the free ranges
must be exactly the maximal runs of free units (returned blocks merge)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "wgpu_solver_backend_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include "ranges.hpp"
int main() {
  hbj::FreeRanges f;
  char op;
  unsigned long long a, b;
  while (std::scanf(" %c", &op) == 1) {
    if (op == 'r') { std::scanf("%llu", &a); f.reset(a); }
    else if (op == 't') {
      std::scanf("%llu", &a);
      uint64_t off = 0;
      if (f.take(a, &off)) std::printf("%llu\n", (unsigned long long)off); else std::printf("-1\n");
    } else if (op == 'g') { std::scanf("%llu %llu", &a, &b); f.give(a, b); }
    else if (op == 'l') {
      for (const auto& kv : f.ranges()) std::printf("%llu:%llu ", (unsigned long long)kv.first,
                                                    (unsigned long long)kv.second);
      std::printf("\n");
    }
  }
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("ranges")
    src, exe = d / "drv.cpp", d / "drv"
    src.write_text(DRIVER)
    subprocess.run([gxx, "-O1", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    return str(exe)


def runs(free):
    """maximal runs of free units -> [(offset_bytes, length_bytes)]"""
    out, i, n = [], 0, free.size
    while i < n:
        if free[i]:
            j = i
            while j < n and free[j]:
                j += 1
            out.append((i * 256, (j - i) * 256))
            i = j
        else:
            i += 1
    return out


def test_first_fit_and_merging_match_a_unit_bitmap(driver):
    rng = np.random.default_rng(3)
    units = 4096
    free = np.ones(units, bool)
    live = []  # (offset, bytes)
    ops, expect = [f"r {units * 256}"], []
    for step in range(3000):
        if live and (rng.random() < 0.45 or len(live) > 60):
            off, nb = live.pop(int(rng.integers(len(live))))
            ops.append(f"g {off} {nb}")
            u = (nb + 255) // 256
            free[off // 256: off // 256 + u] = True
        else:
            nb = int(rng.choice([1, 255, 256, 257, 1000, 4096, 20000, 100000]) + rng.integers(0, 300))
            u = (nb + 255) // 256
            ops.append(f"t {nb}")
            run = np.convolve(free.astype(np.int64), np.ones(u, np.int64), "valid") == u
            hit = np.flatnonzero(run)
            if hit.size:
                off = int(hit[0]) * 256
                free[off // 256: off // 256 + u] = False
                live.append((off, nb))
                expect.append(("t", off))
            else:
                expect.append(("t", -1))
        ops.append("l")
        expect.append(("l", runs(free)))
    out = subprocess.run([driver], input="\n".join(ops) + "\n", capture_output=True, text=True,
                         check=True).stdout.splitlines()
    assert len(out) == len(expect)
    for line, (kind, want) in zip(out, expect):
        if kind == "t":
            assert int(line) == want
        else:
            got = [tuple(int(v) for v in kv.split(":")) for kv in line.split()]
            assert got == want
    assert any(w == -1 for k, w in expect if k == "t")  # the arena did fill up
