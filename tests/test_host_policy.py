"""Host-side pieces of plan and pool set-up, compiled alone with g++: the
scheduling arithmetic of pools (fdtd2d_amd/csrc/pool_policy.h: the turns of
time-sliced shots, the list copy a rebuild writes) and the block scan for
plasma runs the tiler uses (runs.h), against brute force. The GPU tests
check the same rules end to end (results equal to the oracle); these pin
them on the CPU."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

HARNESS = r"""
#include <cstdio>
#include "pool_policy.h"
using namespace fdtd2d::policy;
int main() {
  int bad = 0;
  // brute force: the first turn at or after g is the smallest multiple of k >= g,
  // and g + to_next_turn(g, k) is the next multiple of k after g
  for (int k = 1; k <= 64; ++k)
    for (int g = 0; g < 1000; ++g) {
      int t = g;
      while (t % k) ++t;
      int u = g + 1;
      while (u % k) ++u;
      if (first_turn(g, k) != t) { printf("first_turn(%d, %d) = %d, want %d\n", g, k, first_turn(g, k), t); ++bad; }
      if (g + to_next_turn(g, k) != u) { printf("to_next_turn(%d, %d) wrong\n", g, k); ++bad; }
    }
  // the copy a rebuild writes
  struct Case { bool valid[4]; int last_g[4]; int n, cur, want; } cases[] = {
    {{true, false, true, true}, {9, -1, 3, 7}, 4, 0, 1},   // an unused copy first
    {{true, true, true, true}, {9, 5, 3, 7}, 4, 0, 2},     // else the least recently read
    {{true, true, true, true}, {1, 5, 3, 7}, 4, 0, 2},     // never the current one
    {{false, false, true, true}, {-1, -1, 3, 7}, 4, 0, 1}, // not the current one, unused or not
    {{true, true, false, false}, {4, 6, -1, -1}, 2, 1, 0}, // two copies: the other one
    {{true, true, true, true}, {8, 2, 6, 4}, 2, 3, 1},     // current copy beyond n: among the first n
  };
  for (const Case &c : cases) {
    const int got = pick_copy(c.valid, c.last_g, c.n, c.cur);
    if (got != c.want) { printf("pick_copy case cur=%d n=%d: %d, want %d\n", c.cur, c.n, got, c.want); ++bad; }
  }
  printf("bad %d\n", bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_pool_policy(tmp_path):
    src = tmp_path / "policy.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "policy"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I", str(ROOT / "fdtd2d_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True, text=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().endswith("bad 0")


RUNS_HARNESS = r"""
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>
#include "runs.h"
int main() {
  std::mt19937 rng(7);
  const float vals[] = {0.f, -0.f, 1e18f, -3.f, NAN, 1e-30f, INFINITY};
  long bad = 0;
  for (int t = 0; t < 100000; ++t) {
    const int n = 1 + rng() % 300;
    std::vector<float> r(n);
    if (rng() % 3 == 0) {
      for (auto &v : r) v = vals[rng() % 7];
    } else {  // one long run, maybe with a stray cell
      const int a = rng() % n, b = rng() % (n + 1);
      for (int i = 0; i < n; ++i) r[i] = (i >= a && i < b) ? 1e18f : 0.f;
      if (rng() % 2) r[rng() % n] = vals[rng() % 7];
    }
    std::vector<int> want, got;
    int run = -1;  // brute force, cell by cell
    for (int i = 0; i < n; ++i) {
      if (r[i] > 0.f) { if (run < 0) run = i; }
      else if (run >= 0) { want.push_back(run); want.push_back(i); run = -1; }
    }
    if (run >= 0) { want.push_back(run); want.push_back(n); }
    const float *ne = r.data();
    fdtd2d::scan_runs(n, [=](int i) { return static_cast<int>(ne[i] > 0.f); },
                      [&](int a, int b) { got.push_back(a); got.push_back(b); });
    bad += got != want;
  }
  printf("bad %ld\n", bad);
  return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_plasma_run_scan(tmp_path):
    src = tmp_path / "runs.cpp"
    src.write_text(RUNS_HARNESS)
    exe = tmp_path / "runs"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", str(ROOT / "fdtd2d_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True, capture_output=True, text=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "bad 0", r.stdout
