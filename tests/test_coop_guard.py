"""The process-wide guard of the persistent (cooperative) Lanczos form,
checked in a host-only build (g++, threads, no GPU): a second context is
refused while a first holds the form, the same context may nest, and the
form is free again after release.  Every entry point that can launch the
persistent form takes this guard before any HIP call."""
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "usbs_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <thread>
#include <atomic>
#include "coop_guard.hpp"
using usbs::CoopGuard;
int main() {
  int a = 0, b = 0;
  int fails = 0;
  {
    CoopGuard g1(&a);
    if (!g1.ok()) ++fails;
    {
      CoopGuard nested(&a);           // same context: allowed
      if (!nested.ok()) ++fails;
    }
    std::atomic<int> other_ok{-1};
    std::thread t([&] { CoopGuard g2(&b); other_ok = g2.ok() ? 1 : 0; });
    t.join();
    if (other_ok != 0) ++fails;       // another context: refused
    CoopGuard none(nullptr);          // per-op call: never blocked, holds nothing
    if (!none.ok()) ++fails;
  }
  {
    CoopGuard g3(&b);                 // released: free again
    if (!g3.ok()) ++fails;
  }
  // many threads race for the form: exactly one holds it at a time
  std::atomic<int> inside{0}, maxinside{0}, wins{0};
  std::thread ts[8];
  int owners[8];
  for (int q = 0; q < 8; ++q)
    ts[q] = std::thread([&, q] {
      for (int it = 0; it < 20000; ++it) {
        CoopGuard g(&owners[q]);
        if (g.ok()) {
          int v = ++inside;
          int mx = maxinside.load();
          while (v > mx && !maxinside.compare_exchange_weak(mx, v)) {}
          ++wins;
          --inside;
        }
      }
    });
  for (auto& t : ts) t.join();
  if (maxinside.load() != 1 || wins.load() < 1) ++fails;
  std::printf("fails=%d wins=%d\n", fails, wins.load());
  return fails;
}
"""


def test_coop_guard_host_build():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "guard_test.cpp")
        exe = os.path.join(d, "guard_test")
        open(src, "w").write(DRIVER)
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-pthread", "-I", CSRC, src, "-o", exe],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("src,entry", [("solve.hip", "usbs_solve"), ("cgal.hip", "usbs_cgal_solve"),
                                       ("engine.hip", "usbs_lanczos_smallest"),
                                       ("engine.hip", "usbs_lanczos_smallest_batch")])
def test_entry_points_take_the_guard_before_hip_calls(src, entry):
    """In each entry point that can run the persistent form, the guard is
    checked before the first HIP call (hipSetDevice)."""
    text = open(os.path.join(CSRC, src)).read()
    start = text.index("int %s(" % entry)
    body = text[start:]
    body = body[:body.index("\n}\n")]
    g = body.find("CoopGuard coop(")
    h = body.find("hipSetDevice")
    assert g >= 0 and h >= 0 and g < h, entry
    assert re.search(r"!coop\.ok\(\)\)[^;]*USBS_ERR_STATE", body), entry
