"""Accuracy of the closures' FP64 logarithms (frehg2_amd/csrc/gw_math.h),
compiled for the host from the same header the kernels include and checked
against long-double libm: at most a few ulp over the ranges the closures
use, and libm's special values."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

PROG = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include "gw_math.h"
static double ulps(double got, long double ref) {
  const double r = (double)ref;
  if (r == 0.0) return got == 0.0 ? 0.0 : 1e300;
  const double u = std::ldexp(1.0, std::ilogb(r) - 52);
  return (double)(std::fabs((long double)got - ref) / (long double)u);
}
int main() {
  uint64_t st = 0x9E3779B97F4A7C15ull;
  auto next = [&]() { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; };
  double worstLog = 0, worstLog1p = 0, xl = 0, xp = 0;
  for (int i = 0; i < 4000000; ++i) {
    // log: log-uniform over [1e-300, 1e300], plus [0.5, 2] densely
    const double a = (next() >> 11) * 0x1.0p-53;
    const double x = (i & 1) ? std::pow(10.0, -300.0 + 600.0 * a) : 0.5 + 1.5 * a;
    const double e = ulps(frehg::gw_log(x), logl((long double)x));
    if (e > worstLog) { worstLog = e; xl = x; }
    // log1p: log-uniform over [1e-20, 1e20], and (-0.9, 0)
    const double b = (next() >> 11) * 0x1.0p-53;
    const double y = (i & 1) ? std::pow(10.0, -20.0 + 40.0 * b) : -0.9 * b;
    const double f = ulps(frehg::gw_log1p(y), log1pl((long double)y));
    if (f > worstLog1p) { worstLog1p = f; xp = y; }
  }
  std::printf("log %.3f %.17g\nlog1p %.3f %.17g\n", worstLog, xl, worstLog1p, xp);
  const double sub = 4.9406564584124654e-324;
  std::printf("special %d %d %d %d %d %d %d\n", std::isinf(frehg::gw_log(0.0)) && frehg::gw_log(0.0) < 0,
              std::isinf(frehg::gw_log(INFINITY)) && frehg::gw_log(INFINITY) > 0,
              std::isnan(frehg::gw_log(-1.0)), std::isnan(frehg::gw_log(NAN)),
              frehg::gw_log(1.0) == 0.0, std::fabs(frehg::gw_log(sub) - (double)logl(sub)) < 1e-13,
              frehg::gw_log1p(0.0) == 0.0 && frehg::gw_log1p(1e-300) == 1e-300);
  return 0;
}
"""


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    d = tmp_path_factory.mktemp("math")
    src = d / "t.cpp"
    src.write_text(PROG)
    exe = d / "t"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT / "frehg2_amd" / "csrc"), "-o",
                    str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {l.split()[0]: l.split()[1:] for l in out.splitlines()}


def test_log_within_one_and_a_half_ulp(results):
    assert float(results["log"][0]) <= 1.5, results["log"]


def test_log1p_within_three_ulp(results):
    assert float(results["log1p"][0]) <= 3.0, results["log1p"]


def test_special_values_follow_libm(results):
    assert results["special"] == ["1"] * 7, results["special"]
