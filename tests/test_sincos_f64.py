"""The FP64 pair kernels' sin/cos (cudaprob3_amd/csrc/sincos_f64.hpp) on the CPU: the header is
compiled with g++ and checked against the C library's 80-bit sinl/cosl (about 11 bits more than
double) and against its double sin/cos, which the reference's getA calls (physics.hpp:462-474).
Bars: within 1.75 ulp of the exact values (1.47 seen), within 2 ulp of the C library's doubles
(1.0 seen), over phases up
to 10^11 rad (energies far below the physical range) and at quarter-turn boundaries."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include "sincos_f64.hpp"
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <initializer_list>

static uint64_t state = 0x9e3779b97f4a7c15ull;
static double uniform() {                      // [0, 1), 53 bits, splitmix64
    uint64_t z = (state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return double((z ^ (z >> 31)) >> 11) * 0x1.0p-53;
}
static double ulp_of(long double v) {
    const double d = std::fabs(double(v));
    return d > 0 ? std::nextafter(d, INFINITY) - d : 4.9e-324;
}
struct Err { double exact = 0, libm = 0; };
static void check(double x, Err& e) {
    double s, c;
    if (!cp3::sincos_reduced_f64(x, s, c)) { std::printf("REFUSED %.17g\n", x); return; }
    const long double rs = sinl((long double)x), rc = cosl((long double)x);
    e.exact = std::fmax(e.exact, double(std::fabs((long double)s - rs)) / ulp_of(rs));
    e.exact = std::fmax(e.exact, double(std::fabs((long double)c - rc)) / ulp_of(rc));
    e.libm = std::fmax(e.libm, std::fabs(s - std::sin(x)) / ulp_of(rs));
    e.libm = std::fmax(e.libm, std::fabs(c - std::cos(x)) / ulp_of(rc));
}
int main() {
    const double spans[] = {0.785, 10.0, 1.0e4, 1.0e7, 1.0e11};
    for (double span : spans) {
        Err e;
        for (int i = 0; i < 400000; i++) check((2.0 * uniform() - 1.0) * span, e);
        std::printf("span %g exact %.3f libm %.3f\n", span, e.exact, e.libm);
    }
    Err e;
    for (int k = -200000; k <= 200000; k++)           // quarter-turn boundaries and multiples of pi/2
        for (double d : {0.0, 1e-9, -1e-9, 0.78539816, -0.78539816}) check(k * 1.5707963267948966 + d, e);
    std::printf("span boundaries exact %.3f libm %.3f\n", e.exact, e.libm);
    Err t;
    for (int i = 0; i < 200000; i++) check(std::pow(10.0, -300.0 + 297.0 * uniform()) * (i % 2 ? 1 : -1), t);
    std::printf("span tiny exact %.3f libm %.3f\n", t.exact, t.libm);
    double s = 7, c = 7;
    const bool big = cp3::sincos_reduced_f64(2.0e12, s, c), inf = cp3::sincos_reduced_f64(INFINITY, s, c),
               nan = cp3::sincos_reduced_f64(NAN, s, c);
    std::printf("refuses %d %d %d untouched %d\n", !big, !inf, !nan, s == 7 && c == 7);
    return 0;
}
"""


@pytest.fixture(scope="module")
def report(tmp_path_factory):
    d = tmp_path_factory.mktemp("sincos")
    src, exe = d / "check.cpp", d / "check"
    src.write_text(PROGRAM)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", "-I",
                        os.path.join(ROOT, "cudaprob3_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    print(r.stdout)
    return r.stdout


def test_within_two_ulp_of_exact_and_of_the_c_library(report):
    assert "REFUSED" not in report
    rows = re.findall(r"span (\S+) exact ([0-9.]+) libm ([0-9.]+)", report)
    assert len(rows) == 7
    for span, exact, libm in rows:
        assert float(exact) <= 1.75, (span, exact)
        assert float(libm) <= 2.0, (span, libm)


def test_out_of_range_arguments_are_left_to_the_caller(report):
    assert "refuses 1 1 1 untouched 1" in report
