"""The stage kernels' own elementary functions, compiled for the host from the
same headers and checked against the C library:

- the logarithm of the tabulated EOS (gpu-vah_amd/csrc/vah_log.h): within 2 ulp
  over 600 decades, exact special values.  The table index it feeds is
  (ln e - lne0) * inv_h, so 2 ulp of ln e move the interpolation weight by
  ~1e-12 at most;
- the regulator's tanh(rho) / rho (gpu-vah_amd/csrc/vah_tanh.h, the damping
  factor of FullyDiscreteKurganovTadmorScheme.cpp:652-665): within 6 ulp of
  tanh(x) / x over [1e-7.5, 30], 0 at +inf, NaN at NaN."""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpu-vah_amd", "csrc")

SRC = r'''
#define __host__
#define __device__
#include LOG_H
#include TANH_H
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
static int64_t bits(double x) { int64_t i; std::memcpy(&i, &x, 8); return i; }
int main() {
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> u(-300.0, 300.0);
    long long worst = 0;
    for (int n = 0; n < 2000000; ++n) {
        const double x = std::pow(10.0, u(g));
        const long long d = std::llabs(bits(vah::eos_log(x)) - bits(std::log(x)));
        if (d > worst) worst = d;
    }
    // around 1 and the reduction boundary sqrt(1/2), where ln x is small or cancels
    for (int n = 0; n < 200000; ++n) {
        const double x = 0.5 + 1.5 * (n / 200000.0);
        const long long d = std::llabs(bits(vah::eos_log(x)) - bits(std::log(x)));
        if (d > worst) worst = d;
    }
    const bool special = vah::eos_log(1.0) == 0.0 && std::isinf(vah::eos_log(0.0)) && vah::eos_log(0.0) < 0 &&
                         std::isnan(vah::eos_log(-1.0)) && std::isnan(vah::eos_log(NAN)) &&
                         std::isinf(vah::eos_log(INFINITY)) && vah::eos_log(INFINITY) > 0 &&
                         std::fabs(vah::eos_log(4.9e-324) - std::log(4.9e-324)) < 1e-12;
    long long worst_t = 0;
    std::uniform_real_distribution<double> v(-7.5, 1.5);
    for (int n = 0; n < 1000000; ++n) {
        const double x = std::pow(10.0, v(g));
        const long long d = std::llabs(bits(vah::tanh_over_x(x)) - bits(std::tanh(x) / x));
        if (d > worst_t) worst_t = d;
    }
    const bool special_t = vah::tanh_over_x(INFINITY) == 0.0 && std::isnan(vah::tanh_over_x(NAN)) &&
                           vah::tanh_over_x(25.0) == 1.0 / 25.0;
    std::printf("%lld %d %lld %d\n", worst, special ? 1 : 0, worst_t, special_t ? 1 : 0);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_eos_log_and_tanh_ratio_against_libm(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(SRC.replace("LOG_H", '"' + os.path.join(CSRC, "vah_log.h") + '"')
                   .replace("TANH_H", '"' + os.path.join(CSRC, "vah_tanh.h") + '"'))
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    worst_log, special_log, worst_tanh, special_tanh = map(int, out)
    assert special_log == 1 and worst_log <= 2
    assert special_tanh == 1 and worst_tanh <= 6


SQRT_SRC = r'''
#define __host__
#define __device__
#include SQRT_H
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
static int64_t bits(double x) { int64_t i; std::memcpy(&i, &x, 8); return i; }
int main() {
    std::mt19937_64 g(17);
    std::uniform_real_distribution<double> ex(-760.0, 1000.0), m(1.0, 2.0), d(-1.0, 1.0);
    long long worst = 0, exact = 0;
    const int n = 2000000;
    for (int k = 0; k < n; ++k) {
        const double x = std::ldexp(m(g), (int)std::floor(ex(g)));
        const double seed = (1.0 / std::sqrt(x)) * (1.0 + std::ldexp(d(g), -20));  // a seed off by < 2^-20
        const long long e = std::llabs(bits(vah::sqrt_refine(x, seed)) - bits(std::sqrt(x)));
        if (e > worst) worst = e;
        if (e == 0) ++exact;
    }
    const bool special = vah::sqrt_refine(0.0, INFINITY) == 0.0 && std::isnan(vah::sqrt_refine(NAN, NAN)) &&
                         vah::sqrt_refine(1.0, 1.0) == 1.0 && vah::sqrt_refine(4.0, 0.5) == 2.0;
    std::printf("%lld %lld %d %d\n", worst, exact, n, special ? 1 : 0);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sqrt_refinement_is_correctly_rounded(tmp_path):
    """The refinement behind the kernels' square roots (sqrt_refine in
    gpu-vah_amd/csrc/vah_sqrt.h, seeded on the device by v_rsq_f64): from seeds
    off by up to 2^-20 it returns the C library's correctly rounded sqrt bit for
    bit over 2^-760 .. 2^1000, and 0 at 0, NaN at NaN."""
    src = tmp_path / "s.cpp"
    src.write_text(SQRT_SRC.replace("SQRT_H", '"' + os.path.join(CSRC, "vah_sqrt.h") + '"'))
    exe = tmp_path / "s"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-o", str(exe), str(src)], check=True)
    worst, exact, n, special = map(int, subprocess.run([str(exe)], check=True, capture_output=True,
                                                      text=True).stdout.split())
    assert special == 1
    assert worst == 0 and exact == n
