"""The host-side generators (initial conditions, the tabulated EOS) built for
the host alone with AddressSanitizer and UndefinedBehaviorSanitizer from the
library's own sources, and run over the lattice shapes the benchmarks and the
reference's initial-condition types use: no out-of-bounds access, no
undefined behaviour, finite positive energy densities."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "gpu-vah_amd", "host")

MAIN = r'''
#include "vah.h"
#include "vah_host.h"
#include <cmath>
#include <cstdio>
#include <vector>
static bool good(const std::vector<double> &e) {
    for (double v : e) if (!(v > 0.0) || !std::isfinite(v)) return false;
    return true;
}
int main() {
    int bad = 0;
    const int shapes[][3] = {{81, 81, 1}, {201, 201, 5}, {17, 31, 2}, {1, 1, 1}, {261, 101, 1}};
    for (auto &s : shapes) {
        vah_lattice L = {s[0], s[1], s[2], 10, 0.1, 0.1, 0.1, 0.01};
        std::vector<double> e((size_t)s[0] * s[1] * s[2]);
        vah_glauber g = {208, 6.62, 0.546, 70.0, 6.0, 0.5, 60.0 / 0.197326938, 0.5, 42ull, 1.5, 1.4, 1e-3};
        int nw = 0;
        if (vah_ic_mc_glauber(&L, &g, e.data(), &nw) != 0 || !good(e)) ++bad;
        vah_glauber o = {197, 6.38, 0.535, 42.0, 7.0, 0.5, 30.0 / 0.197326938, 0.5, 0ull, 1.5, 1.4, 1e-3};
        if (vah_ic_optical_glauber(&L, &o, e.data()) != 0 || !good(e)) ++bad;
    }
    std::vector<double> p(10000), T(10000);
    double lne0 = 0, h = 0;
    if (vah_eos_lqcd_table(10000, 1e-8, 1e4, p.data(), T.data(), &lne0, &h) != 0) ++bad;
    for (int i = 0; i < 10000; ++i) if (!(p[i] > 0 && T[i] > 0)) { ++bad; break; }
    std::printf("%d\n", bad);
    return bad != 0;
}
'''


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_generators_under_sanitizers(tmp_path):
    src = tmp_path / "main.cpp"
    src.write_text(MAIN)
    exe = tmp_path / "t"
    cmd = ["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src),
           os.path.join(HOST, "initial_conditions.cpp"), os.path.join(HOST, "eos_table.cpp")]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "0"
