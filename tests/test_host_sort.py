"""The upload path's radix sort (csrc/host_sort.h) against std::stable_sort
by key on random keys with ties, +-0.0, +-inf and subnormals: the same
permutation (ties in input order) and the same key values.  Built with g++
into a temporary directory; no GPU."""
import shutil
import subprocess
from pathlib import Path

import pytest

CSRC = Path(__file__).resolve().parents[1] / "hadronic_afterburner_toolkit_amd" / "csrc"

PROGRAM = r"""
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include "host_sort.h"
int main() {
    std::mt19937_64 g(3);
    std::uniform_real_distribution<double> u(-3, 3);
    long long checked = 0;
    for (int t = 0; t < 200; t++) {
        const int n = t % 7 == 0 ? (int)(t % 3) : (int)(g() % 20000);
        std::vector<std::pair<double, int>> v(n);
        for (int k = 0; k < n; k++) {
            double x = u(g);
            switch (g() % 10) {
                case 0: x = 0.0; break;
                case 1: x = -0.0; break;
                case 2: x = std::round(x * 4) / 4; break;  // many ties
                case 3: if (t % 2) x = g() % 2 ? INFINITY : -INFINITY; break;
                case 4: x = 1e-310 * (double)(g() % 3) * (g() % 2 ? 1 : -1); break;
                case 5: if (t % 5 == 0) x = 1.5; break;  // long runs of one key
                default: break;
            }
            v[k] = {x, k};
        }
        auto w = v;
        std::stable_sort(w.begin(), w.end(),
                         [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                             return a.first < b.first;
                         });
        std::vector<std::pair<uint64_t, int>> ra, rb;
        hbt::radix_sort_keys(v, ra, rb);
        for (int k = 0; k < n; k++) {
            if (v[k].second != w[k].second || !(v[k].first == w[k].first)) {
                std::printf("mismatch case %d position %d\n", t, k);
                return 1;
            }
        }
        checked += n;
    }
    std::printf("ok %lld\n", checked);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_radix_sort_equals_stable_sort(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{CSRC}", str(src), "-o", str(exe)],
                   check=True, timeout=120)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
