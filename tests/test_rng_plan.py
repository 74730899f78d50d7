"""The mixing plan's random stream (CPU only).

The reference draws partner ids and rotation angles from std::mt19937 with
uniform_int_distribution<int>() and uniform_real_distribution<double>(0,1)
(Random.h:11-25) in a fixed order (HBT_correlation.cpp:206-215, :495-497).
Three implementations must agree draw for draw: the oracle's C restatement,
the engine's (libstdc++ through hipcc), and libstdc++ compiled by g++ as the
reference is.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import hbt

GXX_PROG = r"""
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    std::mt19937 g(%d);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::uniform_int_distribution<int> k;
    for (int i = 0; i < 2000; i++) {
        int a = k(g);
        double b = u(g);
        std::printf("%%d %%a\n", a, b);
    }
}
"""


@pytest.fixture(scope="module")
def gxx_draws(tmp_path_factory):
    d = tmp_path_factory.mktemp("gxx")
    out = {}
    for seed in (0, 7, 123456789):
        src = d / f"r{seed}.cpp"
        src.write_text(GXX_PROG % seed)
        exe = d / f"r{seed}"
        subprocess.run(["g++", "-O2", "-std=c++11", str(src), "-o", str(exe)], check=True)
        lines = subprocess.run([str(exe)], capture_output=True, text=True,
                               check=True).stdout.split("\n")
        out[seed] = [(int(l.split()[0]), float.fromhex(l.split()[1])) for l in lines if l]
    return out


def test_first_draws_seed0(oracle_mod):
    """std::mt19937(0): rand_int_uniform() then rand_uniform() on one stream."""
    r = oracle.OracleRng(0)
    assert r.rand_int_uniform() == 1178568022
    assert r.rand_uniform() == 0.71518936510962539
    r = oracle.OracleRng(0)
    assert r.rand_uniform() == 0.59284461651668263


@pytest.mark.parametrize("seed", [0, 7, 123456789])
def test_oracle_rng_equals_libstdcxx(oracle_mod, gxx_draws, seed):
    r = oracle.OracleRng(seed)
    for a, b in gxx_draws[seed]:
        assert r.rand_int_uniform() == a
        assert r.rand_uniform() == b


@pytest.mark.parametrize("seed", [0, 7, 123456789])
def test_engine_rng_equals_libstdcxx(gxx_draws, seed):
    r = hbt.Random(seed)
    for a, b in gxx_draws[seed]:
        assert r.rand_int_uniform() == a
        assert r.rand_uniform() == b


@pytest.mark.parametrize("nev,mixed_nev,M", [(10, 10, 0), (1, 1, 0), (2, 2, 0), (7, 12, 0),
                                             (50, 50, 0), (20, 20, 10)])
def test_plan_engine_equals_oracle(oracle_mod, nev, mixed_nev, M):
    a = oracle.OracleRng(99)
    b = hbt.Random(99)
    for _ in range(3):  # the stream continues across buffers
        Ma, ia, ca = a.draw_plan(nev, mixed_nev, M)
        Mb, ib, cb = b.draw_plan(nev, mixed_nev, M)
        assert Ma == Mb
        assert np.array_equal(ia, ib)
        assert np.array_equal(ca.view(np.uint64), cb.view(np.uint64))
        if mixed_nev > 1:  # an event never mixes with itself (:210)
            rows = np.repeat(np.arange(nev), Ma)
            assert not np.any(ia == rows)


def test_plan_single_event_mixes_with_itself(oracle_mod):
    """mixed_nev == 1: the rejection loop is skipped (HBT_correlation.cpp:210)."""
    M, ids, cs = oracle.OracleRng(3).draw_plan(1, 1)
    assert M == 1 and list(ids) == [0]
    c, s = cs[0], cs[1]
    assert abs(c * c + s * s - 1.0) < 1e-15
