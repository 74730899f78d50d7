"""GPU parity on the shapes of the batched-uneven and large-histogram
workloads, against the CPU oracle, at sizes the oracle finishes in seconds.

  - two species (pi+ and K+, 6:1, K+ with T = 0.30 GeV) from one sample,
    analysed as separate identical-particle histograms, multiplicities
    log-uniform, 16 K_phi bins about each buffer's event plane;
  - the large-histogram grid: q 61^3 on [-0.2, 0.2] GeV, 8 K_T bins on
    [0.15, 1.0] GeV, 16 K_phi bins, 20 rotations per event.

Bars (written here): counts (numerator, denominator, K_T x K_phi counters)
bit-exact; FP64 sums within 1e-12 and precision-form sums within 1e-6 of
the oracle under err = |S - S_ref| / max(|S_ref|, sum|.|_bin).  With K_phi
bins the precision form evaluates pairs with the literal FP64 cascade.
"""
import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PI_K = ((211, 0.13957, 0.25, 6.0), (321, 0.49368, 0.30, 1.0))
MODES = {"fast_fp64": (_lib.HBT_PREC_FP64, 1e-12), "fast_fp32c": (_lib.HBT_PREC_FP32C, 1e-6)}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


def rel(a, b, floor):
    d = np.maximum(np.abs(b), floor)
    d = np.where(d > 0, d, 1.0)
    return float((np.abs(a - b) / d).max()) if a.size else 0.0


def check(groups, params, seed, prec, tol):
    g = hbt.HBTCorrelation(params, precision=prec, kernel=_lib.HBT_KERNEL_FAST)
    g.calculate_groups(groups, hbt.Random(seed))
    r = g.results()
    g.close()
    o = oracle.Oracle(params)
    rng = oracle.OracleRng(seed)
    for b in groups:
        o.calculate_buffer(b, rng)
    a = o.arrays()
    assert a["num_count"].sum() > 0 and a["den"].sum() > 0
    assert np.array_equal(r.num_count, a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, a["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_ktphi, a["cnt_num_ktphi"].astype(np.uint64))
    assert np.array_equal(r.cnt_den_ktphi, a["cnt_den_ktphi"].astype(np.uint64))
    for key, fl in SUMS:
        assert rel(getattr(r, key), a[key], a[fl]) <= tol, key
    return r


@pytest.mark.parametrize("mode", list(MODES))
def test_two_species_uneven_kphi(mode):
    prec, tol = MODES[mode]
    full = samples.generate(61, 60, samples.log_uniform_mult(20, 600), species=PI_K)
    params = dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
                  azimuthal_flag=1, n_Kphi=16, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18, number_of_mixed_events=0)
    for monval in (211, 321):
        buf = full.select(monval)
        print(f"two species: monval {monval}, {buf.nev} events, {buf.n} particles, {mode}",
              flush=True)
        groups = [buf.events(g, min(20, buf.nev - g)) for g in range(0, buf.nev, 20)]
        check(groups, params, 100 + monval, prec, tol)


@pytest.mark.parametrize("mode", list(MODES))
def test_large_histogram_grid(mode):
    prec, tol = MODES[mode]
    full = samples.generate(62, 6, 1000.0)
    params = dict(qnpts=61, q_min=-0.2, q_max=0.2, n_KT=9, KT_min=0.15, KT_max=1.0,
                  azimuthal_flag=1, n_Kphi=16, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18, number_of_mixed_events=20)
    print(f"large grid: 6 events, {full.n} particles, groups of 3, {mode}", flush=True)
    groups = [full.events(g, 3) for g in range(0, 6, 3)]
    check(groups, params, 77, prec, tol)


def test_kphi_fold_mod_pi():
    """The 8 bins of Phi_K mod pi are the sums of the 16 bins k and k+8 of
    the reference's [0, 2pi) binning (exact: folding follows the indexing)."""
    params = dict(qnpts=11, q_min=-0.15, q_max=0.15, n_KT=3, KT_min=0.2, KT_max=0.6,
                  azimuthal_flag=1, n_Kphi=16, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18)
    buf = samples.generate(63, 4, 300.0)
    g = hbt.HBTCorrelation(params, precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_FAST)
    g.calculate_groups([buf], hbt.Random(5))
    r = g.results()
    g.close()
    f = hbt.fold_kphi_mod_pi(r, params)
    q3 = 11 ** 3
    nc = r.num_count.reshape(3, 16, q3)
    assert np.array_equal(f.num_count.reshape(3, 8, q3), nc[:, :8] + nc[:, 8:])
    assert np.array_equal(f.cnt_num_ktphi.reshape(3, 8),
                          r.cnt_num_ktphi.reshape(3, 16)[:, :8] + r.cnt_num_ktphi.reshape(3, 16)[:, 8:])
    assert np.allclose(f.num.reshape(3, 8, q3), r.num.reshape(3, 16, q3)[:, :8]
                       + r.num.reshape(3, 16, q3)[:, 8:], rtol=0, atol=0)
