"""GPU parity at the headline workload's size: group 0 of the bench workload
(2000 pi+ events Poisson(600), seed 2025, groups of 50 consecutive events:
~30k particles, ~4.5e8 same-event pairs; 10 rotated partners per event from
Random(2026), the first draws of the bench's plan stream), q 41^3 on
[-0.15, 0.15] GeV, K_T edges 0.2..0.7 GeV, LCMS boost, |y| < 1, no cap.

It runs through the batched path the bench times (upload_groups,
prepare_batch, one launch) in the plain FP64, fast FP64 and precision forms,
the fast forms also with slab binning (hbt_binning = 1: same-event pairs as
records summed per q_out plane in LDS), against the oracle on the same
inputs and plan:
  - counts (numerator, denominator, K_T counters): bit-exact;
  - sums: err = |S - S_ref| / max(|S_ref|, sum|.|_bin) within 1e-12 (FP64)
    and 1e-6 (precision form); the max and RMS of err over the bins with
    pairs are printed for every field.
The same group is also analysed in 16 K_phi bins (atomics and slab binning,
also the K_T x K_phi counters bit-exact).
"""
import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=10, long_comoving_boost=1)
MODES = {"plain_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_PLAIN, 1e-12, 0),
         "fast_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 0),
         "fast_fp32c": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 0),
         "fast_fp64_slab": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 1),
         "fast_fp32c_slab": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 1)}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


@pytest.fixture(scope="module")
def headline_group(oracle_mod):
    full = samples.generate(2025, 2000, 600.0)
    g0 = full.events(0, 50)
    M, ids, cs = hbt.Random(2026).draw_plan(g0.nev, g0.nev, 10)
    print(f"headline group 0: {g0.nev} events, {g0.n} particles, M={M}; oracle ...", flush=True)
    o = oracle_mod.Oracle(PARAMS)
    o.combine_same(g0)
    o.run_mixed_plan(g0, M, ids, cs)
    ref = o.arrays()
    print(f"  oracle: {ref['same_st'][0]:.3e} same + {ref['mixed_st'][0]:.3e} mixed pairs, "
          f"{int(ref['num_count'].sum())} / {int(ref['den'].sum())} binned", flush=True)
    return g0, (M, ids, cs), ref


@pytest.mark.parametrize("mode", list(MODES))
def test_headline_group_vs_oracle(headline_group, mode):
    g0, (M, ids, cs), ref = headline_group
    prec, kern, tol, binning = MODES[mode]
    print(f"headline group, {mode}", flush=True)
    g = hbt.HBTCorrelation(dict(PARAMS, hbt_binning=binning), precision=prec, kernel=kern)
    g.upload_groups([g0])
    g.prepare_batch(np.array([M], np.int32), ids, cs)
    g.launch_batch()
    g.sync()
    r = g.results()
    es, em, _ = g.batch_info()
    slab = g.slab_info()
    g.close()
    assert slab["on"] == (binning == 1)
    if binning:
        print(f"  slab binning: {slab['direct']} same-event pairs added directly", flush=True)
    assert es == ref["same_st"][0] and em == ref["mixed_st"][0]
    assert np.array_equal(r.num_count, ref["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, ref["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_kt, ref["cnt_num_kt"])
    assert np.array_equal(r.cnt_den_kt, ref["cnt_den_kt"])
    report = []
    for key, fl in SUMS:
        a, b = getattr(r, key), ref[key]
        d = np.maximum(np.abs(b), ref[fl])
        sel = d > 0
        err = np.abs(a - b)[sel] / d[sel]
        mx = float(err.max()) if err.size else 0.0
        rms = float(np.sqrt(np.mean(err * err))) if err.size else 0.0
        report.append(f"{key} max {mx:.2e} rms {rms:.2e}")
        assert mx <= tol, (key, mx)
    print("  per-bin error: " + "; ".join(report), flush=True)


# The same group analysed in 16 K_phi bins about its event plane: slab
# binning with one slab per (K_T, K_phi, q_out) plane (6 x 16 x 41 = 3936),
# and device-scope atomics, against the oracle (also the K_T x K_phi
# counters, bit-exact).
PARAMS_KPHI = dict(PARAMS, azimuthal_flag=1, n_Kphi=16)
MODES_KPHI = {k: MODES[k] for k in ("fast_fp32c", "fast_fp64_slab", "fast_fp32c_slab")}


@pytest.fixture(scope="module")
def headline_group_kphi(oracle_mod, headline_group):
    g0, (M, ids, cs), _ = headline_group
    print("headline group 0 with 16 K_phi bins; oracle ...", flush=True)
    o = oracle_mod.Oracle(PARAMS_KPHI)
    o.set_psi_ref(oracle_mod.psi_ref(g0, 2))  # the group's event plane, as per buffer
    o.combine_same(g0)
    o.run_mixed_plan(g0, M, ids, cs)
    return o.arrays()


@pytest.mark.parametrize("mode", list(MODES_KPHI))
def test_headline_group_kphi_vs_oracle(headline_group, headline_group_kphi, mode):
    g0, (M, ids, cs), _ = headline_group
    ref = headline_group_kphi
    prec, kern, tol, binning = MODES_KPHI[mode]
    print(f"headline group, 16 K_phi bins, {mode}", flush=True)
    g = hbt.HBTCorrelation(dict(PARAMS_KPHI, hbt_binning=binning), precision=prec, kernel=kern)
    g.upload_groups([g0])
    g.prepare_batch(np.array([M], np.int32), ids, cs)
    g.launch_batch()
    g.sync()
    r = g.results()
    slab = g.slab_info()
    g.close()
    assert slab["on"] == (binning == 1)
    assert np.array_equal(r.num_count, ref["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, ref["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_ktphi, ref["cnt_num_ktphi"].astype(np.uint64))
    for key, fl in SUMS:
        a, b = getattr(r, key), ref[key]
        d = np.maximum(np.abs(b), ref[fl])
        sel = d > 0
        err = np.abs(a - b)[sel] / d[sel]
        assert (float(err.max()) if err.size else 0.0) <= tol, key
