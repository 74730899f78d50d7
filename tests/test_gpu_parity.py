"""GPU parity: the HIP engine against the reference's golden dumps and the
CPU oracle, through the C ABI.

Bars (tolerances written here):
  - counts (numerator, denominator, q_inv, pair counters): bit-exact;
  - FP64 modes (plain and fast kernels): every FP64 sum within 1e-12 of the
    reference under err = |S_gpu - S_ref| / max(|S_ref|, sum|.|_bin) (the
    floor sum|cos| / sum|q| of each bin comes from the oracle, which is
    bit-exact with the reference);
  - precision form (FP32C): counts bit-exact, sums within 1e-6 under the
    same metric.
The golden cases include binding pair caps (3D per K_T and per K_T x K_phi,
and the q_inv cap), whose admitted pairs depend on the enumeration order;
those calls run the engine's ordered path.  The fast forms also run with
slab binning (hbt_binning = 1), buffer after buffer through
calculate_HBT_correlation_function, each launch's record regions sized from
the previous buffer's demand (the q_inv cases bin with atomics; the K_phi
cases write records per K_T x K_phi x q_out plane).
"""
import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

MODES = {
    "plain_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_PLAIN, 1e-12, 0),
    "fast_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 0),
    "fast_fp32c": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 0),
    "fast_fp64_slab": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 1),
    "fast_fp32c_slab": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 1),
}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


def rel_err(gpu, ref, floor):
    d = np.maximum(np.abs(ref), floor)
    d = np.where(d > 0, d, 1.0)
    return np.abs(gpu - ref) / d


def run_engine_case(name, tmp_path, prec, kern, binning=0):
    spec, ref, outs, dig = cases.load(name)
    par = spec["params"]
    paths = cases.write_inputs(spec, dig, str(tmp_path))
    ebs = int(par["event_buffer_size"])
    bufs = samples.read_buffers(paths["same"], ebs, 211)
    mixb = samples.read_buffers(paths["mixed"], ebs, 211) if "mixed" in paths else None
    g = hbt.HBTCorrelation(dict(cases.hbt_params(spec), hbt_binning=binning), precision=prec,
                           kernel=kern)
    rng = hbt.Random(int(par["randomSeed"]))
    for k, b in enumerate(bufs):
        g.calculate_HBT_correlation_function(b, rng, mixb[k] if mixb else None)
    return spec, cases.ref_arrays_as_u64(ref), g, bufs, mixb


def oracle_floors(spec, bufs, mixb):
    o = oracle.Oracle(cases.hbt_params(spec))
    rng = oracle.OracleRng(int(spec["params"]["randomSeed"]))
    for k, b in enumerate(bufs):
        o.calculate_buffer(b, rng, mixb[k] if mixb else None)
    return o.arrays()


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", cases.CASE_NAMES)
def test_engine_vs_reference_golden(name, mode, tmp_path):
    prec, kern, tol, binning = MODES[mode]
    print(f"case {name} mode {mode}", flush=True)
    spec, ref, g, bufs, mixb = run_engine_case(name, tmp_path, prec, kern, binning)
    r = g.results()
    assert np.array_equal(r.num_count, ref["num_count"]), "num_count"
    assert np.array_equal(r.den, ref["den"]), "den"
    for key in ("cnt_num_kt", "cnt_den_kt", "cnt_num_qinv", "cnt_den_qinv"):
        assert np.array_equal(getattr(r, key), ref[key]), key
    if "cnt_num_ktphi" in ref:
        assert np.array_equal(r.cnt_num_ktphi, ref["cnt_num_ktphi"])
        assert np.array_equal(r.cnt_den_ktphi, ref["cnt_den_ktphi"])
    o = oracle_floors(spec, bufs, mixb)
    for key, fl in SUMS:
        err = rel_err(getattr(r, key), ref[key], o[fl])
        assert float(err.max()) <= tol, (key, float(err.max()))
    if "qinv_count" in ref:
        assert np.array_equal(r.qinv_count, ref["qinv_count"])
        assert np.array_equal(r.qinv_den, ref["qinv_den"])
        assert np.allclose(r.qinv_sum, ref["qinv_sum"], rtol=1e-12, atol=1e-12)
        qfloor = np.abs(ref["qinv_count"]).astype(float)  # |cos| <= 1 per pair
        assert float(rel_err(r.qinv_num, ref["qinv_num"], qfloor).max()) <= 1e-12


@pytest.mark.parametrize("mode", list(MODES))
def test_config0_scale_vs_oracle(mode):
    """Config-0 source model (Poisson(150) pi+, groups of 10 events, q grid
    41^3 on [-0.15,0.15], K_T edges 0.2..0.7), 6 groups, reference mixing
    rule, against the oracle on the same inputs and random stream."""
    prec, kern, tol, binning = MODES[mode]
    params = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
                  HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18)
    full = samples.generate(20240, 60, 150.0)
    g = hbt.HBTCorrelation(dict(params, hbt_binning=binning), precision=prec, kernel=kern)
    o = oracle.Oracle(params)
    rg, ro = hbt.Random(77), oracle.OracleRng(77)
    for grp in range(6):
        b = full.events(10 * grp, 10)
        g.calculate_HBT_correlation_function(b, rg)
        o.calculate_buffer(b, ro)
    r, a = g.results(), o.arrays()
    assert np.array_equal(r.num_count, a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, a["den"].astype(np.uint64))
    for key, fl in SUMS:
        err = rel_err(getattr(r, key), a[key], a[fl])
        assert float(err.max()) <= tol, (key, float(err.max()))
    s, m = g.stage_counts()
    assert int(s[0]) == a["same_st"][0] and int(m[0]) == a["mixed_st"][0]
    assert int(s[4]) == a["same_st"][4] and int(m[4]) == a["mixed_st"][4]
    if kern == _lib.HBT_KERNEL_PLAIN:
        assert [int(v) for v in s] == a["same_st"]
        assert [int(v) for v in m] == a["mixed_st"]


def test_modes_agree_on_uneven_events():
    """Multiplicities log-uniform on [20, 600], one buffer of 12 events:
    every mode (slab binning included) gives identical counts."""
    params = dict(qnpts=31, n_KT=5, KT_min=0.2, KT_max=0.7, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18)
    b = samples.generate(99, 12, samples.log_uniform_mult(20, 600))
    res = {}
    for mode, (prec, kern, _, binning) in MODES.items():
        g = hbt.HBTCorrelation(dict(params, hbt_binning=binning), precision=prec, kernel=kern)
        g.calculate_HBT_correlation_function(b, hbt.Random(1))
        res[mode] = g.results()
    for mode in ("fast_fp64", "fast_fp32c", "fast_fp64_slab", "fast_fp32c_slab"):
        assert np.array_equal(res[mode].num_count, res["plain_fp64"].num_count)
        assert np.array_equal(res[mode].den, res["plain_fp64"].den)


@pytest.mark.parametrize("mode", ["fast_fp64", "fast_fp32c", "plain_fp64"])
def test_batch_equals_sequential_buffers(mode):
    """The batched form (all groups resident, one persistent launch) gives the
    same result as the groups processed one after another, and the hipGraph
    replays add up exactly."""
    prec, kern, _, _ = MODES[mode]
    params = dict(qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18)
    full = samples.generate(555, 24, samples.log_uniform_mult(20, 400))
    groups = [full.events(0, 5), full.events(5, 9), full.events(14, 1), full.events(15, 9)]
    seq = hbt.HBTCorrelation(params, precision=prec, kernel=kern)
    rs = hbt.Random(3)
    for b in groups:
        seq.calculate_HBT_correlation_function(b, rs)
    bat = hbt.HBTCorrelation(params, precision=prec, kernel=kern)
    bat.calculate_groups(groups, hbt.Random(3))
    a, b = seq.results(), bat.results()
    assert np.array_equal(a.num_count, b.num_count)
    assert np.array_equal(a.den, b.den)
    assert np.allclose(a.num, b.num, rtol=1e-12, atol=1e-12)
    # replay the prepared batch twice more: exactly 3x the counts
    bat.launch_batch()
    bat.sync()
    bat.launch_batch()
    bat.sync()
    c = bat.results()
    assert np.array_equal(c.num_count, 3 * b.num_count)
    assert np.array_equal(c.den, 3 * b.den)
    # and the oracle agrees with the sequential run
    o = oracle.Oracle(params)
    ro = oracle.OracleRng(3)
    for g in groups:
        o.calculate_buffer(g, ro)
    assert np.array_equal(a.num_count, o.arrays()["num_count"].astype(np.uint64))
    assert np.array_equal(a.den, o.arrays()["den"].astype(np.uint64))


def test_autotune_keeps_results():
    params = dict(qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8, HBTrap_min=-1.0, HBTrap_max=1.0,
                  needed_number_of_pairs=1e18, number_of_mixed_events=4)
    full = samples.generate(556, 16, 300.0)
    groups = [full.events(0, 8), full.events(8, 8)]
    ref = None
    for tune in (False, True):
        g = hbt.HBTCorrelation(params, precision=_lib.HBT_PREC_FP32C, kernel=_lib.HBT_KERNEL_FAST)
        g.upload_groups(groups)
        rng = hbt.Random(9)
        plans = [rng.draw_plan(b.nev, b.nev, 4) for b in groups]
        g.prepare_batch([p[0] for p in plans], np.concatenate([p[1] for p in plans]),
                        np.concatenate([p[2] for p in plans]))
        if tune:
            t = g.autotune()
            assert t["R"] in (1, 2, 4) and t["ms"] > 0
            z = g.results()
            assert int(z.num_count.sum()) == 0  # autotune clears the histograms
        g.launch_batch()
        g.sync()
        r = g.results()
        if ref is None:
            ref = r
        else:
            assert np.array_equal(r.num_count, ref.num_count)
            assert np.array_equal(r.den, ref.den)
