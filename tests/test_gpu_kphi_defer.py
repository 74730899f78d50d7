"""K_phi bins at bin edges.  The device's atan2 may differ from the
reference's glibc atan2 in the last bits, so a pair whose K_phi angle lies
within a guard band of a bin edge (or of 0 and 2 pi) is decided on the host,
with glibc's atan2 on the same FP64 K_x, K_y (hbt_config.kphi_guard_prad;
default 1e-12 rad).  A wide band (2e-3 rad) sends many pairs through that
path, and the results must still be the reference's: counts bit-exact, FP64
sums within 1e-12 and precision-form sums within 1e-6 (the metric of
tests/test_gpu_parity.py), in every form: plain and fast kernels, the
batched launch, the request server, and under binding pair caps, where
each such pair is admitted in the reference's enumeration order."""
import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

WIDE = 2_000_000_000  # 2e-3 rad
MODES = {
    "plain_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_PLAIN, 1e-12),
    "fast_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12),
    "fast_fp32c": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6),
}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


def rel(a, b, floor):
    d = np.maximum(np.abs(b), floor)
    d = np.where(d > 0, d, 1.0)
    return float((np.abs(a - b) / d).max()) if a.size else 0.0


def fixture_run(name, tmp_path, prec, kern, guard):
    spec, ref, outs, dig = cases.load(name)
    par = spec["params"]
    paths = cases.write_inputs(spec, dig, str(tmp_path))
    bufs = samples.read_buffers(paths["same"], int(par["event_buffer_size"]), 211)
    params = dict(cases.hbt_params(spec), hbt_kphi_guard_prad=guard)
    g = hbt.HBTCorrelation(params, precision=prec, kernel=kern)
    rng = hbt.Random(int(par["randomSeed"]))
    for b in bufs:
        g.calculate_HBT_correlation_function(b, rng, None)
    return spec, cases.ref_arrays_as_u64(ref), g, bufs


@pytest.mark.parametrize("mode", list(MODES))
def test_wide_guard_equals_reference(mode, tmp_path):
    prec, kern, tol = MODES[mode]
    print(f"K_phi guard band 2e-3 rad, fixture_kphi_qinv, {mode}", flush=True)
    spec, ref, g, bufs = fixture_run("fixture_kphi_qinv", tmp_path, prec, kern, WIDE)
    r = g.results()
    assert g.deferred_pairs() > 0
    g.close()
    for key in ("num_count", "den", "cnt_num_ktphi", "cnt_den_ktphi", "qinv_count", "qinv_den"):
        assert np.array_equal(getattr(r, key), ref[key]), key
    o = oracle.Oracle(cases.hbt_params(spec))
    rng = oracle.OracleRng(int(spec["params"]["randomSeed"]))
    for b in bufs:
        o.calculate_buffer(b, rng)
    a = o.arrays()
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], a[fl]) <= tol, key


def test_default_guard_defers_nothing_here(tmp_path):
    spec, ref, g, _ = fixture_run("fixture_kphi_qinv", tmp_path, _lib.HBT_PREC_FP64,
                                  _lib.HBT_KERNEL_FAST, 0)
    r = g.results()
    assert g.deferred_pairs() == 0
    g.close()
    assert np.array_equal(r.num_count, ref["num_count"]) and np.array_equal(r.den, ref["den"])


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", ["fixture_cap_kphi_qinv", "synth_cap_kphi"])
def test_wide_guard_under_binding_cap_equals_reference(name, mode, tmp_path):
    """Binding per-(K_T, K_phi) caps (and the q_inv cap): with the wide band
    many pairs at K_phi edges go to the host, which places each in the cap's
    enumeration order (its arrival counted before the cut, its position
    looked up by the position pass, binned only if admitted): counts and
    counters equal the reference's golden values, sums within the mode's
    bar (HBT_correlation.cpp:402-429, :651-678)."""
    prec, kern, tol = MODES[mode]
    print(f"K_phi guard band 2e-3 rad under a binding cap, {name}, {mode}", flush=True)
    spec, ref, g, bufs = fixture_run(name, tmp_path, prec, kern, WIDE)
    r = g.results()
    deferred = g.deferred_pairs()
    g.close()
    if kern == _lib.HBT_KERNEL_FAST and prec == _lib.HBT_PREC_FP64:
        assert deferred > 0
    for key in ("num_count", "den", "cnt_num_kt", "cnt_den_kt", "cnt_num_ktphi",
                "cnt_den_ktphi", "cnt_num_qinv", "cnt_den_qinv"):
        assert np.array_equal(getattr(r, key), ref[key]), key
    if "qinv_count" in ref:
        assert np.array_equal(r.qinv_count, ref["qinv_count"])
        assert np.array_equal(r.qinv_den, ref["qinv_den"])
    o = oracle.Oracle(cases.hbt_params(spec))
    rng = oracle.OracleRng(int(spec["params"]["randomSeed"]))
    for b in bufs:
        o.calculate_buffer(b, rng)
    a = o.arrays()
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], a[fl]) <= tol, key


PI = ((211, 0.13957, 0.25, 1.0),)
KPHI_PARAMS = dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=5, KT_min=0.2, KT_max=0.6,
                   azimuthal_flag=1, n_Kphi=12, HBTrap_min=-1.0, HBTrap_max=1.0,
                   needed_number_of_pairs=1e18, number_of_mixed_events=0)


def oracle_groups(groups, params, seed):
    o = oracle.Oracle(params)
    rng = oracle.OracleRng(seed)
    for b in groups:
        o.calculate_buffer(b, rng)
    return o.arrays()


@pytest.mark.parametrize("binning", [0, 1])
@pytest.mark.parametrize("prec,tol", [(_lib.HBT_PREC_FP64, 1e-12), (_lib.HBT_PREC_FP32C, 1e-6)])
def test_wide_guard_batched(prec, tol, binning):
    """Also with slab binning (binning 1): the K_phi instance writes slab
    records for the pairs it bins while it sends the guard-band pairs to the
    host, in the same launch."""
    full = samples.generate(71, 40, samples.log_uniform_mult(30, 400), species=PI)
    groups = [full.events(g, 10) for g in range(0, 40, 10)]
    params = dict(KPHI_PARAMS, hbt_kphi_guard_prad=WIDE, hbt_binning=binning)
    print(f"K_phi guard band, batched launch, precision {prec}, binning {binning}", flush=True)
    g = hbt.HBTCorrelation(params, precision=prec, kernel=_lib.HBT_KERNEL_FAST)
    g.calculate_groups(groups, hbt.Random(7))
    r = g.results()
    # the precision form decides K_phi with FP32 atan2 outside its own 4e-6 rad
    # band, so only the FP64 form is sure to send pairs to the host here
    assert prec != _lib.HBT_PREC_FP64 or g.deferred_pairs() > 0
    assert g.slab_info()["on"] == (binning == 1)
    g.close()
    a = oracle_groups(groups, KPHI_PARAMS, 7)
    assert np.array_equal(r.num_count, a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, a["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_ktphi, a["cnt_num_ktphi"].astype(np.uint64))
    for key, fl in SUMS:
        assert rel(getattr(r, key), a[key], a[fl]) <= tol, key


def test_wide_guard_server():
    buf = samples.generate(72, 30, 250.0, species=PI)
    params = dict(KPHI_PARAMS, hbt_kphi_guard_prad=WIDE)
    print("K_phi guard band, request server", flush=True)
    srv = hbt.HBTServer(params, precision=_lib.HBT_PREC_FP64, hbm_bytes=1 << 30)
    rid = srv.submit(buf, 10, 9)
    srv.wait(rid, 120.0)
    r = srv.result(rid)
    st = srv.stats()
    srv.release(rid)
    srv.close()
    assert st["deferred"] > 0
    a = oracle_groups([buf.events(g, 10) for g in range(0, 30, 10)], KPHI_PARAMS, 9)
    assert np.array_equal(r.num_count, a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, a["den"].astype(np.uint64))
    for key, fl in SUMS:
        assert rel(getattr(r, key), a[key], a[fl]) <= 1e-12, key
