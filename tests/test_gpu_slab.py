"""Slab binning (csrc/hbt_slab.h): binned pairs written as records per
(block, slab) and summed in LDS instead of device-scope atomics
(hbt_binning = 1).  Against the oracle on the same inputs, in both fast forms
and with device-scope atomics (hbt_binning = 0, the default) beside it:
  - counts (bins, denominators, K_T counters): bit-exact;
  - sums: err = |S - S_ref| / max(|S_ref|, sum|.|_bin) within 1e-12 (FP64)
    and 1e-6 (precision form: in slab mode its per-pair values are FP32).
A sample whose pairs all fall into a few bins overflows the regions on the
first launch (those pairs take the direct path, results unchanged); the
next launch is sized from the measured demand."""
import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=10, long_comoving_boost=1)
FORMS = {"fp64": (_lib.HBT_PREC_FP64, 1e-12), "fp32c": (_lib.HBT_PREC_FP32C, 1e-6)}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


def plans(groups, seed, M):
    rng = hbt.Random(seed)
    Mg, ids, cs = [], [], []
    for b in groups:
        m, i, c = rng.draw_plan(b.nev, b.nev, M)
        Mg.append(m)
        ids.append(i)
        cs.append(c)
    return np.array(Mg, np.int32), np.concatenate(ids), np.concatenate(cs)


def oracle_batch(groups, plan):
    Mg, ids, cs = plan
    o = oracle.Oracle(PARAMS)
    r0 = 0
    for g, b in enumerate(groups):
        M = int(Mg[g])
        o.combine_same(b)
        o.run_mixed_plan(b, M, ids[r0:r0 + b.nev * M], cs[2 * r0:2 * (r0 + b.nev * M)])
        r0 += b.nev * M
    return o.arrays()


def run(groups, plan, prec, binning, launches=1):
    g = hbt.HBTCorrelation(dict(PARAMS, hbt_binning=binning), precision=prec,
                           kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(groups)
    g.prepare_batch(*plan)
    infos = []
    for _ in range(launches):
        g.launch_batch()
        g.sync()
        infos.append(g.slab_info())
    r = g.results()
    g.close()
    return r, infos


def check(r, a, tol, scale=1):
    assert np.array_equal(r.num_count, scale * a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, scale * a["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_kt, scale * a["cnt_num_kt"])
    assert np.array_equal(r.cnt_den_kt, scale * a["cnt_den_kt"])
    for key, fl in SUMS:
        ref = scale * a[key]
        d = np.maximum(np.abs(ref), scale * a[fl])
        d = np.where(d > 0, d, 1.0)
        err = float((np.abs(getattr(r, key) - ref) / d).max())
        assert err <= tol, (key, err)


@pytest.mark.parametrize("form", list(FORMS))
def test_slab_binning_equals_oracle_and_direct(form):
    prec, tol = FORMS[form]
    full = samples.generate(606, 120, 400.0)
    groups = [full.events(20 * k, 20) for k in range(6)]
    plan = plans(groups, 61, 10)
    print(f"slab binning vs direct atomics, {form}", flush=True)
    a = oracle_batch(groups, plan)
    rs, infos = run(groups, plan, prec, 1, launches=2)
    assert all(i["on"] for i in infos)
    check(rs, a, tol, scale=2)
    rd, infod = run(groups, plan, prec, 0)
    assert not infod[0]["on"]
    check(rd, a, tol)


@pytest.mark.parametrize("form", list(FORMS))
def test_overflowing_regions_take_the_direct_path(form):
    """Every particle near p = (0.45, 0, 0) GeV: the pairs crowd into a few
    bins of one slab, far above a uniform share, so the first launch's
    regions overflow; the results stay the oracle's, and the next launch,
    sized from that launch's demand, fits."""
    prec, tol = FORMS[form]
    rng = np.random.default_rng(5)
    base = samples.generate(607, 40, 120.0)
    n = base.n
    px = 0.45 + rng.normal(0.0, 0.004, n)
    py = rng.normal(0.0, 0.004, n)
    pz = rng.normal(0.0, 0.004, n)
    E = np.sqrt(samples.PION_MASS ** 2 + px * px + py * py + pz * pz)
    crowd = samples.Buffer(ev_off=base.ev_off, px=px, py=py, pz=pz, E=E, x=base.x, y=base.y,
                           z=base.z, t=base.t)
    groups = [crowd.events(20 * k, 20) for k in range(2)]
    plan = plans(groups, 62, 10)
    print(f"crowded sample, {form}", flush=True)
    a = oracle_batch(groups, plan)
    rs, infos = run(groups, plan, prec, 1, launches=2)
    assert infos[0]["on"] and infos[0]["direct"] > 0
    assert infos[1]["direct"] < infos[0]["direct"]
    check(rs, a, tol, scale=2)


@pytest.mark.parametrize("form", list(FORMS))
def test_autotuner_picks_the_histogram_privatisation(form):
    """hbt_binning = 2: hbt_autotune times the launch shapes with device-scope
    atomics, then the best one with slab binning, and keeps the faster; the
    results after the tuned launch are the oracle's either way."""
    prec, tol = FORMS[form]
    full = samples.generate(608, 120, 400.0)
    groups = [full.events(20 * k, 20) for k in range(6)]
    plan = plans(groups, 63, 10)
    print(f"autotuned privatisation, {form}", flush=True)
    a = oracle_batch(groups, plan)
    g = hbt.HBTCorrelation(dict(PARAMS, hbt_binning=2), precision=prec,
                           kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(groups)
    g.prepare_batch(*plan)
    tune = g.autotune()
    g.launch_batch()
    g.sync()
    info = g.slab_info()
    r = g.results()
    g.close()
    print(f"  tuning {tune}, slab binning {info['on']}", flush=True)
    assert info["on"] in (True, False)
    check(r, a, tol)

