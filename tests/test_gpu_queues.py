"""Per-XCD work queues of the fast kernel (build_items / assign_queues in
csrc/hbt_engine.hip): every item of a group goes to one queue, groups are
dealt to queues by load, and waves that find their own queue empty take
items from the others.  Batches of uneven groups exercise the split and the
stealing: a power-of-two count (4 groups: 4 queues), a count that is not
(3 groups: 2 queues, one holding two groups), one large group with many
small ones (its queue is drained by every XCD), and a single group (one
queue).  Against the CPU oracle on the same batch:
counts bit-exact, FP64 sums within 1e-12, the precision form's within 1e-6,
under err = |S - S_ref| / max(|S_ref|, sum|.|_bin)."""
import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=31, q_min=-0.12, q_max=0.12, n_KT=5, KT_min=0.2, KT_max=0.6,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=6, long_comoving_boost=1)
FORMS = {"fp64": (_lib.HBT_PREC_FP64, 1e-12), "fp32c": (_lib.HBT_PREC_FP32C, 1e-6)}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))


def batch_groups(sizes, mult, seed):
    """Groups of the given event counts from one sample."""
    full = samples.generate(seed, sum(sizes), mult)
    out, e = [], 0
    for n in sizes:
        out.append(full.events(e, n))
        e += n
    return out


def run_and_check(groups, prec, tol, seed, nqueues):
    rng = hbt.Random(seed)
    Mg, ids, cs = [], [], []
    for b in groups:
        m, i, c = rng.draw_plan(b.nev, b.nev, PARAMS["number_of_mixed_events"])
        Mg.append(m)
        ids.append(i)
        cs.append(c)
    g = hbt.HBTCorrelation(PARAMS, precision=prec, kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(groups)
    g.prepare_batch(np.array(Mg, np.int32), np.concatenate(ids), np.concatenate(cs))
    g.set_tuning(1, 256, 1, 8)
    qoff = g.batch_queues()
    assert len(qoff) - 1 == nqueues, qoff
    assert qoff[0] == 0 and qoff[-1] == g.batch_info()[2] and all(np.diff(qoff) > 0)
    g.launch_batch()
    g.sync()
    r = g.results()
    g.close()
    o = oracle.Oracle(PARAMS)
    for k, b in enumerate(groups):
        o.combine_same(b)
        o.run_mixed_plan(b, int(Mg[k]), ids[k], cs[k])
    a = o.arrays()
    assert np.array_equal(r.num_count, a["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, a["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_kt, a["cnt_num_kt"])
    assert np.array_equal(r.cnt_den_kt, a["cnt_den_kt"])
    for key, fl in SUMS:
        d = np.maximum(np.abs(a[key]), a[fl])
        d = np.where(d > 0, d, 1.0)
        err = float((np.abs(getattr(r, key) - a[key]) / d).max())
        assert err <= tol, (key, err)


CASES = {  # event counts per group, mean multiplicity, queues
    "four_groups": ([8, 8, 8, 8], 300.0, 4),
    "three_groups": ([6, 10, 4], 300.0, 2),
    "one_large_many_small": ([30, 2, 2, 2, 2, 2, 2, 2, 2], 250.0, 8),
    "single_group": ([12], 300.0, 1),
}


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("case", list(CASES))
def test_queues_equal_oracle(case, form):
    sizes, mult, nq = CASES[case]
    prec, tol = FORMS[form]
    print(f"queues: {case} ({sizes} events x Poisson({mult})), {form}", flush=True)
    groups = batch_groups(sizes, mult, 800 + len(sizes))
    run_and_check(groups, prec, tol, 90 + len(sizes), nq)
