"""An upload issued while a batch is still running (hbt_launch_batch without
hbt_sync) waits for that batch: the batch's results are those of the data it
was launched on, and the new buffer is the one the next batch reads.

Bars (written here): counts bit-exact, FP64 sums within 1e-12 under
err = |S - S_ref| / max(|S_ref|, n_bin * max|term|) per bin with pairs (the
order of the atomic additions differs between runs).
"""
import numpy as np
import pytest

from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=10, long_comoving_boost=1)


def plans(groups, seed, M):
    rng = hbt.Random(seed)
    Mg, ids, cs = [], [], []
    for b in groups:
        m, i, c = rng.draw_plan(b.nev, b.nev, M)
        Mg.append(m)
        ids.append(i)
        cs.append(c)
    return np.array(Mg, np.int32), np.concatenate(ids), np.concatenate(cs)


def workload(seed):
    full = samples.generate(seed, 120, 400.0)
    return [full.events(g * 20, 20) for g in range(6)]


def same(r, ref):
    assert np.array_equal(r.num_count, ref.num_count)
    assert np.array_equal(r.den, ref.den)
    assert int(ref.num_count.sum()) > 0 and int(ref.den.sum()) > 0
    m = ref.num_count > 0
    n = ref.num_count[m].astype(np.float64)
    # sum|term| of a bin is at most its pair count times the largest |term|:
    # 1 for cos, q_max + dq/2 for the q sums
    for key, scale in (("num", 1.0), ("q_out_sum", 0.16), ("q_side_sum", 0.16),
                       ("q_long_sum", 0.16)):
        a, b = getattr(r, key)[m], getattr(ref, key)[m]
        floor = np.maximum(np.abs(b), scale * n)
        assert float((np.abs(a - b) / floor).max()) <= 1e-12, key


LAUNCHES = 50  # queued back to back, so the upload finds them still running


def fresh(groups, plan):
    g = hbt.HBTCorrelation(PARAMS, precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(groups)
    g.prepare_batch(*plan)
    for _ in range(LAUNCHES):
        g.launch_batch()
    g.sync()
    r = g.results()
    g.close()
    return r


def test_upload_waits_for_running_batch():
    A, B = workload(71), workload(72)
    pa, pb = plans(A, 5, 10), plans(B, 6, 10)
    ref = {"A": fresh(A, pa), "B": fresh(B, pb)}
    data = {"A": (A, pa), "B": (B, pb)}
    g = hbt.HBTCorrelation(PARAMS, precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(A)
    g.prepare_batch(*pa)
    for cur, nxt in (("A", "B"), ("B", "A")):
        print(f"upload of {nxt} while batches on {cur} run", flush=True)
        g.reset()
        for _ in range(LAUNCHES):
            g.launch_batch()
        g.upload_groups(data[nxt][0])  # no sync before it
        g.sync()
        same(g.results(), ref[cur])
        g.prepare_batch(*data[nxt][1])
        g.reset()
        for _ in range(LAUNCHES):
            g.launch_batch()
        g.sync()
        same(g.results(), ref[nxt])
    g.close()
