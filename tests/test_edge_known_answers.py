"""Known answers of single hand-placed pairs on the cascade's decision edges,
run through the oracle one pair at a time (CPU).

Each pair of tests/golden/edge_inputs.py is its own two-event buffer:
same-event, it is the buffer's only pair; mixed, event 0 is mixed with event
1 under the rotation (cos, sin) = (1, 0), which leaves p_x and p_y unchanged
(p_x * 1 - p_y * 0 = p_x), so the mixed pair is the same pair.  The expected
outcome of every pair follows from the reference's comparisons alone
(HBT_correlation.cpp):

* same-event window: keep lo <= v <= hi (:363-364, :373-374, :392-393);
* mixed window: keep lo <= v < hi (:612-613, :621-622, :640-641);
* K_T: keep KT_min^2 <= K_perp^2 <= KT_max^2 (:318, :582), index
  int((K_perp - KT_min)/dKT), n_KT - 1 at K_perp = KT_max (:324);
* bin index int((v - (q_min - dq/2))/dq) (:369, :618).

The reference's golden run on the same buffers (edge_3d) pins the same-event
side bit for bit; the mixed side, which the reference only reaches through
random rotations, is pinned here.
"""
import math

import numpy as np
import pytest

import edge_inputs as E
import oracle

P = E.EDGE_PARAMS
D = E.derived(P)


def _one_pair_buffer(a, b):
    return E._buffer([(a, b)])


def _run(a, b):
    """(same-event arrays, mixed arrays) of one designed pair."""
    buf = _one_pair_buffer(a, b)
    o_same = oracle.Oracle(P)
    o_same.combine_same(buf)
    o_mix = oracle.Oracle(P)
    ids = np.array([1, 0], np.int32)
    cs = np.array([1.0, 0.0, 1.0, 0.0])
    # event 0 -> partner 1 forms (a, b); event 1 -> partner 0 forms (b, a):
    # keep only event 0's call
    o_mix.combine_mixed(buf, 0, buf, 1, ids[:1], cs[:2])
    return o_same.arrays(), o_mix.arrays()


def _bin(kt, io, is_, il):
    q = P["qnpts"]
    return ((kt * q + io) * q + is_) * q + il


def _qidx(v):
    return int((v - D["base"]) / D["dq"])


def _kt_index(k):
    return int((k - P["KT_min"]) / D["dKT"])


def _designed():
    """(pair, coordinate, value) of every designed pair."""
    out = []
    vals = E.edge_values(P)
    pairs = E.edge_pairs_3d(P)
    k = 0
    for v in vals:
        if E._exact_split(v) is not None:
            out.append((pairs[k], "out", v))
            k += 1
        out.append((pairs[k], "side", v))
        out.append((pairs[k + 1], "long", v))
        k += 2
    for a, b in pairs[k:]:
        out.append(((a, b), "kt", a[0]))
    assert len(out) == len(pairs)
    return out


@pytest.fixture(scope="module", autouse=True)
def _build(oracle_mod):
    return oracle_mod


def _expect(a, b, coord, v, mixed):
    """(binned?, (kt, io, is, il)) from the reference's comparisons."""
    kx = 0.5 * (a[0] + b[0])
    k2 = kx * kx
    if not (k2 >= P["KT_min"] * P["KT_min"] and k2 <= P["KT_max"] * P["KT_max"]):
        return False, None
    kt = _kt_index(math.sqrt(k2))
    if coord == "kt":
        return True, (kt, _qidx(0.0), _qidx(0.0), _qidx(0.02))
    keep = v >= D["lo"] and (v < D["hi"] if mixed else v <= D["hi"])
    if not keep or _qidx(v) >= P["qnpts"]:
        return False, None
    # the other two components of each design
    if coord == "out":
        return True, (kt, _qidx(v), _qidx(0.0), _qidx(0.02))
    if coord == "side":
        return True, (kt, _qidx(0.0), _qidx(v), _qidx(0.026))
    return True, (kt, _qidx(0.0), _qidx(0.0), _qidx(v))


@pytest.mark.parametrize("k", range(len(_designed())))
def test_designed_pair_outcome(k):
    (a, b), coord, v = _designed()[k]
    same, mixed = _run(a, b)
    for arr, is_mixed, key in ((same, False, "num_count"), (mixed, True, "den")):
        binned, where = _expect(a, b, coord, v, is_mixed)
        total = int(arr[key].sum())
        assert total == (1 if binned else 0), (coord, v, is_mixed, total)
        if binned:
            hit = int(np.flatnonzero(arr[key])[0])
            assert hit == _bin(*where), (coord, v, is_mixed, hit, where)


def test_edges_are_where_the_rules_split():
    """The designed values straddle every rule: same keeps hi, mixed drops
    it; lo is kept by both, one ulp below lo by neither; K_perp = KT_max is
    kept with index n_KT - 1; one ulp above it is dropped."""
    outcomes = {}
    for (a, b), coord, v in _designed():
        if coord != "side":
            continue
        same, mixed = _run(a, b)
        outcomes[v] = (int(same["num_count"].sum()), int(mixed["den"].sum()))
    assert outcomes[D["hi"]] == (1, 0)
    assert outcomes[D["lo"]] == (1, 1)
    assert outcomes[math.nextafter(D["lo"], -1.0)] == (0, 0)
    assert outcomes[math.nextafter(D["hi"], 1.0)] == (0, 0)
    assert outcomes[math.nextafter(D["hi"], -1.0)] == (1, 1)
    kt = {}
    for (a, b), coord, v in _designed():
        if coord == "kt":
            same, _ = _run(a, b)
            kt[v] = np.flatnonzero(same["num_count"])
    q3 = P["qnpts"] ** 3
    assert kt[P["KT_max"]].size == 1 and kt[P["KT_max"]][0] // q3 == P["n_KT"] - 1
    assert kt[math.nextafter(P["KT_max"], 1.0)].size == 0
    assert kt[P["KT_min"]].size == 1 and kt[P["KT_min"]][0] // q3 == 0
    assert kt[math.nextafter(P["KT_min"], 0.0)].size == 0
