"""GPU parity on pairs placed at the edges the fast kernel's pruning relies on.

The fast kernel drops work it can prove holds no binnable pair:
  - particles whose |p_T| lies outside [KT_min - D/2, KT_max + D/2]
    (D = sqrt(2) H, H the q window half-width);
  - same-event j's whose p_x differs from the tile's by more than D;
  - mixed partners whose |p_T| differs from the tile's by more than D;
  - pairs failing the conservative FP32/FP64 prefilter (K_T window, |q_T|^2 <= 2 H^2,
    the beta bound of |q_long| <= H).
Each event here holds pairs built from (K_T, q_out, q_side, q_long) at those edges:
K_T within 1e-9 of the K_T window's ends, |q_out| and |q_side| within 1e-9 of the
window's half-width, q_T along the x axis (|dp_x| = |q_T| = sqrt(2) H at the corner),
q_long at its edge.  The fast kernel (FP64 and precision form) must give the
oracle's counts bit for bit (also with slab binning), and sums within the usual bars:
err = |S - S_ref| / max(|S_ref|, sum|.|_bin) <= 1e-12 (FP64), 1e-6 (precision form).

Some q components here are exactly zero by construction, so a bin can hold only
pairs whose q_out (say) is rounding noise of order 1e-17 GeV, of either sign,
and sum|q_out|_bin is noise too.  The precision form computes q_out through
another (refined reciprocal square root) path than the reference, with an
absolute error of ~1e-16 |q_T|, so its bar here also admits an absolute error of
1e-14 GeV per pair in the bin (far below any q resolution); the FP64 kernel
uses the reference's expressions and keeps the pure relative bar.
"""
import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

MODES = {"fast_fp64": (_lib.HBT_PREC_FP64, 1e-12, 0.0, 0),
         "fast_fp32c": (_lib.HBT_PREC_FP32C, 1e-6, 1e-14, 0),
         "fast_fp64_slab": (_lib.HBT_PREC_FP64, 1e-12, 0.0, 1),
         "fast_fp32c_slab": (_lib.HBT_PREC_FP32C, 1e-6, 1e-14, 1)}
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))
M_PI = 0.13957


def edge_buffer(params, nev, seed):
    """Events of edge pairs: p1 = K + q/2, p2 = K - q/2 (transverse), q built from
    (q_out, q_side) along and across K; K_z = 0, q_z = q_long (LCMS = lab when
    the boost is off)."""
    rng = np.random.default_rng(seed)
    dq = (params["q_max"] - params["q_min"]) / (params["qnpts"] - 1)
    hi = params["q_max"] + dq / 2 - 1e-8
    a = hi * (1 - 1e-9)
    kts = (params["KT_min"] * (1 + 1e-9), params["KT_max"] * (1 - 1e-9),
           0.5 * (params["KT_min"] + params["KT_max"]))
    qts = [(s1 * a, s2 * a) for s1 in (1, -1) for s2 in (1, -1)] + \
          [(a, 0.0), (-a, 0.0), (0.0, a), (0.0, -a)]
    qls = (0.0, a, -a)
    cols = {k: [] for k in ("px", "py", "pz", "E", "x", "y", "z", "t")}
    off = [0]
    for _ in range(nev):
        for kt in kts:
            for qo, qs in qts:
                # K direction: q_T along +x, or a random direction
                for phi in (np.arctan2(-qs, qo) if (qo or qs) else 0.0, rng.uniform(0, 2 * np.pi)):
                    kx, ky = kt * np.cos(phi), kt * np.sin(phi)
                    ux, uy = np.cos(phi), np.sin(phi)
                    qx, qy = qo * ux - qs * uy, qo * uy + qs * ux
                    for ql in qls:
                        for sgn in (1, -1):
                            px, py, pz = kx + sgn * qx / 2, ky + sgn * qy / 2, sgn * ql / 2
                            cols["px"].append(px)
                            cols["py"].append(py)
                            cols["pz"].append(pz)
                            cols["E"].append(np.sqrt(M_PI**2 + px * px + py * py + pz * pz))
                            for c in ("x", "y", "z"):
                                cols[c].append(rng.normal(0.0, 3.0))
                            cols["t"].append(rng.uniform(0.0, 10.0))
        # shuffle inside the event so pair orientation and tile placement vary
        n0 = off[-1]
        n1 = len(cols["px"])
        perm = n0 + rng.permutation(n1 - n0)
        for c in cols:
            cols[c][n0:n1] = [cols[c][i] for i in perm]
        off.append(n1)
    arr = {c: np.asarray(v, dtype=np.float64) for c, v in cols.items()}
    return samples.Buffer(ev_off=np.asarray(off, np.int64), **arr)


def rel(a, b, floor, count, tol, abs_per_pair):
    d = np.maximum(np.maximum(np.abs(b), floor), (abs_per_pair / tol) * count)
    d = np.where(d > 0, d, 1.0)
    return float((np.abs(a - b) / d).max()) if a.size else 0.0


@pytest.mark.parametrize("boost", [0, 1])
@pytest.mark.parametrize("mode", list(MODES))
def test_edge_pairs_match_oracle(mode, boost):
    prec, tol, abs_per_pair, binning = MODES[mode]
    params = dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
                  HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
                  number_of_mixed_events=0, long_comoving_boost=boost)
    buf = edge_buffer(params, 4, 11 + boost)
    print(f"edges: {buf.nev} events, {buf.n} particles, boost {boost}, {mode}", flush=True)
    g = hbt.HBTCorrelation(dict(params, hbt_binning=binning), precision=prec,
                           kernel=_lib.HBT_KERNEL_FAST)
    g.calculate_groups([buf], hbt.Random(5))
    r = g.results()
    g.close()
    o = oracle.Oracle(params)
    o.calculate_buffer(buf, oracle.OracleRng(5))
    ref = o.arrays()
    if boost == 0:
        # every designed pair is binnable: (3 K_T x 8 q_T x 2 directions x 3 q_long) per event
        assert ref["num_count"].sum() >= 4 * 3 * 8 * 2 * 3
    assert np.array_equal(r.num_count, ref["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, ref["den"].astype(np.uint64))
    cnt = ref["num_count"].astype(np.float64)
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], ref[fl], cnt, tol, abs_per_pair) <= tol, key


EXACT_MODES = {"plain_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_PLAIN, 1e-12, 0.0, 0),
               "fast_fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 0.0, 0),
               "fast_fp32c": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 1e-14, 0),
               "fast_fp64_slab": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, 1e-12, 0.0, 1),
               "fast_fp32c_slab": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 1e-6, 1e-14, 1)}


@pytest.mark.parametrize("kind", ["3d", "kphi"])
@pytest.mark.parametrize("mode", list(EXACT_MODES))
def test_exact_edges_same_and_mixed(mode, kind):
    """Pairs placed exactly on lo+1e-8 and hi-1e-8 (and one ulp around them),
    on interior q bin edges, at K_perp = KT_min / KT_max, and (kphi) with the
    pair axis exactly on a K_phi bin edge or wrapping to 2 pi
    (tests/golden/edge_inputs.py).  Same-event pairs as the reference forms
    them; mixed pairs with the explicit rotation (cos, sin) = (1, 0), under
    which each designed pair is formed again and decided by the mixed
    comparisons (>= on hi).  Counts bit-exact with the oracle, sums within the
    mode's bar."""
    import edge_inputs as E
    prec, kern, tol, abs_per_pair, binning = EXACT_MODES[mode]
    params = dict(E.EDGE_PARAMS if kind == "3d" else E.KPHI_PARAMS, hbt_binning=binning)
    buf = E.build(kind)
    print(f"exact edges {kind}: {buf.n} particles, {mode}", flush=True)
    ids = np.array([1, 0], np.int32)
    cs = np.array([1.0, 0.0, 1.0, 0.0])
    g = hbt.HBTCorrelation(params, precision=prec, kernel=kern)
    g.upload(buf)
    g.run_same()
    g.run_mixed(1, ids, cs)
    r = g.results()
    deferred = g.deferred_pairs()
    g.close()
    o = oracle.Oracle(params)
    if kind == "kphi":
        psi = oracle.psi_ref(buf, 2)
        assert psi == 0.0  # mirror images cancel sin(2 phi) exactly
        o.set_psi_ref(psi)
    o.combine_same(buf)
    o.run_mixed_plan(buf, 1, ids, cs)
    ref = o.arrays()
    assert ref["num_count"].sum() > 0 and ref["den"].sum() > 0
    assert np.array_equal(r.num_count, ref["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, ref["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_kt, ref["cnt_num_kt"])
    assert np.array_equal(r.cnt_den_kt, ref["cnt_den_kt"])
    if kind == "kphi":
        assert np.array_equal(r.cnt_num_ktphi, ref["cnt_num_ktphi"])
        assert np.array_equal(r.cnt_den_ktphi, ref["cnt_den_ktphi"])
        # pair axes on bin edges are decided on the host
        assert deferred > 0
    cnt = ref["num_count"].astype(np.float64)
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], ref[fl], cnt, tol, abs_per_pair) <= tol, key
