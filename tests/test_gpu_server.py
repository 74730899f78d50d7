"""GPU tests of the request server (hbt_server_* in include/hbt_hip.h).

A request is one whole analysis: its events grouped into buffers of
events_per_group consecutive events, each buffer processed as
calculate_HBT_correlation_function does, plans drawn from Random(seed)
buffer after buffer.  Bars (written here):
  - against the CPU oracle on the same request: counts bit-exact, FP64 sums
    within 1e-12 under err = |S - S_ref| / max(|S_ref|, sum|.|_bin);
  - the same request run alone and inside a mix of other requests that
    arrive while it runs: counts bit-identical, sums within 1e-12 of each
    other (only the order of the FP64 additions may differ), in both
    precisions;
  - admission by HBM footprint: with an arena that holds about two requests
    every request still completes, resident bytes never exceed the arena,
    and a request larger than the arena is refused at submit.
"""
import time

import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=21, q_min=-0.1, q_max=0.1, n_KT=5, KT_min=0.2, KT_max=0.6,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=0, long_comoving_boost=1)
SUMS = (("q_out_sum", "abs_q_out"), ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long"),
        ("num", "abs_num"))
COUNTS = ("num_count", "den", "cnt_num_kt", "cnt_den_kt")


def request(seed, nev, mult, sigma=4.0):
    return samples.generate(seed, nev, mult, sigma=sigma)


def oracle_request(buf, epg, seed, params=PARAMS):
    o = oracle.Oracle(params)
    rng = oracle.OracleRng(seed)
    for g in range(0, buf.nev, epg):
        o.calculate_buffer(buf.events(g, min(epg, buf.nev - g)), rng)
    return o.arrays()


def rel(a, b, floor):
    d = np.maximum(np.abs(b), floor)
    d = np.where(d > 0, d, 1.0)
    return float((np.abs(a - b) / d).max()) if a.size else 0.0


def same_result(a, b, tol):
    """Counts identical; sums within tol of each other per bin, relative to
    max(|S|, pair count of the bin) (every term is at most 1 in size)."""
    for k in COUNTS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    floor = b.num_count.astype(np.float64)
    for k, _ in SUMS:
        assert rel(getattr(a, k), getattr(b, k), floor) <= tol, k


def test_server_matches_oracle():
    buf = request(11, 30, 120.0)
    print("server vs oracle: 30 events Poisson(120), groups of 10, seed 5", flush=True)
    srv = hbt.HBTServer(PARAMS, precision=_lib.HBT_PREC_FP64, hbm_bytes=1 << 30)
    rid = srv.submit(buf, 10, 5)
    assert srv.wait(rid, 60.0)
    r = srv.result(rid)
    info = srv.info(rid)
    srv.release(rid)
    srv.close()
    ref = oracle_request(buf, 10, 5)
    assert info["state"] == "done"
    assert info["enum_same"] == int(ref["same_st"][0])
    assert info["enum_mixed"] == int(ref["mixed_st"][0])
    assert np.array_equal(r.num_count, ref["num_count"].astype(np.uint64))
    assert np.array_equal(r.den, ref["den"].astype(np.uint64))
    assert np.array_equal(r.cnt_num_kt, ref["cnt_num_kt"].astype(np.uint64))
    assert np.array_equal(r.cnt_den_kt, ref["cnt_den_kt"].astype(np.uint64))
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], ref[fl]) <= 1e-12, key


@pytest.mark.parametrize("prec", [_lib.HBT_PREC_FP64, _lib.HBT_PREC_FP32C])
def test_alone_equals_in_mix(prec):
    probe = request(21, 40, 200.0, sigma=3.0)
    others = [request(100 + k, n, m, sigma=s) for k, (n, m, s) in
              enumerate([(300, 2500.0, 5.0), (20, 80.0, 2.0), (60, 400.0, 6.5), (10, 50.0, 4.0),
                         (80, 250.0, 3.5)])]
    print(f"alone vs mix, precision {prec}", flush=True)
    # alone
    srv = hbt.HBTServer(PARAMS, precision=prec, hbm_bytes=2 << 30, step_pairs=2e8)
    rid = srv.submit(probe, 10, 7)
    assert srv.wait(rid, 60.0)
    alone = srv.result(rid)
    srv.release(rid)
    # in a mix: the big request first, the probe while it runs, the rest after
    ids = [srv.submit(others[0], 10, 1000)]
    time.sleep(0.002)
    pid = srv.submit(probe, 10, 7)
    ids += [srv.submit(b, 10, 1001 + k) for k, b in enumerate(others[1:])]
    for i in ids + [pid]:
        assert srv.wait(i, 120.0)
    mixed = srv.result(pid)
    st = srv.stats()
    infos = [srv.info(i) for i in ids + [pid]]
    for i in ids + [pid]:
        srv.release(i)
    srv.close()
    same_result(mixed, alone, 1e-12)
    # the probe joined the launches of the large request (2e10 pairs) that
    # was running when it arrived, and finished first
    big = infos[0]
    p = infos[-1]
    assert p["t_active"] < big["t_done"]
    assert p["t_done"] < big["t_done"]
    assert st["completed"] == 1 + len(ids) + 1
    assert st["resident_peak"] >= 2


def test_footprint_admission():
    bufs = [request(300 + k, 30, 400.0) for k in range(5)]
    srv0 = hbt.HBTServer(PARAMS, precision=_lib.HBT_PREC_FP64, hbm_bytes=1 << 30)
    rid = srv0.submit(bufs[0], 10, 3)
    foot = srv0.info(rid)["footprint"]
    assert srv0.wait(rid, 60.0)
    alone = srv0.result(rid)
    srv0.release(rid)
    srv0.close()
    arena = int(2.5 * foot)
    print(f"footprint {foot} B, arena {arena} B", flush=True)
    srv = hbt.HBTServer(PARAMS, precision=_lib.HBT_PREC_FP64, hbm_bytes=arena, step_pairs=1e8)
    ids = [srv.submit(b, 10, 3 + k) for k, b in enumerate(bufs)]
    for i in ids:
        assert srv.wait(i, 120.0)
    first = srv.result(ids[0])
    st = srv.stats()
    infos = [srv.info(i) for i in ids]
    for i in ids:
        srv.release(i)
    assert st["arena_peak"] <= arena
    assert st["resident_peak"] <= 2
    assert all(i["state"] == "done" for i in infos)
    # a later request was admitted only after an earlier one left
    assert infos[2]["t_admit"] >= min(infos[0]["t_done"], infos[1]["t_done"])
    same_result(first, alone, 1e-12)
    with pytest.raises(hbt.HbtError):
        srv.submit(request(999, 200, 2000.0), 10, 1)  # needs more than the arena
    srv.close()


@pytest.mark.parametrize("prec", [_lib.HBT_PREC_FP64, _lib.HBT_PREC_FP32C])
def test_request_in_parts_equals_oracle_and_whole(prec):
    """part_particles small enough that every group of the request is a part
    of its own: six parts prepared, uploaded and launched one after the
    other, all adding to one histogram block.  Against the oracle (counts
    bit-exact, FP64 sums within 1e-12, the precision form's within 1e-6) and
    against the same request in one part (counts identical, sums within
    1e-12 of each other), also with another request in parts running
    beside it."""
    buf = request(41, 60, 300.0, sigma=3.5)
    other = request(42, 40, 500.0, sigma=5.0)
    print(f"request in parts, precision {prec}", flush=True)
    whole_srv = hbt.HBTServer(PARAMS, precision=prec, hbm_bytes=1 << 30)
    rid = whole_srv.submit(buf, 10, 9)
    assert whole_srv.wait(rid, 60.0)
    whole = whole_srv.result(rid)
    whole_srv.release(rid)
    whole_srv.close()
    srv = hbt.HBTServer(PARAMS, precision=prec, hbm_bytes=1 << 30, step_pairs=2e7,
                        part_particles=2000)
    oid = srv.submit(other, 10, 10)
    rid = srv.submit(buf, 10, 9)
    assert srv.wait(rid, 60.0) and srv.wait(oid, 60.0)
    parts = srv.result(rid)
    info = srv.info(rid)
    st = srv.stats()
    srv.release(rid)
    srv.release(oid)
    srv.close()
    ref = oracle_request(buf, 10, 9)
    assert info["state"] == "done"
    assert info["enum_same"] == int(ref["same_st"][0])
    assert info["enum_mixed"] == int(ref["mixed_st"][0])
    assert st["resident_peak"] >= 2
    for k in COUNTS:
        assert np.array_equal(getattr(parts, k), ref[k].astype(np.uint64)), k
    tol = 1e-12 if prec == _lib.HBT_PREC_FP64 else 1e-6
    for key, fl in SUMS:
        assert rel(getattr(parts, key), ref[key], ref[fl]) <= tol, key
    same_result(parts, whole, 1e-12)


def test_kphi_request_in_parts_decides_edge_pairs_on_the_host():
    """16 K_phi bins with a 1e-3 rad guard band: many pairs of every part go
    to the request's shared deferral buffer and are decided on the host once
    the last part has run.  Counts bit-exact against the oracle."""
    params = dict(PARAMS, azimuthal_flag=1, n_Kphi=16, hbt_kphi_guard_prad=10 ** 9)
    buf = request(43, 40, 250.0, sigma=4.0)
    print("K_phi request in parts, guard 1e-3 rad", flush=True)
    srv = hbt.HBTServer(params, precision=_lib.HBT_PREC_FP64, hbm_bytes=1 << 30,
                        part_particles=2000)
    rid = srv.submit(buf, 10, 12)
    assert srv.wait(rid, 60.0)
    r = srv.result(rid)
    st = srv.stats()
    srv.release(rid)
    srv.close()
    ref = oracle_request(buf, 10, 12, params)
    assert st["deferred"] > 0
    for k in ("num_count", "den"):
        assert np.array_equal(getattr(r, k), ref[k].astype(np.uint64)), k
    for key, fl in SUMS:
        assert rel(getattr(r, key), ref[key], ref[fl]) <= 1e-12, key
