"""HBT pair-correlation throughput on one MI355X.

Workload (the precision form's single large workload): 2000 pi+ events,
multiplicity Poisson(600), source model of hadronic_afterburner_toolkit_amd.
samples, grouped by 50 consecutive events (oversample 50: 40 groups of
~30k particles, ~4.5e8 same-event pairs each), 10 rotated mixed partners
per event, q grid 41^3 on [-0.15, 0.15] GeV, K_T edges 0.2..0.7 GeV (step
0.1; the [0.5,0.7] bin of the quoted edges is the sum of the last two),
|y| < 1, pair cap off.

One step = one pass of the hot path over the whole batch: every same-event
pair of the 40 groups and every mixed pair of the 2000 events, one
persistent launch replayed from a hipGraph, inputs resident in HBM.

Printed: one JSON line; ``value`` is the precision form's pairs/s (same +
mixed enumerated pairs), with the FP64 kernels, the per-bin error of the
precision form against the FP64 oracle, the roofline of the dominant kernel
and the CPU oracle timed on a bounded sample beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
FP64_PEAK_TFLOPS = 78.6   # MI355X published FP64 vector peak

PARAMS = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              number_of_mixed_events=10, long_comoving_boost=1)


BENCH_BINS = 41 ** 3 * 5  # histogram bins of the headline workload (q 41^3, 5 K_T bins)


def make_workload(nev, mult, per_group, seed):
    from hadronic_afterburner_toolkit_amd import samples
    full = samples.generate(seed, nev, mult)
    return [full.events(g * per_group, per_group) for g in range(nev // per_group)]


def draw_plans(groups, seed, M):
    from hadronic_afterburner_toolkit_amd import hbt
    rng = hbt.Random(seed)
    Mg, ids, cs = [], [], []
    for b in groups:
        m, i, c = rng.draw_plan(b.nev, b.nev, M)
        Mg.append(m)
        ids.append(i)
        cs.append(c)
    return np.array(Mg, np.int32), np.concatenate(ids), np.concatenate(cs)


def flops(stages_same, stages_mixed):
    """Algorithmic FLOPs of the cascade per launch (+ - * / sqrt count 1 each,
    attributed to the first stage that needs them; cos not counted):
    same  = 7 N_enum + 12 N_KT + 5 N_out + 17 N_side + 17 N_bin
    mixed = 7 N_enum + 12 N_KT + 5 N_out + 17 N_side +  1 N_bin"""
    s, m = [float(v) for v in stages_same], [float(v) for v in stages_mixed]
    return (7 * s[0] + 12 * s[1] + 5 * s[2] + 17 * s[3] + 17 * s[4]
            + 7 * m[0] + 12 * m[1] + 5 * m[2] + 17 * m[3] + 1 * m[4])


def run_mode(groups, plan, prec, kern, steps, warmup, autotune, log, tuning=None, binning=0):
    from hadronic_afterburner_toolkit_amd import hbt
    g = hbt.HBTCorrelation(dict(PARAMS, hbt_binning=binning), precision=prec, kernel=kern)
    t0 = time.perf_counter()
    g.upload_groups(groups)
    g.prepare_batch(*plan)
    upload_s = time.perf_counter() - t0
    # again on the same context, from packed columns: the steady state of a
    # caller that uploads buffer after buffer (device and pinned buffers kept)
    packed = hbt.pack_groups(groups)
    t0 = time.perf_counter()
    g.upload_packed(packed)
    upload_warm_s = time.perf_counter() - t0
    g.prepare_batch(*plan)
    prepare_warm_s = time.perf_counter() - t0 - upload_warm_s
    if tuning:
        g.set_tuning(*tuning)
    tune = g.autotune() if autotune else g.tuning()
    es, em, nitems = g.batch_info()
    graph = g.batch_graph()
    for _ in range(warmup):
        g.launch_batch()
    g.sync()
    g.reset()
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        g.launch_batch()
        dev_ms.append(g.sync())
    wall = time.perf_counter() - t0
    slab = g.slab_info()
    log(f"  prec={prec} kernel={kern} binning={binning} steps={steps} wall={wall:.3f}s "
        f"dev_ms/step={np.mean(dev_ms):.2f} tune={tune} slab={slab}")
    return g, dict(wall=wall, dev_ms=dev_ms, enum_same=es, enum_mixed=em, nitems=nitems, slab=slab,
                   tune=tune, graph=graph, upload_s=upload_s, upload_warm_s=upload_warm_s,
                   prepare_warm_s=prepare_warm_s)


def serve(args, log):
    """Serving form: a Poisson stream of independent analyses (rate 20/s),
    N events log-uniform on [50, 5000], multiplicity Poisson(mu) with mu
    uniform on [100, 1500], source radius sigma uniform on [2, 7] fm, groups
    of 10 events, 10 mixed rotations per event, same q/K_T grid.  Requests
    are generated before the clock starts, submitted at their arrival times
    by a pool of submitter threads, and served by the request server.
    Latency = arrival -> result in host memory."""
    import concurrent.futures as cf
    from hadronic_afterburner_toolkit_amd import _lib, hbt, samples
    rng = np.random.default_rng(args.seed + 7)
    n = args.serve_requests
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
    nevs = np.exp(rng.uniform(np.log(50), np.log(5000), n)).astype(int)
    mus = rng.uniform(100, 1500, n)
    sigmas = rng.uniform(2.0, 7.0, n)
    log(f"serving: generating {n} requests")
    reqs = [samples.generate(args.seed + 1000 + k, int(nevs[k]), float(mus[k]),
                             sigma=float(sigmas[k])) for k in range(n)]
    srv = hbt.HBTServer(PARAMS, precision=_lib.HBT_PREC_FP32C, hbm_bytes=64 << 30,
                        max_resident=32, step_pairs=args.serve_step_pairs)
    # warm-up request (first launch, code load), not timed
    w = srv.submit(reqs[0].events(0, 20), 10, 1)
    srv.wait(w, 60.0)
    srv.release(w)
    picked = [None] * n
    infos = [None] * n
    t0 = time.perf_counter()

    def client(k):
        """One client: submit at arrival, wait, take the result, release."""
        picked[k] = time.perf_counter()
        rid = srv.submit(reqs[k], 10, int(args.seed + k))
        srv.wait(rid, 600.0)
        infos[k] = srv.info(rid)
        srv.result(rid)
        srv.release(rid)

    with cf.ThreadPoolExecutor(max_workers=64) as pool:
        futs = []
        for k in range(n):
            dt = t0 + arrivals[k] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            futs.append(pool.submit(client, k))
        for f in futs:
            f.result()
    pairs = 0
    t_end = 0.0
    lat = [None] * n
    stages = {"front": [], "host_prep": [], "admission_wait": [], "upload": [],
              "to_first_launch": [], "launches": [], "readback": []}
    for k in range(n):
        info = infos[k]
        assert info["state"] == "done", info
        for name, a, b in (("host_prep", "t_submit", "t_queued"),
                           ("admission_wait", "t_queued", "t_admit"),
                           ("upload", "t_admit", "t_active"),
                           ("to_first_launch", "t_active", "t_first_launch"),
                           ("launches", "t_first_launch", "t_last_launch"),
                           ("readback", "t_last_launch", "t_done")):
            stages[name].append(1e3 * (info[b] - info[a]))
        front = picked[k] - (t0 + arrivals[k])  # arrival -> submit call
        stages["front"].append(1e3 * front)
        lat[k] = front + (info["t_done"] - info["t_submit"])
        pairs += info["enum_same"] + info["enum_mixed"]
        t_end = max(t_end, arrivals[k] + lat[k])
    st = srv.stats()
    srv.close()
    lat_ms = np.array(lat) * 1e3
    out = {"requests": n, "rate_hz": args.serve_rate, "p50_ms": float(np.percentile(lat_ms, 50)),
           "p99_ms": float(np.percentile(lat_ms, 99)), "mean_ms": float(lat_ms.mean()),
           "max_ms": float(lat_ms.max()), "pairs": float(pairs), "makespan_s": float(t_end),
           "pairs_per_s": float(pairs / t_end), "busy_frac": float(st["busy_s"] / t_end),
           "launches": int(st["launches"]), "resident_peak": int(st["resident_peak"]),
           "arena_peak_bytes": int(st["arena_peak"]), "precision": "fp32c",
           "step_pairs": args.serve_step_pairs,
           "stages_ms_p50_p99": {k: [float(np.percentile(v, 50)), float(np.percentile(v, 99))]
                                 for k, v in stages.items()}}
    log(f"  serving: p50 {out['p50_ms']:.1f} ms p99 {out['p99_ms']:.1f} ms "
        f"{out['pairs_per_s']:.3e} pairs/s busy {out['busy_frac']:.2f}")
    return out


PI_K = ((211, 0.13957, 0.25, 6.0), (321, 0.49368, 0.30, 1.0))  # monval, mass, T, weight


def batch_line(groups, params, seed, prec, steps, tuning, autotune, log, label):
    """One analysis in the batched form: groups resident, plans from one
    Random(seed) stream, one persistent launch per step (hipGraph replay)."""
    from hadronic_afterburner_toolkit_amd import _lib, hbt
    rng = hbt.Random(seed)
    Mg, ids, cs = [], [], []
    for b in groups:
        m, i, c = rng.draw_plan(b.nev, b.nev, int(params.get("number_of_mixed_events", 0)))
        Mg.append(m)
        ids.append(i)
        cs.append(c)
    g = hbt.HBTCorrelation(params, precision=prec, kernel=_lib.HBT_KERNEL_FAST)
    g.upload_groups(groups)
    g.prepare_batch(np.array(Mg, np.int32), np.concatenate(ids), np.concatenate(cs))
    if tuning:
        g.set_tuning(*tuning)
    tune = g.autotune() if autotune else g.tuning()
    es, em, _ = g.batch_info()
    g.launch_batch()
    g.sync()
    ms = []
    for _ in range(steps):
        g.launch_batch()
        ms.append(g.sync())
    res = g.results()
    # binned pairs per launch (= atomic accesses), from the kernel's counters
    sc, mc = g.stage_counts()
    launches = max(1, int(sc[0]) // es if es else int(mc[0]) // max(1, em))
    slab = g.slab_info()  # slab binning or device-scope atomics (the autotuner's pick)
    g.close()
    t = float(np.mean(ms)) * 1e-3
    log(f"  {label}: {es + em:.3e} pairs in {t * 1e3:.1f} ms -> {(es + em) / t:.3e} pairs/s")
    return dict(pairs=es + em, same_pairs=es, mixed_pairs=em, ms=t * 1e3,
                pairs_per_s=(es + em) / t, tuning=tune, slab=slab,
                binned_same=int(sc[4]) // launches, binned_mixed=int(mc[4]) // launches), res


def add_atomic_floor(line, rates, nbins=None):
    """The launch's atomic floor: binned pairs over the measured rates, those
    of the micro-benchmark's histogram size nearest to nbins (in ratio) when
    it measured several.  It is the floor of binning every pair with
    device-scope atomics; a launch with slab binning (line["slab"]["on"])
    adds only its mixed pairs that way and can run below it."""
    if not rates:
        return
    same, mixed, at = rates["same"], rates["mixed"], BENCH_BINS
    if nbins and rates.get("by_bins"):
        at = min(rates["by_bins"], key=lambda n: abs(np.log(n / nbins)))
        if abs(np.log(at / nbins)) < abs(np.log(BENCH_BINS / nbins)):
            same, mixed = rates["by_bins"][at]
        else:
            at = BENCH_BINS
    line["atomic_floor_ms"] = 1e3 * (line["binned_same"] / same + line["binned_mixed"] / mixed)
    line["atomic_floor_bins"] = at


def batched_uneven(args, log):
    """Batched uneven form: 512 events, multiplicity log-uniform on [20, 4000],
    pi+ and K+ (6:1, K+ with T = 0.30 GeV) analysed as separate identical-
    particle histograms, groups of 20 events, 16 K_phi bins about the event
    plane (folded to 8 bins of Phi_K mod pi afterwards), K_T edges 0.2..0.7
    (step 0.1), the reference's partner rule (11 per event)."""
    from hadronic_afterburner_toolkit_amd import _lib, samples
    full = samples.generate(args.seed + 20, 512, samples.log_uniform_mult(20, 4000), species=PI_K)
    # binning 2: the autotuner times slab binning against atomics
    params = dict(PARAMS, azimuthal_flag=1, n_Kphi=16, number_of_mixed_events=0, hbt_binning=2)
    out = {"events": 512, "per_group": 20}
    tot_pairs = tot_t = 0.0
    for monval, name in ((211, "pi+"), (321, "K+")):
        buf = full.select(monval)
        groups = [buf.events(g, min(20, buf.nev - g)) for g in range(0, buf.nev, 20)]
        line, _ = batch_line(groups, params, args.seed + monval, _lib.HBT_PREC_FP32C,
                             max(1, args.steps), None, not args.no_autotune, log,
                             f"batched uneven {name}")
        out[name] = line
        tot_pairs += line["pairs"]
        tot_t += line["ms"] * 1e-3
    out["pairs_per_s"] = tot_pairs / tot_t
    return out


def kphi_binning(args, log, groups):
    """The headline workload analysed in 16 K_phi bins as well (6 x 16 x 41
    = 3936 slabs), the precision form binned with device-scope atomics and
    with slab binning, at one fixed launch shape (R 1, jchunk 256, kchunk 4,
    4 blocks per CU) so that the two differ only in the binning."""
    from hadronic_afterburner_toolkit_amd import _lib
    out = {}
    for b, name in ((0, "atomics"), (1, "slab")):
        params = dict(PARAMS, azimuthal_flag=1, n_Kphi=16, hbt_binning=b)
        out[name], _ = batch_line(groups, params, args.seed + 3, _lib.HBT_PREC_FP32C,
                                  max(1, args.steps), (1, 256, 4, 4), False, log,
                                  f"K_phi headline, {name}")
    out["slab_speedup"] = out["atomics"]["ms"] / out["slab"]["ms"]
    return out


def large_histogram(args, log, tuning):
    """Large-histogram form: 1000 pi+ events Poisson(4000), groups of 100,
    20 rotations per event, q 61^3 on [-0.2, 0.2] GeV, 8 K_T bins on
    [0.15, 1.0] GeV, 16 K_phi bins: 3.2e7 bins (2.1 GB of histogram).
    One launch takes seconds, so instead of the autotuner's sweep the
    headline's launch shape runs at R = 1 and R = 2 and the faster counts."""
    from hadronic_afterburner_toolkit_amd import _lib, samples
    full = samples.generate(args.seed + 40, 1000, 4000.0)
    params = dict(PARAMS, qnpts=61, q_min=-0.2, q_max=0.2, n_KT=9, KT_min=0.15, KT_max=1.0,
                  azimuthal_flag=1, n_Kphi=16, number_of_mixed_events=20)
    groups = [full.events(g, 100) for g in range(0, 1000, 100)]
    best = None
    for R in (1, 2):
        line, _ = batch_line(groups, params, args.seed + 41, _lib.HBT_PREC_FP32C, 1,
                             (R,) + tuple(tuning[1:]), False, log, f"large histogram R={R}")
        if best is None or line["ms"] < best["ms"]:
            best = line
    best.update(events=1000, per_group=100, bins=61 ** 3 * 9 * 16)
    return best


BF_PARAMS = dict(particle_alpha=211, particle_beta=211, BpT_min=0.2, BpT_max=3.0,
                 Brap_max=2.0, Bnpts=21, rap_type=1)  # parameters.dat:122-128 with pi+
BF_SPECIES = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0),
              (321, 0.49368, 0.30, 1.0), (-321, 0.49368, 0.30, 1.0))
BF_FAST_VALU_PER_PAIR = 15.25  # vector instructions per pair, k_bf fast form (gfx950 assembly: 122 per 8 rows)
MI355X_CUS, MI355X_GHZ = 256, 2.4


def balance_function(args, log):
    """Balance-function pair histograms (BalanceFunction.cpp:120-197): 2000
    events, Poisson(900) particles of pi+-, K+- (3:3:1:1), alpha = beta = 211
    with the reference's parameters.dat cuts (pT in [0.2, 3] GeV, 21 Delta y
    bins up to 2, rapidity), one buffer; the fast and the exact forms, and
    the oracle on the first 40 events (one thread)."""
    from hadronic_afterburner_toolkit_amd import _lib, hbt, samples
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    buf = samples.generate(args.seed + 70, 2000, 900.0, species=BF_SPECIES)
    out = {"workload": "2000 events Poisson(900) pi+-,K+- 3:3:1:1, alpha=beta=211, "
                       "Bnpts 21, Brap_max 2, pT 0.2-3 GeV, rapidity",
           "events": buf.nev, "particles": int(buf.n)}
    hists = {}
    for name, kern in (("fast", _lib.HBT_BF_KERNEL_FAST), ("exact", _lib.HBT_BF_KERNEL_EXACT)):
        bf = hbt.BalanceFunction(BF_PARAMS, kernel=kern)
        runs = []
        for rep in range(4):
            bf.reset()
            t0 = time.perf_counter()
            bf.calculate_balance_function(buf, hbt.Random(args.seed + 71))
            wall = time.perf_counter() - t0
            runs.append((bf.info(), wall))
        hists[name] = bf.results()["hist"]
        bf.close()
        runs = runs[1:]  # the first call allocates
        info = runs[0][0]
        kms = float(np.median([r[0]["kernel_ms"] for r in runs]))
        line = {"kernel_ms": kms, "prep_ms": float(np.median([r[0]["prep_ms"] for r in runs])),
                "host_ms": float(np.median([r[0]["host_ms"] for r in runs])),
                "call_ms": 1e3 * float(np.median([r[1] for r in runs])),
                "pairs": int(info["pairs"]), "pairs_per_s": info["pairs"] / (kms * 1e-3),
                "fallback_fraction": info["fallback"] / max(1, info["pairs"]),
                "launches": int(info["launches"])}
        line["pairs_per_s_call"] = info["pairs"] / (line["call_ms"] * 1e-3)
        out[name] = line
        log(f"  balance function {name}: {kms:.3f} ms, {line['pairs_per_s']:.3e} pairs/s "
            f"(call {line['call_ms']:.1f} ms)")
    out["identical_fast_vs_exact"] = bool(np.array_equal(hists["fast"], hists["exact"]))
    ceil = MI355X_CUS * MI355X_GHZ * 1e9 * 64 / BF_FAST_VALU_PER_PAIR
    out["valu_roofline"] = {"bound": "valu", "ops_per_pair": BF_FAST_VALU_PER_PAIR,
                            "ceiling_pairs_per_s": ceil,
                            "frac": out["fast"]["pairs_per_s"] / ceil, "kernel": "k_bf fast"}
    if not args.no_cpu:
        oracle.build()
        sub = buf.events(0, 40)
        o = oracle.BfOracle(BF_PARAMS)
        t0 = time.perf_counter()
        o.calculate_buffer(sub, oracle.OracleRng(args.seed + 71))
        dt = time.perf_counter() - t0
        bf = hbt.BalanceFunction(BF_PARAMS)
        bf.calculate_balance_function(sub, hbt.Random(args.seed + 71))
        pairs = bf.info()["pairs"]
        same = bool(np.array_equal(bf.results()["hist"], o.arrays()["hist"]))
        bf.close()
        out["cpu_baseline"] = {"value": pairs / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"first 40 events: {pairs:.3e} pairs in {dt:.2f}s, "
                                         "oracle/hbt_oracle.c -O2 (lists + pair loops)",
                               "gpu_equals_oracle": same}
        out["speedup_kernel_vs_cpu_1core"] = out["fast"]["pairs_per_s"] / (pairs / dt)
        log(f"  balance function cpu oracle: {pairs / dt:.3e} pairs/s")
    return out


def dropin_line(groups, seed, log):
    """The drop-in path as Analysis::HBTAnalysis drives it
    (src/Analysis.cpp:821-833): buffer after buffer,
    calculate_HBT_correlation_function (upload, psi_ref, same-event pairs,
    mixing plan from the shared random stream, mixed pairs), then the
    histograms read back once.  The 40 bench groups as 40 buffers, precision
    form; cold = a fresh context (allocations included), warm = the same 40
    buffers again on that context."""
    from hadronic_afterburner_toolkit_amd import _lib, hbt
    out = {"buffers": len(groups)}
    g = hbt.HBTCorrelation(PARAMS, precision=_lib.HBT_PREC_FP32C, kernel=_lib.HBT_KERNEL_FAST)
    for label in ("cold", "warm"):
        rng = hbt.Random(seed)
        s0, m0 = g.stage_counts()
        t0 = time.perf_counter()
        for b in groups:
            g.calculate_HBT_correlation_function(b, rng)
        g.results()
        dt = time.perf_counter() - t0
        s1, m1 = g.stage_counts()
        pairs = float(int(s1[0]) - int(s0[0]) + int(m1[0]) - int(m0[0]))
        out[label] = {"s": dt, "ms_per_buffer": 1e3 * dt / len(groups), "pairs": pairs,
                      "pairs_per_s": pairs / dt}
        log(f"  drop-in {label}: {dt:.3f} s for {len(groups)} buffers "
            f"({1e3 * dt / len(groups):.2f} ms per buffer)")
    g.close()
    return out


def cpu_baseline(group, plan_g, log):
    """The CPU oracle on group 0 (its same-event pairs and its events' mixed
    pairs with the same plan), one thread."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    oracle.build()
    M, ids, cs = plan_g
    o = oracle.Oracle(PARAMS)
    t0 = time.perf_counter()
    o.combine_same(group)
    o.run_mixed_plan(group, M, ids, cs)
    dt = time.perf_counter() - t0
    a = o.arrays()
    pairs = a["same_st"][0] + a["mixed_st"][0]
    log(f"  cpu oracle: {pairs:.3e} pairs in {dt:.2f}s")
    return o, a, pairs / dt, dt, pairs


def cpu_baseline_multicore(groups, plan, log):
    """The CPU oracle on several groups at once, one thread per group (groups
    are independent once the plans are drawn; the oracle's C calls release
    the GIL): threads = min(16, cores, groups), the box's CPU share."""
    import threading
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    nthreads = max(1, min(16, os.cpu_count() or 1, len(groups)))
    Mg, ids, cs = plan
    offs = np.concatenate([[0], np.cumsum([int(Mg[g]) * groups[g].nev for g in range(len(groups))])])
    pairs = [0.0] * nthreads
    arrays = [None] * nthreads

    def work(g):
        M = int(Mg[g])
        o = oracle.Oracle(PARAMS)
        o.combine_same(groups[g])
        o.run_mixed_plan(groups[g], M, ids[offs[g]:offs[g + 1]], cs[2 * offs[g]:2 * offs[g + 1]])
        a = o.arrays()
        pairs[g] = float(a["same_st"][0] + a["mixed_st"][0])
        arrays[g] = a

    th = [threading.Thread(target=work, args=(g,)) for g in range(nthreads)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    log(f"  cpu oracle, {nthreads} threads: {sum(pairs):.3e} pairs in {dt:.2f}s")
    # the groups' histograms add up (counts, sums and the per-bin floors)
    total = {k: np.sum([np.asarray(a[k]) for a in arrays], axis=0) for k in arrays[0]}
    return {"value": sum(pairs) / dt, "unit": "pairs/s", "cores": nthreads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": model,
            "sample": f"groups 0..{nthreads - 1} of {len(groups)}, one thread each: "
                      f"{sum(pairs):.3e} same+mixed pairs in {dt:.1f}s, oracle/hbt_oracle.c -O2"}, total


def precision_vs_oracle(groups, plan, a):
    """The precision form on these groups as one batch against the FP64
    oracle's arrays a, per bin: max and RMS of |S - S_ref| / max(|S_ref|,
    sum|.|_bin) over the bins with pairs (num and the three q sums), and
    whether the counts are identical."""
    from hadronic_afterburner_toolkit_amd import _lib, hbt
    gs = hbt.HBTCorrelation(PARAMS, precision=_lib.HBT_PREC_FP32C, kernel=_lib.HBT_KERNEL_FAST)
    gs.upload_groups(groups)
    gs.prepare_batch(*plan)
    gs.launch_batch()
    gs.sync()
    rr = gs.results()
    gs.close()
    errs = []
    for key, fl_key in (("num", "abs_num"), ("q_out_sum", "abs_q_out"),
                        ("q_side_sum", "abs_q_side"), ("q_long_sum", "abs_q_long")):
        ref = a[key]
        den = np.maximum(np.abs(ref), a[fl_key])
        sel = den > 0
        errs.append(np.abs(getattr(rr, key) - ref)[sel] / den[sel])
    e = np.concatenate(errs)
    return {"max_rel_bin_err": float(e.max()), "rms_rel_bin_err": float(np.sqrt(np.mean(e * e))),
            "counts_identical": bool(np.array_equal(rr.num_count, a["num_count"].astype(np.uint64))
                                     and np.array_equal(rr.den, a["den"].astype(np.uint64))),
            "groups": len(groups),
            "metric": "|S-S_ref|/max(|S_ref|,sum|.|_bin), bins with pairs, num and q sums"}


def ingest(args, log):
    """Sample ingest (host only): a 200-event Poisson(600) sample written as
    mode-10 text (gzip and plain, %.7e like the reference's sample files) and
    as mode-9 binary records, read back with the engine's readers into SoA
    buffers, beside the oracle's restatement of the reference's mode-10
    reader (byte-at-a-time gz_readline, strtod per field)."""
    import tempfile
    from hadronic_afterburner_toolkit_amd import hbt, samples
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    buf = samples.generate(args.seed + 60, 200, 600.0)
    cols = ("mass", "t", "x", "y", "z", "E", "px", "py", "pz")
    vals = np.stack([getattr(buf, k) for k in cols], axis=1)
    out = {"particles": int(buf.n), "events": int(buf.nev)}
    with tempfile.TemporaryDirectory() as d:
        lines = []
        for e in range(buf.nev):
            a, b = int(buf.ev_off[e]), int(buf.ev_off[e + 1])
            lines.append(f"{b - a}\n")
            for i in range(a, b):
                lines.append("%d %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e %.7e\n"
                             % ((buf.monval[i],) + tuple(vals[i])))
        text = "".join(lines).encode()
        import gzip
        paths = {"text_plain": os.path.join(d, "s.txt"), "text_gz": os.path.join(d, "s.gz"),
                 "binary_mode9": os.path.join(d, "s.bin")}
        with open(paths["text_plain"], "wb") as f:
            f.write(text)
        with open(paths["text_gz"], "wb") as f:
            f.write(gzip.compress(text, 6))
        with open(paths["binary_mode9"], "wb") as f:
            for e in range(buf.nev):
                a, b = int(buf.ev_off[e]), int(buf.ev_off[e + 1])
                np.array([b - a], np.int32).tofile(f)
                rec = np.zeros(b - a, dtype=[("m", "<i4"), ("v", "<f4", (9,))])
                rec["m"] = buf.monval[a:b]
                rec["v"] = vals[a:b]
                rec.tofile(f)
        out["text_bytes"] = len(text)
        for name, path in paths.items():
            t0 = time.perf_counter()
            got = hbt.read_buffers(path, 10 ** 9, 211, read_in_mode=9 if "mode9" in name else 10)
            dt = time.perf_counter() - t0
            assert sum(b.n for b, _ in got) == buf.n
            out[name] = {"s": dt, "particles_per_s": buf.n / dt}
        t0 = time.perf_counter()
        ref = oracle.read_mode10(paths["text_gz"], 10 ** 9, 211)
        dt = time.perf_counter() - t0
        assert sum(b.n for b, _ in ref) == buf.n
        out["oracle_text_gz"] = {"s": dt, "particles_per_s": buf.n / dt,
                                 "what": "oracle/hbt_oracle.c restatement of the reference's "
                                         "mode-10 reader (gz_readline, strtod)"}
    out["speedup_text_gz_vs_oracle"] = out["oracle_text_gz"]["s"] / out["text_gz"]["s"]
    log(f"  ingest: text.gz {out['text_gz']['particles_per_s']:.3e} particles/s, "
        f"oracle reader {out['oracle_text_gz']['particles_per_s']:.3e}")
    return out


def copy_calibration(nbytes, reps):
    """A plain device copy of known size (torch's copy kernel), for scaling
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": reps}), flush=True)


TRAFFIC_SUMMARY = os.path.join(REPO, "profiles", "k_fast_traffic.md")
SQ_SUMMARY = os.path.join(REPO, "profiles", "k_fast_sq_counters.md")
ATOMICS_BENCH = os.path.join(REPO, "hadronic_afterburner_toolkit_amd", "hist_atomics_bench.e")


def atomic_ceiling(log):
    """Chip-wide rates of histogram atomics, in (instruction, cache line)
    accesses per second, from the histogram-atomics micro-benchmark run as a
    child process before this process touches the GPU: "same" = the f64coop5
    pattern (5 lanes per bin, 12 bins per FP64 atomic instruction, the fast
    kernel's same-event scatter), "mixed" = u32carry_1_per_lane (the
    denominator's u32 add with carry).  None when the binary is missing or
    fails."""
    if not os.path.exists(ATOMICS_BENCH):
        return None
    try:
        r = subprocess.run([ATOMICS_BENCH], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True, timeout=120)
    except (OSError, subprocess.TimeoutExpired) as e:
        log(f"atomics micro-benchmark failed: {e}")
        return None
    rates, same_n, mixed_n = {}, {}, {}
    for line in r.stdout.splitlines():
        parts = line.split()
        if r.returncode == 0 and len(parts) >= 4:
            if parts[0] == "f64coop5":
                rates["same"] = float(parts[3])
            elif parts[0] == "u32carry_1_per_lane":
                rates["mixed"] = float(parts[3])
            elif parts[0].startswith("f64coop5_bins_"):
                same_n[int(parts[0].rsplit("_", 1)[1])] = float(parts[3])
            elif parts[0].startswith("u32carry_bins_"):
                mixed_n[int(parts[0].rsplit("_", 1)[1])] = float(parts[3])
    if len(rates) != 2:
        log("atomics micro-benchmark printed no f64coop5 / u32carry_1_per_lane line")
        return None
    # the same two patterns at the K_phi workloads' histogram sizes
    rates["by_bins"] = {n: (same_n[n], mixed_n[n]) for n in same_n if n in mixed_n}
    return rates


def kernel_build_hash():
    """sha256 (16 hex digits) of the sources the fast kernel and its slab
    binning are built from (csrc/*.h, csrc/hbt_engine.hip)."""
    import glob
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(REPO, "hadronic_afterburner_toolkit_amd", "csrc")
    for p in sorted(glob.glob(os.path.join(csrc, "*.h")) + [os.path.join(csrc, "hbt_engine.hip")]):
        with open(p, "rb") as f:
            h.update(os.path.basename(p).encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def profile_value(path, key, kernel_prefix):
    """A per-step figure of the headline mode from a committed rocprofv3
    counter summary (lines '<key> <mode> N <build hash>'), with the build
    hash of the sources it was measured on and the current one; the value is
    None when they differ (the kernel changed since the counters were
    read)."""
    cur = kernel_build_hash()
    try:
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == key and parts[1] == kernel_prefix:
                    built = parts[3] if len(parts) > 3 else None
                    return (float(parts[2]) if built == cur else None), built, cur
    except OSError:
        pass
    return None, None, cur


def measured_traffic(mode):
    """HBM bytes per step of the headline mode (fp32c, fp64, or with _slab:
    the step's k_fast, k_slab_bin and k_slab_reduce) from
    profiles/k_fast_traffic.md."""
    return profile_value(TRAFFIC_SUMMARY, "traffic_bytes_per_step", mode)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--events", type=int, default=2000)
    ap.add_argument("--mult", type=float, default=600.0)
    ap.add_argument("--per-group", type=int, default=50)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--only", choices=["fp32c", "fp64", "fp32c_slab", "fp64_slab", "plain"],
                    default=None)
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving form")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched-uneven and large-histogram lines")
    ap.add_argument("--serve-requests", type=int, default=60)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-step-pairs", type=float, default=2e9)
    ap.add_argument("--tuning", default=None, metavar="R,JCHUNK,KCHUNK,BLOCKS_PER_CU",
                    help="fixed launch shape of the fast kernel (implies --no-autotune)")
    ap.add_argument("--only-bf", action="store_true",
                    help="only the balance-function line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the HBT engine runs on a single MI355X here "
              "(sharding over several GPUs is a design only: groups split across GPUs, histograms summed at the end)", file=sys.stderr)
        sys.exit(2)

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(1, args.steps))
        return
    if args.only_bf:
        print(json.dumps({"balance_function": balance_function(args, log)}), flush=True)
        return

    tuning = [int(v) for v in args.tuning.split(",")] if args.tuning else None
    # the roofline needs the plain kernel's stage counts
    atom_ceil = atomic_ceiling(log) if args.only in (None, "plain") else None
    from hadronic_afterburner_toolkit_amd import _lib
    log(f"generating {args.events} events, Poisson({args.mult}), groups of {args.per_group}")
    groups = make_workload(args.events, args.mult, args.per_group, args.seed)
    plan = draw_plans(groups, args.seed + 1, PARAMS["number_of_mixed_events"])
    res = {}
    # fp32c / fp64: the fast kernel with the histogram privatisation the
    # autotuner measures faster (hbt_binning 2: device-scope atomics or slab
    # binning; atomics when the launch shape is given with --tuning); the
    # *_slab modes force slab binning (binned pairs summed per (block, q_out
    # plane) in LDS), for comparison
    modes = {"fp32c": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, args.steps, 2),
             "fp64": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, args.steps, 2),
             "fp32c_slab": (_lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, args.steps, 1),
             "fp64_slab": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_FAST, args.steps, 1),
             "plain": (_lib.HBT_PREC_FP64, _lib.HBT_KERNEL_PLAIN, max(1, min(2, args.steps)), 0)}
    ctx = {}
    for name, (prec, kern, k, binning) in modes.items():
        if args.only and name != args.only:
            continue
        log(f"mode {name}")
        ctx[name], res[name] = run_mode(groups, plan, prec, kern, k,
                                        args.warmup if name != "plain" else 1,
                                        not (args.no_autotune or tuning), log, tuning, binning)

    def pairs_per_s(r):
        return (r["enum_same"] + r["enum_mixed"]) * len(r["dev_ms"]) / r["wall"]

    head = "fp32c" if "fp32c" in res else next(iter(res))
    r = res[head]
    out = {
        "metric": "pairs/sec (same+mixed) FP32-comp vs FP64 at <=1e-6 bin err; p50/p99 latency",
        "value": pairs_per_s(r),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": len(r["dev_ms"]),
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["wall"] / len(r["dev_ms"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32c" else "f64",
        "data": "synthetic",
        "config": {"workload": "precision form: 2000 pi+ events Poisson(600), oversample 50 "
                               "(40 groups ~30k particles), 10 mixed rotations, q 41^3, 4(5) K_T bins",
                   "events": args.events, "mult": args.mult, "per_group": args.per_group,
                   "pairs_per_step": r["enum_same"] + r["enum_mixed"],
                   "same_pairs_per_step": r["enum_same"], "mixed_pairs_per_step": r["enum_mixed"],
                   "tuning": r["tune"]},
    }
    for name in res:
        out[f"{name}_pairs_per_s"] = pairs_per_s(res[name])
        out[f"{name}_dev_ms_per_step"] = float(np.mean(res[name]["dev_ms"]))
        out[f"{name}_upload_s"] = res[name]["upload_s"]
        out[f"{name}_upload_warm_s"] = res[name]["upload_warm_s"]
        out[f"{name}_prepare_warm_s"] = res[name]["prepare_warm_s"]
        out[f"{name}_batch_graph"] = res[name]["graph"]
        out[f"{name}_slab"] = res[name]["slab"]
    if "fp32c" in res and "plain" in res:
        out["speedup_vs_plain_fp64"] = out["fp32c_pairs_per_s"] / out["plain_pairs_per_s"]
    for form in ("fp32c", "fp64"):
        if form in res and form + "_slab" in res:
            out[f"speedup_slab_vs_atomics_{form}"] = (out[f"{form}_slab_pairs_per_s"]
                                                     / out[f"{form}_pairs_per_s"])
    if "fp32c" in res and "fp64" in res:
        out["speedup_vs_fast_fp64"] = out["fp32c_pairs_per_s"] / out["fp64_pairs_per_s"]
        a, b = ctx["fp32c"].results(), ctx["fp64"].results()
        out["counts_identical_fp32c_vs_fp64"] = bool(
            np.array_equal(a.num_count, b.num_count) and np.array_equal(a.den, b.den))
        cnt = b.num_count.astype(float)
        sel = cnt > 0
        per_pair = np.abs(a.num - b.num)[sel] / cnt[sel]
        out["num_err_per_pair_max"] = float(per_pair.max()) if per_pair.size else 0.0

    # roofline of the dominant kernel: stage counts from the plain kernel
    # (it counts every stage of the cascade), FLOP model per launch
    if "plain" in res:
        s, m = ctx["plain"].stage_counts()
        nsteps = len(res["plain"]["dev_ms"])
        s = [int(v) // nsteps for v in s]
        m = [int(v) // nsteps for v in m]
        fl = flops(s, m)
        dev_s = float(np.mean(r["dev_ms"])) * 1e-3
        ach = fl / dev_s / 1e12
        peak = FP32_PEAK_TFLOPS if head == "fp32c" else FP64_PEAK_TFLOPS
        # the step's kernels: k_fast alone with device-scope atomic binning;
        # with slab binning also k_slab_bin and k_slab_reduce (the mode the
        # autotuner chose, in the profiles' keys)
        slab_on = bool(r["slab"]["on"])
        pn = "float" if head == "fp32c" else "double"
        pmode = head + ("_slab" if slab_on else "")
        kname = (f"k_fast<{pn}> + k_slab_bin<{pn}> + k_slab_reduce" if slab_on
                 else f"k_fast<{pn}>")
        traffic, t_build, cur_build = measured_traffic(pmode)
        out["roofline"] = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                           "frac": ach / peak, "traffic": traffic,
                           "traffic_source": "from_profile (profiles/k_fast_traffic.md)",
                           "traffic_build": t_build, "kernel_build": cur_build,
                           "kernel": kname, "scope": "one step: every kernel of the batch launch",
                           "profile_mode": pmode,
                           "flop_per_launch": fl, "stages_same": s, "stages_mixed": m}
        # issue-side view of the same step: wave64 VALU instructions per step
        # (SQ_INSTS_VALU summed over its kernels, committed counter summary of
        # this build), each holding its SIMD 4 cycles, over the step on 1024
        # SIMDs at the 2.4 GHz peak clock
        insts, _, _ = profile_value(SQ_SUMMARY, "valu_insts_per_step", pmode)
        out["roofline"]["valu_issue_frac"] = (
            insts * 4 / (MI355X_CUS * 4 * MI355X_GHZ * 1e9 * dev_s) if insts else None)
        out["roofline"]["valu_insts_per_launch"] = insts
        # device-scope atomics: with atomic binning every binned pair is one
        # atomic cache-line access (same-event: count + four sums in one FP64
        # instruction; mixed: the denominator's u32 add); with slab binning
        # only the mixed pairs' adds are.  Floors from the chip-wide rates of
        # those patterns measured in this run.
        acc_same = 0 if slab_on else s[4]
        floor = ((acc_same / atom_ceil["same"] + m[4] / atom_ceil["mixed"]) if atom_ceil else None)
        floor_all = ((s[4] / atom_ceil["same"] + m[4] / atom_ceil["mixed"]) if atom_ceil else None)
        amode = head if slab_on else head + "_slab"
        a_dev = float(np.mean(res[amode]["dev_ms"])) * 1e-3 if amode in res else None
        out["atomic_roofline"] = {
            "bound": "memory_side_atomics", "binning": "slab" if slab_on else "atomics",
            "accesses_per_launch": acc_same + m[4], "accesses_same": acc_same,
            "accesses_mixed": m[4], "unit": "atomic line accesses/s",
            "achieved": (acc_same + m[4]) / dev_s,
            "floor_ms": 1e3 * floor if floor else None,
            "frac": floor / dev_s if floor else None,
            "atomic_binning_floor_ms": 1e3 * floor_all if floor_all else None,
            "step_below_atomic_binning_floor": bool(floor_all and dev_s < floor_all),
            "other_binning_mode": amode, "other_binning_ms": a_dev * 1e3 if a_dev else None,
            "kernel": f"k_fast<{pn}>",
            "ceiling_source": "hist_atomics_bench.e f64coop5 (same-event) and "
                              "u32carry_1_per_lane (mixed), measured in this run"}
    if not args.no_extra and not args.only:
        out["dropin"] = dropin_line(groups, args.seed + 1, log)
        out["dropin"]["kernel_ms_per_buffer"] = float(np.mean(res[head]["dev_ms"])) / len(groups)
    if not args.no_serve:
        out["serving"] = serve(args, log)
    if not args.no_extra:
        t = res[head]["tune"]
        out["batched_uneven"] = batched_uneven(args, log)
        out["large_histogram"] = large_histogram(
            args, log, (t["R"], t["jchunk"], t["kchunk"], t["blocks_per_cu"]))
        out["kphi_binning"] = kphi_binning(args, log, groups)
        for line, nb in ((out["batched_uneven"]["pi+"], BENCH_BINS * 16),
                         (out["batched_uneven"]["K+"], BENCH_BINS * 16),
                         (out["large_histogram"], out["large_histogram"]["bins"]),
                         (out["kphi_binning"]["atomics"], BENCH_BINS * 16)):
            add_atomic_floor(line, atom_ceil, nb)
    if not args.no_extra:
        out["ingest"] = ingest(args, log)
        out["balance_function"] = balance_function(args, log)
    if not args.no_cpu:
        g0 = groups[0]
        M = int(plan[0][0])
        plan_g = (M, plan[1][:g0.nev * M], plan[2][:2 * g0.nev * M])
        o, a, rate, dt, pairs = cpu_baseline(g0, plan_g, log)
        out["cpu_baseline"] = {"value": rate, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"group 0 of {len(groups)} ({g0.nev} events, "
                                         f"{g0.n} particles): {pairs:.3e} same+mixed pairs in "
                                         f"{dt:.1f}s, oracle/hbt_oracle.c -O2"}
        out["cpu_baseline_multicore"], multi = cpu_baseline_multicore(groups, plan, log)
        # precision form vs the FP64 oracle per bin: on group 0, and as one
        # batch on the groups the multicore oracle ran (its histograms summed)
        if "fp32c" in res:
            n_multi = int(out["cpu_baseline_multicore"]["cores"])
            rows = sum(int(plan[0][g]) * groups[g].nev for g in range(n_multi))
            for label, gsel, pl, ref_a in (
                    ("fp32c_vs_fp64_oracle", [g0], (np.array([M], np.int32), plan_g[1], plan_g[2]), a),
                    (f"fp32c_vs_fp64_oracle_{n_multi}_groups", groups[:n_multi],
                     (plan[0][:n_multi], plan[1][:rows], plan[2][:2 * rows]), multi)):
                out[label] = precision_vs_oracle(gsel, pl, ref_a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
