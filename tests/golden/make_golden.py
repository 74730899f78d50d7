"""Writes the golden fixtures under tests/golden/ by running the reference's
own HBT path (compiled from its sources by oracle/Makefile into
oracle/_ref/) on fixed inputs.  Runs only where /root/reference exists.

For each case: tests/golden/<case>/
  params.json   the parameter overrides and input description
  ref.npz       raw histograms and counters dumped by oracle/ref_harness
                (bit-exact doubles), compressed
  outputs.json  sha256 and first lines of every .dat file the reference wrote
  input.sha256  digest of the mode-10 input file(s) the case used

Usage: python tests/golden/make_golden.py [case ...]
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from hadronic_afterburner_toolkit_amd import samples  # noqa: E402

sys.path.insert(0, HERE)
import edge_inputs  # noqa: E402

REF = "/root/reference"
HARNESS = os.path.join(REPO, "oracle", "_ref", "ref_harness")
FIXTURE_GZ = os.path.join(HERE, "ref_fixture", "particle_samples.gz")

CASES = {
    # the reference's own HBT fixture with the reference's default HBT keys
    "fixture_3d": dict(input="fixture", params=dict(randomSeed=0)),
    "fixture_q11": dict(input="fixture", params=dict(randomSeed=3, qnpts=11, n_KT=3,
                                                      KT_min=0.2, KT_max=0.6)),
    "fixture_kphi_qinv": dict(input="fixture", params=dict(
        randomSeed=5, qnpts=11, n_KT=3, KT_min=0.2, KT_max=0.6, azimuthal_flag=1, n_Kphi=4,
        invariant_radius_flag=1)),
    "fixture_budget700": dict(input="fixture", params=dict(randomSeed=21, event_buffer_size=700,
                                                            qnpts=15)),
    "fixture_qinv": dict(input="fixture", params=dict(
        randomSeed=9, qnpts=15, invariant_radius_flag=1, HBTrap_min=-1.0, HBTrap_max=1.0)),
    # config-0 source model, small: several buffers, K_T edges 0.2..0.7 step 0.1
    "synth_cfg0": dict(input=("synth", 1001, 30, 150.0), params=dict(
        randomSeed=7, qnpts=41, n_KT=6, KT_min=0.2, KT_max=0.7, HBTrap_min=-1.0,
        HBTrap_max=1.0, needed_number_of_pairs=1e18, event_buffer_size=1000)),
    "synth_noboost": dict(input=("synth", 1002, 12, 120.0), params=dict(
        randomSeed=11, qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8, long_comoving_boost=0,
        HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18, event_buffer_size=600)),
    # binding pair cap (order-dependent semantics), counters persist across buffers
    "synth_cap": dict(input=("synth", 1003, 16, 100.0), params=dict(
        randomSeed=13, qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8, HBTrap_min=-1.0,
        HBTrap_max=1.0, needed_number_of_pairs=40, event_buffer_size=500)),
    # binding caps with K_phi counters (per K_T x K_phi bin) and the q_inv cap
    # (50 x needed per K_T bin)
    "fixture_cap_kphi_qinv": dict(input="fixture", params=dict(
        randomSeed=23, qnpts=11, n_KT=3, KT_min=0.2, KT_max=0.6, azimuthal_flag=1, n_Kphi=4,
        invariant_radius_flag=1, needed_number_of_pairs=6)),
    "synth_cap_qinv": dict(input=("synth", 1007, 12, 100.0), params=dict(
        randomSeed=29, qnpts=15, n_KT=3, KT_min=0.2, KT_max=0.6, HBTrap_min=-1.0,
        HBTrap_max=1.0, invariant_radius_flag=1, needed_number_of_pairs=5,
        event_buffer_size=400)),
    # separate mixed-event file (read_in_real_mixed_events=1)
    "synth_mixedfile": dict(input=("synth2", 1004, 1005, 10, 120.0), params=dict(
        randomSeed=17, qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8, HBTrap_min=-1.0,
        HBTrap_max=1.0, needed_number_of_pairs=1e18, event_buffer_size=800,
        read_in_real_mixed_events=1)),
    # budget met exactly at the end of the file: one extra empty buffer
    "synth_budget_exact": dict(input=("synth", 1006, 6, 80.0), params=dict(
        randomSeed=19, qnpts=15, n_KT=3, KT_min=0.2, KT_max=0.6, HBTrap_min=-1.0,
        HBTrap_max=1.0, needed_number_of_pairs=1e18, event_buffer_size="total")),
    # hand-placed pairs exactly on the window edges lo+1e-8 / hi-1e-8, one ulp
    # around them, on interior q bin edges, at K_perp = KT_min and KT_max
    # (edge_inputs.py); the bench's grid
    "edge_3d": dict(input=("edges", "3d"), params=dict(
        randomSeed=31, qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
        HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18)),
    # pair axes exactly on K_phi bin edges about psi_ref = 0, and one that
    # wraps to 2 pi (skipped by the reference)
    "edge_kphi": dict(input=("edges", "kphi"), params=dict(
        randomSeed=37, qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
        HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18, azimuthal_flag=1,
        n_Kphi=4)),
    # binding per-(K_T, K_phi) caps on a larger sample: with a wide K_phi
    # guard band the engine decides many pairs on the host, which must be
    # admitted in the reference's enumeration order
    "synth_cap_kphi": dict(input=("synth", 1008, 20, 120.0), params=dict(
        randomSeed=41, qnpts=15, n_KT=3, KT_min=0.2, KT_max=0.6, HBTrap_min=-1.0,
        HBTrap_max=1.0, azimuthal_flag=1, n_Kphi=6, needed_number_of_pairs=300,
        event_buffer_size=700)),
}

BASE = dict(read_in_mode=10, particle_monval=211, analyze_flow=0, analyze_HBT=1,
            analyze_balance_function=0, analyze_ebe_yield=0, resonance_feed_down_flag=0,
            resonance_weak_feed_down_flag=0, read_in_real_mixed_events=0,
            event_buffer_size=100000, echo_level=0)


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def read_dump(path):
    with open(path, "rb") as f:
        raw = f.read()
    hdr = np.frombuffer(raw[:32], dtype=np.int32)
    magic, nkt, nphi, q, az, qi, nbuf, _ = [int(v) for v in hdr]
    assert magic == 0x48425431
    off = 32
    psi, tcalc = np.frombuffer(raw[off:off + 16], dtype=np.float64)
    off += 16
    np_ = nphi if az else 1
    nb = nkt * np_ * q ** 3
    out = {}
    for name in ("num_count", "q_out_sum", "q_side_sum", "q_long_sum", "num", "den"):
        out[name] = np.frombuffer(raw[off:off + 8 * nb], dtype=np.float64).copy()
        off += 8 * nb
    if qi:
        for name in ("qinv_count", "qinv_sum", "qinv_num", "qinv_den"):
            out[name] = np.frombuffer(raw[off:off + 8 * nkt * q], dtype=np.float64).copy()
            off += 8 * nkt * q
    for name in ("cnt_num_kt", "cnt_den_kt", "cnt_num_qinv", "cnt_den_qinv"):
        out[name] = np.frombuffer(raw[off:off + 8 * nkt], dtype=np.uint64).copy()
        off += 8 * nkt
    if az:
        for name in ("cnt_num_ktphi", "cnt_den_ktphi"):
            out[name] = np.frombuffer(raw[off:off + 8 * nkt * nphi], dtype=np.uint64).copy()
            off += 8 * nkt * nphi
    nb_ = int(np.frombuffer(raw[off:off + 4], dtype=np.int32)[0])
    off += 4
    out["nev_per_buffer"] = np.frombuffer(raw[off:off + 4 * nb_], dtype=np.int32).copy()
    out["psi_ref"] = np.float64(psi)
    out["t_calc"] = np.float64(tcalc)
    return out


def run_case(name, spec, keep_dir=None):
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    res = os.path.join(work, "results")
    os.makedirs(res)
    os.symlink(os.path.join(REF, "EOS"), os.path.join(work, "EOS"))
    shutil.copy(os.path.join(REF, "parameters.dat"), os.path.join(work, "parameters.dat"))
    inp = spec["input"]
    params = dict(BASE)
    params.update(spec["params"])
    digests = {}
    if inp == "fixture":
        shutil.copy(FIXTURE_GZ, os.path.join(res, "particle_samples.gz"))
    elif inp[0] == "synth":
        _, seed, nev, mult = inp
        buf = samples.generate(seed, nev, mult)
        samples.write_mode10(os.path.join(res, "particle_samples.gz"), buf)
        if params.get("event_buffer_size") == "total":
            params["event_buffer_size"] = int(buf.n)
    elif inp[0] == "edges":
        samples.write_mode10(os.path.join(res, "particle_samples.gz"), edge_inputs.build(inp[1]))
    elif inp[0] == "synth2":
        _, seed, seed2, nev, mult = inp
        samples.write_mode10(os.path.join(res, "particle_samples.gz"),
                             samples.generate(seed, nev, mult))
        samples.write_mode10(os.path.join(res, "particle_samples_mixed_event.gz"),
                             samples.generate(seed2, nev + 3, mult))
    for fn in sorted(os.listdir(res)):
        with gzip.open(os.path.join(res, fn), "rb") as f:
            digests[fn] = hashlib.sha256(f.read()).hexdigest()
    dump = os.path.join(work, "dump.bin")
    args = [HARNESS, work, dump] + [f"{k}={v}" for k, v in params.items()]
    r = subprocess.run(args, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{name}: ref_harness failed: {r.stderr[-2000:]}")
    d = read_dump(dump)
    outdir = os.path.join(HERE, name)
    os.makedirs(outdir, exist_ok=True)
    np.savez_compressed(os.path.join(outdir, "ref.npz"), **d)
    outs = {}
    for fn in sorted(os.listdir(res)):
        if fn.endswith(".dat"):
            p = os.path.join(res, fn)
            with open(p) as f:
                head = [next(f, "") for _ in range(3)]
            outs[fn] = dict(sha256=sha(p), head=head, size=os.path.getsize(p))
    with open(os.path.join(outdir, "outputs.json"), "w") as f:
        json.dump(outs, f, indent=1, sort_keys=True)
    with open(os.path.join(outdir, "params.json"), "w") as f:
        json.dump(dict(input=inp, params=params), f, indent=1)
    with open(os.path.join(outdir, "input.sha256"), "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)
    if keep_dir:
        shutil.copytree(work, os.path.join(keep_dir, name), dirs_exist_ok=True)
    shutil.rmtree(work)
    return d


def make_psi_case():
    """psi_ref golden of HBT_unittest.cc:23-48 runs on the UrQMD-zipped
    fixture (unit_tests/parameters.dat: read_in_mode=2).  The reference reads
    it; the (px, py) of the pi+ it kept are stored with its psi_ref values,
    so the oracle's psi_ref can be pinned without restating that reader."""
    work = tempfile.mkdtemp(prefix="golden_psi_")
    os.makedirs(os.path.join(work, "results"))
    os.symlink(os.path.join(REF, "EOS"), os.path.join(work, "EOS"))
    shutil.copy(os.path.join(REF, "unit_tests", "parameters.dat"),
                os.path.join(work, "parameters.dat"))
    shutil.copy(os.path.join(REF, "unit_tests", "test_gzip_reader", "particle_list.dat"),
                os.path.join(work, "results", "particle_list.dat"))
    dump = os.path.join(work, "psi.bin")
    keys = dict(harness_psi_only=1, analyze_flow=0, analyze_HBT=1, analyze_balance_function=0,
                analyze_ebe_yield=0, read_in_real_mixed_events=0, long_comoving_boost=1,
                HBTrap_min=-0.5, HBTrap_max=0.5)
    subprocess.run([HARNESS, work, dump] + [f"{k}={v}" for k, v in keys.items()],
                   stdout=subprocess.DEVNULL, check=True)
    with open(dump, "rb") as f:
        raw = f.read()
    psi = np.frombuffer(raw[:24], dtype=np.float64).copy()
    nev = int(np.frombuffer(raw[24:28], dtype=np.int32)[0])
    off, counts, xy = 28, [], []
    for _ in range(nev):
        n = int(np.frombuffer(raw[off:off + 4], dtype=np.int32)[0])
        off += 4
        xy.append(np.frombuffer(raw[off:off + 16 * n], dtype=np.float64).reshape(n, 2))
        off += 16 * n
        counts.append(n)
    xy = np.concatenate(xy)
    outdir = os.path.join(HERE, "psi_urqmd")
    os.makedirs(outdir, exist_ok=True)
    np.savez_compressed(os.path.join(outdir, "particles.npz"), px=xy[:, 0], py=xy[:, 1],
                        ev_off=np.concatenate([[0], np.cumsum(counts)]).astype(np.int64),
                        psi_ref=psi)
    shutil.rmtree(work)
    print("psi_urqmd:", psi, counts)


def main(argv):
    if not os.path.isdir(REF):
        sys.exit("the reference tree is not present")
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "all", "ref"], check=True)
    names = argv or list(CASES) + ["psi"]
    if "psi" in names:
        make_psi_case()
        names = [n for n in names if n != "psi"]
    for n in names:
        d = run_case(n, CASES[n])
        print(f"{n}: buffers={list(d['nev_per_buffer'])} binned_same="
              f"{int(d['num_count'].sum())} binned_mixed={int(d['den'].sum())} "
              f"t_calc={float(d['t_calc']):.3f}s")


if __name__ == "__main__":
    main(sys.argv[1:])
