"""The drop-in CLI (hbt_afterburner.e: ParameterReader, particleSamples,
HBT_correlation over the C ABI) end to end on the GPU, against the files the
oracle writes for the same inputs (the oracle's files are byte-identical to
the reference's: tests/test_oracle_golden.py).

GPU sums are accumulated in a different order than the reference's loop, so
a printed 8-digit value can differ in its last digit when a sum sits on a
rounding boundary; the bar for the FP64 mode is: identical file names and
line counts, every number within 2e-8 relative (one unit in the 8th digit)
and at least 99.9% of lines byte-identical.  For the precision form the
cos sums may differ by 1e-6 per pair (written tolerance: 1e-4 absolute on
the sum column of these small cases), the other columns as in FP64.
"""
import os
import subprocess

import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import samples

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "hadronic_afterburner_toolkit_amd", "hbt_afterburner.e")


def write_parameters(path, params):
    with open(path, "w") as f:
        f.write("# HBT path parameters (key = value  # comment)\n")
        for k, v in params.items():
            f.write(f"{k} = {v}    # {k}\n")


@pytest.mark.parametrize("name", ["fixture_3d", "fixture_kphi_qinv", "synth_cfg0",
                                  "synth_mixedfile", "synth_budget_exact", "synth_cap_qinv"])
@pytest.mark.parametrize("precision", [0, 1])
def test_cli_matches_reference_files(name, precision, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    par = dict(spec["params"])
    work = tmp_path / "work"
    res = work / "results"
    res.mkdir(parents=True)
    paths = cases.write_inputs(spec, dig, str(res))
    keys = dict(cases.hbt_params(spec))
    keys.update(analyze_HBT=1, read_in_mode=10, particle_monval=211,
                event_buffer_size=int(par["event_buffer_size"]),
                read_in_real_mixed_events=int(par.get("read_in_real_mixed_events", 0)),
                randomSeed=int(par["randomSeed"]), hbt_precision=precision)
    write_parameters(str(work / "parameters.dat"), keys)
    print(f"cli case {name} precision {precision}", flush=True)
    r = subprocess.run([CLI], cwd=str(work), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # oracle files for the same inputs
    ebs = int(par["event_buffer_size"])
    bufs = oracle.read_mode10(paths["same"], ebs, 211)
    mixb = oracle.read_mode10(paths["mixed"], ebs, 211) if "mixed" in paths else None
    o = oracle.Oracle(cases.hbt_params(spec))
    rng = oracle.OracleRng(int(par["randomSeed"]))
    for k, (b, _) in enumerate(bufs):
        o.calculate_buffer(b, rng, mixb[k][0] if mixb else None)
    od = tmp_path / "oracle"
    od.mkdir()
    o.write_outputs(str(od), 0)
    got = sorted(f for f in os.listdir(res) if f.endswith(".dat"))
    assert got == sorted(os.listdir(od)) == sorted(outs)
    total = same = 0
    for fn in got:
        a = (res / fn).read_text().splitlines()
        b = (od / fn).read_text().splitlines()
        assert len(a) == len(b)
        sum_col = 1 if "_inv_" in fn else 3
        for la, lb in zip(a, b):
            total += 1
            if la == lb:
                same += 1
                continue
            va = np.array([float(x) for x in la.split()])
            vb = np.array([float(x) for x in lb.split()])
            ok = np.isclose(va, vb, rtol=2e-8, atol=0) | (np.isnan(va) & np.isnan(vb))
            if precision == 1:
                ok[sum_col] |= abs(va[sum_col] - vb[sum_col]) <= 1e-4
            assert ok.all(), (fn, la, lb)
    if precision == 0:
        assert same >= 0.999 * total, (same, total)


def test_cli_mode9_binary(tmp_path):
    """read_in_mode 9 (results/particle_list.bin, float records) end to end:
    the CLI's files against the oracle's on the same particles (the floats
    widened to double) and the same buffering; bar as for FP64 above."""
    full = samples.generate(1011, 14, 120.0)
    f32 = {k: getattr(full, k).astype(np.float32) for k in
           ("mass", "t", "x", "y", "z", "E", "px", "py", "pz")}
    work = tmp_path / "work"
    res = work / "results"
    res.mkdir(parents=True)
    with open(res / "particle_list.bin", "wb") as f:
        for e in range(full.nev):
            a, b = int(full.ev_off[e]), int(full.ev_off[e + 1])
            np.array([b - a], np.int32).tofile(f)
            rec = np.zeros(b - a, dtype=[("m", "<i4"), ("v", "<f4", (9,))])
            rec["m"] = full.monval[a:b]
            rec["v"] = np.stack([f32[k][a:b] for k in ("mass", "t", "x", "y", "z", "E", "px",
                                                        "py", "pz")], axis=1)
            rec.tofile(f)
    hp = cases.hbt_params({"params": dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=4, KT_min=0.2,
                                          KT_max=0.8, HBTrap_min=-1.0, HBTrap_max=1.0,
                                          needed_number_of_pairs=1e18)})
    keys = dict(hp, analyze_HBT=1, read_in_mode=9, particle_monval=211, event_buffer_size=500,
                read_in_real_mixed_events=0, randomSeed=31, hbt_precision=0)
    write_parameters(str(work / "parameters.dat"), keys)
    r = subprocess.run([CLI], cwd=str(work), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    wide = samples.Buffer(ev_off=full.ev_off, **{k: v.astype(np.float64) for k, v in f32.items()},
                          monval=full.monval)
    o = oracle.Oracle(hp)
    rng = oracle.OracleRng(31)
    for first, count in samples.buffers_by_budget(wide, 500):
        o.calculate_buffer(wide.events(first, count), rng)
    od = tmp_path / "oracle"
    od.mkdir()
    o.write_outputs(str(od), 0)
    got = sorted(f for f in os.listdir(res) if f.endswith(".dat"))
    assert got == sorted(os.listdir(od)) and got
    for fn in got:
        a = (res / fn).read_text().splitlines()
        b = (od / fn).read_text().splitlines()
        assert len(a) == len(b)
        for la, lb in zip(a, b):
            if la != lb:
                va = np.array([float(x) for x in la.split()])
                vb = np.array([float(x) for x in lb.split()])
                assert (np.isclose(va, vb, rtol=2e-8, atol=0) | (np.isnan(va) & np.isnan(vb))).all(), (fn, la, lb)


def test_cli_species_class(tmp_path):
    """particle_monval = 9998 (positive hadrons of several masses) through the
    CLI, which loads EOS/pdg.dat for the class as the reference's
    particle_decay does; against the oracle on the same particles, picked
    by charge from the same table, with the same buffering (the budget counts
    all species).  Bar as for FP64 above."""
    species = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0),
               (111, 0.13498, 0.25, 1.0), (321, 0.49368, 0.3, 1.0), (-321, 0.49368, 0.3, 1.0),
               (2212, 0.93827, 0.35, 0.6), (-2212, 0.93827, 0.35, 0.4))
    full = samples.generate(1021, 12, 140.0, species=species)
    work = tmp_path / "work"
    res = work / "results"
    res.mkdir(parents=True)
    (work / "EOS").mkdir()
    table = os.path.join(cases.HERE, "species_table.dat")
    with open(table) as f:
        (work / "EOS" / "pdg.dat").write_text(f.read())
    samples.write_mode10(str(res / "particle_samples.gz"), full)
    hp = cases.hbt_params({"params": dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=4, KT_min=0.2,
                                          KT_max=0.8, HBTrap_min=-1.0, HBTrap_max=1.0,
                                          needed_number_of_pairs=1e18)})
    keys = dict(hp, analyze_HBT=1, read_in_mode=10, particle_monval=9998, event_buffer_size=700,
                read_in_real_mixed_events=0, randomSeed=37, hbt_precision=0)
    write_parameters(str(work / "parameters.dat"), keys)
    print("cli species class 9998", flush=True)
    r = subprocess.run([CLI], cwd=str(work), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    sp = oracle.species_read(table)
    charge = dict(zip(sp["monval"].tolist(), sp["charge"].tolist()))
    keep_all = np.array([charge[int(m)] > 0 for m in full.monval])
    o = oracle.Oracle(hp)
    rng = oracle.OracleRng(37)
    for first, count in samples.buffers_by_budget(full, 700):
        b = full.events(first, count)
        keep = keep_all[int(full.ev_off[first]):int(full.ev_off[first + count])]
        cnt = [int(keep[int(b.ev_off[e]):int(b.ev_off[e + 1])].sum()) for e in range(b.nev)]
        sel = np.flatnonzero(keep)
        kept = samples.Buffer(ev_off=np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64),
                              px=b.px[sel], py=b.py[sel], pz=b.pz[sel], E=b.E[sel], x=b.x[sel],
                              y=b.y[sel], z=b.z[sel], t=b.t[sel], monval=b.monval[sel],
                              mass=b.mass[sel])
        o.calculate_buffer(kept, rng)
    od = tmp_path / "oracle"
    od.mkdir()
    o.write_outputs(str(od), 0)
    got = sorted(f for f in os.listdir(res) if f.endswith(".dat"))
    assert got == sorted(os.listdir(od)) and got
    for fn in got:
        a = (res / fn).read_text().splitlines()
        b = (od / fn).read_text().splitlines()
        assert len(a) == len(b)
        for la, lb in zip(a, b):
            if la != lb:
                va = np.array([float(x) for x in la.split()])
                vb = np.array([float(x) for x in lb.split()])
                assert (np.isclose(va, vb, rtol=2e-8, atol=0)
                        | (np.isnan(va) & np.isnan(vb))).all(), (fn, la, lb)


@pytest.mark.parametrize("name", cases.BF_CASE_NAMES)
def test_cli_balance_function_files_equal_reference(name, tmp_path):
    """analyze_balance_function = 1 through the CLI (drop-in BalanceFunction
    and particleSamples over the C ABI, mixing draws from the shared random
    stream): the three files byte-identical to the reference's on every
    balance-function golden case (species classes read EOS/pdg.dat)."""
    import hashlib
    import io
    import gzip
    spec, ref, outs, dig = cases.load(name)
    full = cases.bf_input(spec, dig)
    work = tmp_path / "work"
    res = work / "results"
    res.mkdir(parents=True)
    text = io.StringIO()
    for e in range(full.nev):
        a, b = int(full.ev_off[e]), int(full.ev_off[e + 1])
        text.write(f"{b - a}\n")
        for i in range(a, b):
            text.write("%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                full.monval[i], full.mass[i], full.t[i], full.x[i], full.y[i], full.z[i],
                full.E[i], full.px[i], full.py[i], full.pz[i]))
    with gzip.open(res / "particle_samples.gz", "wb") as f:
        f.write(text.getvalue().encode())
    table = cases.bf_species_table(spec)
    if table:
        (work / "EOS").mkdir()
        with open(table) as f:
            (work / "EOS" / "pdg.dat").write_text(f.read())
    keys = dict(spec["params"], analyze_HBT=0, analyze_balance_function=1)
    write_parameters(str(work / "parameters.dat"), keys)
    print(f"cli balance function {name}", flush=True)
    r = subprocess.run([CLI], cwd=str(work), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for fn, meta in outs.items():
        with open(res / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == meta["sha256"], fn
