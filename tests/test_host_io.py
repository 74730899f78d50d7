"""The engine library's host pieces against the reference (CPU only):
text writers byte-identical to the reference's files, the mode-10 reader
equal to the reference's buffering, psi_ref."""
import hashlib
import os

import numpy as np
import pytest

import cases
from hadronic_afterburner_toolkit_amd import hbt


def oracle_result_arrays(o):
    a = o.arrays()
    out = {k: a[k] for k in ("q_out_sum", "q_side_sum", "q_long_sum", "num", "qinv_sum",
                             "qinv_num")}
    for k in ("num_count", "den", "qinv_count", "qinv_den"):
        out[k] = a[k].astype(np.uint64)
    for k in ("cnt_num_kt", "cnt_den_kt", "cnt_num_qinv", "cnt_den_qinv", "cnt_num_ktphi",
              "cnt_den_ktphi"):
        out[k] = a[k]
    return out


@pytest.mark.parametrize("name", cases.CASE_NAMES)
def test_writer_byte_identical_to_reference(oracle_mod, name, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    par = spec["params"]
    paths = cases.write_inputs(spec, dig, str(tmp_path))
    ebs = int(par["event_buffer_size"])
    bufs = oracle_mod.read_mode10(paths["same"], ebs, 211)
    mixb = oracle_mod.read_mode10(paths["mixed"], ebs, 211) if "mixed" in paths else None
    params = cases.hbt_params(spec)
    o = oracle_mod.Oracle(params)
    rng = oracle_mod.OracleRng(int(par["randomSeed"]))
    for k, (b, _) in enumerate(bufs):
        o.calculate_buffer(b, rng, mixb[k][0] if mixb else None)
    out = tmp_path / "out"
    out.mkdir()
    hbt.write_outputs(params, oracle_result_arrays(o), str(out))
    got = sorted(os.listdir(out))
    assert got == sorted(outs)
    for fn in got:
        with open(out / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == outs[fn]["sha256"], fn


def test_writer_eco_output(oracle_mod, tmp_path):
    """ecoOutput=1 writes two columns (HBT_correlation.cpp:767-770), same as the oracle."""
    spec, ref, outs, dig = cases.load("fixture_q11")
    params = cases.hbt_params(spec)
    (b, _), = oracle_mod.read_mode10(cases.FIXTURE_GZ, 100000, 211)
    o = oracle_mod.Oracle(params)
    o.calculate_buffer(b, oracle_mod.OracleRng(3))
    d1, d2 = tmp_path / "a", tmp_path / "b"
    d1.mkdir()
    d2.mkdir()
    hbt.write_outputs(params, oracle_result_arrays(o), str(d1), eco_output=1)
    o.write_outputs(str(d2), 1)
    for fn in sorted(os.listdir(d2)):
        assert (d1 / fn).read_bytes() == (d2 / fn).read_bytes()
        assert len((d1 / fn).read_text().splitlines()[0].split()) == 2


@pytest.mark.parametrize("name", cases.CASE_NAMES)
def test_reader_equals_reference_buffering(oracle_mod, name, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    par = spec["params"]
    paths = cases.write_inputs(spec, dig, str(tmp_path))
    ebs = int(par["event_buffer_size"])
    mine = hbt.read_buffers(paths["same"], ebs, 211)
    theirs = oracle_mod.read_mode10(paths["same"], ebs, 211)
    assert [b.nev for b, _ in mine] == list(ref["nev_per_buffer"])
    assert len(mine) == len(theirs)
    for (a, ta), (b, tb) in zip(mine, theirs):
        assert ta == tb
        assert np.array_equal(a.ev_off, b.ev_off)
        for x, y in zip(a.soa(), b.soa()):
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_reader_unit_test_counts():
    """particleSamples_unittest.cc:199-206: 619 and 635 pi+."""
    (b, total), = hbt.read_buffers(cases.FIXTURE_GZ, 100000, 211)
    assert list(np.diff(b.ev_off)) == [619, 635] and total == 1524


def test_reader_species_class_needs_table(tmp_path):
    with pytest.raises(hbt.HbtError):
        hbt.read_buffers(cases.FIXTURE_GZ, 100000, 9999)


SPECIES_MIX = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0), (111, 0.13498, 0.25, 1.0),
               (321, 0.49368, 0.3, 1.0), (-321, 0.49368, 0.3, 1.0), (2212, 0.93827, 0.35, 0.6),
               (-2212, 0.93827, 0.35, 0.4), (3122, 1.11568, 0.35, 0.3),
               (-3122, 1.11568, 0.35, 0.2), (7777777, 1.5, 0.3, 0.1))


@pytest.mark.parametrize("poi", [9999, 9998, -9998, 9997, -9997, 9996, -9996])
def test_reader_species_classes(poi, tmp_path):
    """decide_to_pick_OSCAR (particleSamples.cpp:625-678) with the species
    table: the kept particles are those of the class, in file order, with the
    reference's buffering (the budget counts all species)."""
    from hadronic_afterburner_toolkit_amd import samples
    table = os.path.join(cases.HERE, "species_table.dat")
    sp = hbt.read_species_table(table)
    q = {int(m): (int(c), int(b), int(st))
         for m, c, b, st in zip(sp["monval"], sp["charge"], sp["baryon"], sp["strange"])}
    pick = {9999: lambda c, b, st: c != 0, 9998: lambda c, b, st: c > 0,
            -9998: lambda c, b, st: c < 0, 9997: lambda c, b, st: b > 0,
            -9997: lambda c, b, st: b < 0, 9996: lambda c, b, st: st > 0,
            -9996: lambda c, b, st: st < 0}[poi]
    full = samples.generate(2801, 9, 90.0, species=SPECIES_MIX)
    path = str(tmp_path / "particle_samples.gz")
    samples.write_mode10(path, full)
    got = hbt.read_buffers(path, 400, poi, species_table=table)
    keep = np.array([pick(*q[int(m)]) for m in full.monval])
    e0 = 0
    for buf, total in got:
        nev = buf.nev
        if nev == 0:
            continue
        a, b = int(full.ev_off[e0]), int(full.ev_off[e0 + nev])
        assert total == b - a
        sel = np.flatnonzero(keep[a:b]) + a
        assert np.array_equal(buf.px, full.px[sel]) and np.array_equal(buf.E, full.E[sel])
        assert np.array_equal(buf.monval, full.monval[sel])
        assert np.array_equal(buf.mass, full.mass[sel])
        cnt = [int(keep[int(full.ev_off[e]):int(full.ev_off[e + 1])].sum())
               for e in range(e0, e0 + nev)]
        assert list(np.diff(buf.ev_off)) == cnt
        e0 += nev
    assert e0 == full.nev and keep.any() and not keep.all()


def test_reader_species_missing_from_table(tmp_path):
    """A monval the table does not hold fails the read, as the reference's
    lookup throws; the reader names it."""
    from hadronic_afterburner_toolkit_amd import samples
    full = samples.generate(2802, 3, 40.0, species=SPECIES_MIX[:2] + ((4444, 1.0, 0.3, 1.0),))
    path = str(tmp_path / "particle_samples.gz")
    samples.write_mode10(path, full)
    with pytest.raises(hbt.HbtError, match="4444"):
        hbt.read_buffers(path, 10 ** 6, 9998,
                         species_table=os.path.join(cases.HERE, "species_table.dat"))


def test_psi_ref_golden():
    z = np.load(os.path.join(cases.HERE, "psi_urqmd", "particles.npz"))
    golden = (-1.6751628499713109, -0.85236384292521539, 0.06674083818478263)
    for n, g in zip((1, 2, 3), golden):
        v = hbt.flow_event_plane_angle(z["px"], z["py"], n)
        assert abs(v - g) < 1e-8 and v == z["psi_ref"][n - 1]


def _number_strings(rng, n):
    """Decimal spellings a sample file may hold: %g / %e / %f at every
    precision, signs, leading and trailing zeros, missing integer or fraction
    digits, long digit strings, and exponents beyond the fast path."""
    out = []
    mags = 10.0 ** rng.uniform(-30, 30, n)
    vals = mags * rng.choice([-1.0, 1.0], n)
    for k, v in enumerate(vals):
        f = k % 9
        if f == 0:
            out.append("%.*g" % (1 + k % 17, v))
        elif f == 1:
            out.append("%.*e" % (k % 16, v))
        elif f == 2:
            out.append("%.*f" % (k % 12, v / 1e25))
        elif f == 3:
            out.append("%.*E" % (k % 10, v))
        elif f == 4:
            out.append("+%.7e" % abs(v))
        elif f == 5:
            out.append("%.17g" % v)
        elif f == 6:
            out.append("000" + ("%.9f" % (abs(v) / 1e25)))
        elif f == 7:
            out.append(("%.6f" % (abs(v) / 1e26)).lstrip("0") or "0")
        else:
            out.append("%d." % (k % 1000))
    out += ["0", "-0", "0.000", "42", "5.", ".5", "-.25", "1e-400", "1e400", "-1e400",
            "0.12345678901234567890123", "123456789012345678901234567890",
            "9007199254740993", "9007199254740992", "1e22", "1e23", "1e-22", "1e-23",
            "4.9406564584124654e-324", "2.2250738585072011e-308", "inf", "-inf"]
    return out


@pytest.mark.parametrize("gz", [False, True])
def test_reader_numbers_bit_identical_to_strtod(tmp_path, gz):
    """Every field is read as strtod reads it: the reader's fast decimal path
    and its fallback against Python's float() (correctly rounded, as glibc
    strtod), on plain and gzip text, with tabs and repeated blanks."""
    import gzip
    rng = np.random.default_rng(7)
    nums = _number_strings(rng, 4000)
    nums += ["1.0"] * ((-len(nums)) % 9)
    rows = [nums[k:k + 9] for k in range(0, len(nums), 9)]
    text = [f"{len(rows)}\n"]
    for k, r in enumerate(rows):
        sep = "\t" if k % 3 == 0 else ("  " if k % 3 == 1 else " ")
        text.append(f"{211 + (k % 2)}{sep}" + sep.join(r) + "\n")
    path = tmp_path / ("s.gz" if gz else "s.txt")
    if gz:
        with gzip.open(path, "wt") as f:
            f.write("".join(text))
    else:
        path.write_text("".join(text))
    (b, total), = hbt.read_buffers(str(path), 10 ** 9, 99999)
    assert total == len(rows) and list(b.ev_off) == [0, len(rows)]
    want = np.array([[float(x) for x in r] for r in rows])  # mass t x y z E px py pz
    cols = {"px": 6, "py": 7, "pz": 8, "E": 5, "x": 2, "y": 3, "z": 4, "t": 1}
    for name, c in cols.items():
        got = getattr(b, name)
        assert np.array_equal(got.view(np.uint64), want[:, c].view(np.uint64)), name


def _write_mode9(path, events):
    with open(path, "wb") as f:
        for mon, vals in events:
            np.array([len(mon)], np.int32).tofile(f)
            rec = np.zeros(len(mon), dtype=[("m", "<i4"), ("v", "<f4", (9,))])
            rec["m"] = mon
            rec["v"] = vals
            rec.tofile(f)


@pytest.mark.parametrize("budget", [1, 250, 700, 10 ** 6])
def test_mode9_binary_reader(tmp_path, budget):
    """Mode 9 (particleSamples.cpp:1203-1245): int count, then per particle an
    int id and 9 floats mass t x y z E px py pz, widened to double; whole
    events while the running total (all species) is below the budget; the
    species filter keeps one id.  Events keep their own particles."""
    rng = np.random.default_rng(3)
    events = []
    for e in range(9):
        n = int(rng.integers(0, 200))
        mon = rng.choice([211, -211, 321], n).astype(np.int32)
        vals = rng.normal(0, 2, (n, 9)).astype(np.float32)
        events.append((mon, vals))
    path = tmp_path / "particle_list.bin"
    _write_mode9(path, events)
    got = hbt.read_buffers(str(path), budget, 211, read_in_mode=9)
    e = 0
    for b, total in got:
        first, tot = e, 0
        while tot < budget and e < len(events):
            tot += len(events[e][0])
            e += 1
        assert total == tot and b.nev == e - first
        keep = [ev[0] == 211 for ev in events[first:e]]
        assert list(np.diff(b.ev_off)) == [int(k.sum()) for k in keep]
        v = np.concatenate([ev[1][k] for ev, k in zip(events[first:e], keep)] or
                           [np.zeros((0, 9), np.float32)]).astype(np.float64)
        for name, c in {"px": 6, "py": 7, "pz": 8, "E": 5, "x": 2, "y": 3, "z": 4, "t": 1}.items():
            assert np.array_equal(getattr(b, name), v[:, c]), name
    # every event read once; a buffer that fills up exactly with the last
    # event is followed by one empty buffer (the end is noticed by a read
    # that finds nothing), which the loop above checks as a 0-event buffer
    assert e == len(events)
    if budget == 10 ** 6:
        assert [b.nev for b, _ in got] == [9]
