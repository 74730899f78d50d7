"""The CPU oracle against the reference's own results (CPU only).

Golden dumps were written by the reference's HBT_correlation compiled from
its sources (tests/golden/make_golden.py).  The oracle must reproduce every
count and every FP64 sum bit for bit, and the reference's text files byte
for byte.  Also pins psi_ref (HBT_unittest.cc:39,43,47) and the reader's
species counts (particleSamples_unittest.cc:199-206).
"""
import hashlib
import os

import numpy as np
import pytest

import cases

FIELDS = ("num_count", "den", "q_out_sum", "q_side_sum", "q_long_sum", "num", "cnt_num_kt",
          "cnt_den_kt", "cnt_num_qinv", "cnt_den_qinv")


COUNTS = ("num_count", "den", "qinv_count", "qinv_den")


def bits(a, key):
    """Counts compare as integers, sums by their exact bit patterns."""
    a = np.asarray(a)
    if key in COUNTS or a.dtype != np.float64:
        return a.astype(np.uint64)
    return a.view(np.uint64)


def run_oracle_case(oracle_mod, name, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    par = spec["params"]
    paths = cases.write_inputs(spec, dig, str(tmp_path))
    ebs = int(par["event_buffer_size"])
    bufs = oracle_mod.read_mode10(paths["same"], ebs, 211)
    mixb = oracle_mod.read_mode10(paths["mixed"], ebs, 211) if "mixed" in paths else None
    o = oracle_mod.Oracle(cases.hbt_params(spec))
    rng = oracle_mod.OracleRng(int(par["randomSeed"]))
    for k, (b, _) in enumerate(bufs):
        o.calculate_buffer(b, rng, mixb[k][0] if mixb else None)
    return spec, cases.ref_arrays_as_u64(ref), outs, o, bufs


@pytest.mark.parametrize("name", cases.CASE_NAMES)
def test_oracle_bit_exact_with_reference(oracle_mod, name, tmp_path):
    spec, ref, outs, o, bufs = run_oracle_case(oracle_mod, name, tmp_path)
    a = o.arrays()
    assert [b.nev for b, _ in bufs] == list(ref["nev_per_buffer"])
    for key in FIELDS:
        assert np.array_equal(bits(ref[key], key), bits(a[key], key)), key
    if "qinv_count" in ref:
        for key in ("qinv_count", "qinv_sum", "qinv_num", "qinv_den"):
            assert np.array_equal(bits(ref[key], key), bits(a[key], key)), key
    if "cnt_num_ktphi" in ref:
        for key in ("cnt_num_ktphi", "cnt_den_ktphi"):
            assert np.array_equal(ref[key], a[key]), key
    # the writers: byte-identical files (sha256 of the reference's output)
    outdir = tmp_path / "out"
    outdir.mkdir()
    assert o.write_outputs(str(outdir), 0) == 0
    got = sorted(os.listdir(outdir))
    assert got == sorted(outs)
    for fn in got:
        with open(outdir / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == outs[fn]["sha256"], fn


def test_cap_binds_as_reference(oracle_mod, tmp_path):
    """needed_number_of_pairs=40: N+1 pairs admitted per K_T bin, counters
    persist across buffers (HBT_correlation.cpp:402-406)."""
    spec, ref, outs, o, bufs = run_oracle_case(oracle_mod, "synth_cap", tmp_path)
    a = o.arrays()
    n_written = int(spec["params"]["n_KT"]) - 1
    assert list(a["cnt_num_kt"][:n_written]) == [41] * n_written
    assert list(a["cnt_den_kt"][:n_written]) == [41] * n_written


def test_psi_ref_golden(oracle_mod):
    """HBT_unittest.cc:23-48: psi_ref for n=1,2,3, tolerance 1e-8, on the
    pi+ the reference read from its UrQMD-zipped fixture
    (unit_tests/parameters.dat: read_in_mode=2; particles stored by
    make_golden.py).  All particles of the buffer, no rapidity cut."""
    from hadronic_afterburner_toolkit_amd.samples import Buffer
    z = np.load(os.path.join(cases.HERE, "psi_urqmd", "particles.npz"))
    zeros = np.zeros_like(z["px"])
    buf = Buffer(ev_off=z["ev_off"], px=z["px"], py=z["py"], pz=zeros, E=zeros + 1, x=zeros,
                 y=zeros, z=zeros, t=zeros)
    golden = (-1.6751628499713109, -0.85236384292521539, 0.06674083818478263)
    for n, g in zip((1, 2, 3), golden):
        assert abs(oracle_mod.psi_ref(buf, n) - g) < 1e-8
        # and bit-exact with the reference build's own value on the same input
        assert oracle_mod.psi_ref(buf, n) == z["psi_ref"][n - 1]


def test_psi_ref_matches_reference_run():
    """psi_ref (n=2) of the azimuthal golden case, as the reference computed it."""
    import oracle
    (buf, _), = oracle.read_mode10(cases.FIXTURE_GZ, 100000, 211)
    spec, ref, _, _ = cases.load("fixture_kphi_qinv")
    assert oracle.psi_ref(buf, 2) == float(ref["psi_ref"])


def test_reader_counts_golden(oracle_mod):
    """particleSamples_unittest.cc:199-206: 2 events, 619 and 635 pi+."""
    (buf, total), = oracle_mod.read_mode10(cases.FIXTURE_GZ, 100000, 211)
    assert buf.nev == 2
    assert list(np.diff(buf.ev_off)) == [619, 635]
    assert total == 744 + 780


def test_reader_buffering(oracle_mod, tmp_path):
    """Whole events are read while the running total of all species is below
    event_buffer_size (particleSamples.cpp:1256): budget 700 takes event 0
    (744 particles) alone, then event 1; the end of file is only seen by the
    next read, so a third, empty buffer follows (as in the golden case
    fixture_budget700, which the reference produced)."""
    bufs = oracle_mod.read_mode10(cases.FIXTURE_GZ, 700, 211)
    assert [b.nev for b, _ in bufs] == [1, 1, 0]
    assert [t for _, t in bufs] == [744, 780, 0]
