"""The CPU oracle's balance function against the reference's own build
(golden vectors from oracle/_ref/ref_bf_harness, tests/golden/make_golden_bf.py):
the eight pair histograms and the b / bbar counts equal, bit for bit, and
the three output files byte-identical (sha256)."""
import hashlib
import os

import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import samples


def bf_params(spec):
    return {k: spec["params"][k] for k in oracle.BF_KEYS}


def run_oracle(name):
    spec, ref, outs, dig = cases.load(name)
    full = cases.bf_input(spec, dig)
    table = cases.bf_species_table(spec)
    o = oracle.BfOracle(bf_params(spec), species=oracle.species_read(table) if table else None)
    rng = oracle.OracleRng(int(spec["params"]["randomSeed"]))
    bufs = samples.buffers_by_budget(full, int(spec["params"]["event_buffer_size"]))
    assert [n for _, n in bufs] == list(ref["nev_per_buffer"])
    for first, count in bufs:
        o.calculate_buffer(full.events(first, count), rng)
    return o, spec, ref, outs


@pytest.mark.parametrize("name", cases.BF_CASE_NAMES)
def test_bf_oracle_equals_reference(name, tmp_path):
    o, spec, ref, outs = run_oracle(name)
    a = o.arrays()
    assert np.array_equal(a["hist"], ref["hist"])
    assert a["N_b"] == int(ref["N_b"]) and a["N_bbar"] == int(ref["N_bbar"])
    assert o.write_outputs(str(tmp_path)) == 0
    assert sorted(os.listdir(tmp_path)) == sorted(outs)
    for fn, meta in outs.items():
        with open(tmp_path / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == meta["sha256"], fn


@pytest.mark.parametrize("name", cases.BF_CASE_NAMES)
def test_bf_engine_writer_byte_identical(name, tmp_path):
    """The engine library's writer (host code) on the reference's histograms
    writes the reference's three files byte for byte."""
    from hadronic_afterburner_toolkit_amd import hbt
    spec, ref, outs, _ = cases.load(name)
    hbt.write_bf_outputs(bf_params(spec), ref["hist"], str(tmp_path))
    assert sorted(os.listdir(tmp_path)) == sorted(outs)
    for fn, meta in outs.items():
        with open(tmp_path / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == meta["sha256"], fn
