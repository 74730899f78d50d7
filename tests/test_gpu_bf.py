"""The balance function on the GPU (hbt_bf_*) against the reference's own
build (golden histograms and files, tests/golden/make_golden_bf.py) and the
CPU oracle: the eight histograms are integer counts and must be equal bit for
bit; N_b / N_bbar equal; the written files byte-identical."""
import hashlib
import os

import numpy as np
import pytest

import cases
import oracle
from hadronic_afterburner_toolkit_amd import hbt, samples

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", cases.BF_CASE_NAMES)
def test_bf_gpu_equals_reference(name, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    full = cases.bf_input(spec, dig)
    params = {k: spec["params"][k] for k in hbt.BF_KEYS}
    print(f"balance function case {name}", flush=True)
    bf = hbt.BalanceFunction(params)
    rng = hbt.Random(int(spec["params"]["randomSeed"]))
    for first, count in samples.buffers_by_budget(full, int(spec["params"]["event_buffer_size"])):
        bf.calculate_balance_function(full.events(first, count), rng)
    r = bf.results()
    assert np.array_equal(r["hist"], ref["hist"])
    assert r["N_b"] == int(ref["N_b"]) and r["N_bbar"] == int(ref["N_bbar"])
    bf.output_balance_function(str(tmp_path))
    bf.close()
    for fn, meta in outs.items():
        with open(tmp_path / fn, "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == meta["sha256"], fn


def test_bf_gpu_equals_oracle_separate_mixed_file():
    """A separate mixed-event buffer (read_in_real_mixed_events = 1) and
    uneven events, against the oracle."""
    species = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0),
               (321, 0.49368, 0.30, 1.0), (-321, 0.49368, 0.30, 1.0))
    same = samples.generate(2101, 9, samples.log_uniform_mult(5, 400), species=species)
    mix = samples.generate(2102, 7, 150.0, species=species)
    params = dict(particle_alpha=321, particle_beta=-211, BpT_min=0.15, BpT_max=2.5,
                  Brap_max=1.8, Bnpts=25, rap_type=1)
    bf = hbt.BalanceFunction(params)
    bf.calculate_balance_function(same, hbt.Random(3), mix=mix)
    r = bf.results()
    bf.close()
    o = oracle.BfOracle(params)
    o.calculate_buffer(same, oracle.OracleRng(3), mix)
    a = o.arrays()
    assert a["hist"].sum() > 0
    assert np.array_equal(r["hist"], a["hist"])
    assert (r["N_b"], r["N_bbar"]) == (a["N_b"], a["N_bbar"])
