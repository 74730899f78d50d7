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
    table = cases.bf_species_table(spec)
    bf = hbt.BalanceFunction(params, species=hbt.read_species_table(table) if table else None)
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


# --- the two per-pair forms (HBT_BF_KERNEL_FAST, HBT_BF_KERNEL_EXACT) ----------------

from hadronic_afterburner_toolkit_amd import _lib  # noqa: E402

PI_K = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0),
        (321, 0.49368, 0.30, 1.0), (-321, 0.49368, 0.30, 1.0))


def run_bf(params, buf, seed, kernel, mix=None):
    bf = hbt.BalanceFunction(params, kernel=kernel)
    bf.calculate_balance_function(buf, hbt.Random(seed), mix=mix)
    r = bf.results()
    r["info"] = bf.info()
    bf.close()
    return r


def run_oracle(params, buf, seed, mix=None):
    o = oracle.BfOracle(params)
    o.calculate_buffer(buf, oracle.OracleRng(seed), mix)
    return o.arrays()


def grid_sample(seed, nev, mult, y_step, phi_step, eta=False):
    """Particles whose rapidity (or pseudo-rapidity) and azimuth sit on grids,
    so that many pair differences fall on bin edges and on Delta y = 0."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(mult, nev).astype(np.int64)
    n = int(counts.sum())
    sp = rng.integers(0, 2, n)
    mon = np.where(sp == 0, 211, -211).astype(np.int32)
    mass = np.full(n, 0.13957)
    pt = rng.uniform(0.25, 2.0, n)
    y = rng.integers(-30, 31, n) * y_step
    phi = rng.integers(-20, 21, n) * phi_step
    px, py = pt * np.cos(phi), pt * np.sin(phi)
    mt = np.sqrt(pt * pt + mass * mass)
    pz = pt * np.sinh(y) if eta else mt * np.sinh(y)
    E = np.sqrt(mass * mass + pt * pt + pz * pz)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    z = np.zeros(n)
    return samples.Buffer(ev_off=off, px=px, py=py, pz=pz, E=E, x=z, y=z, z=z, t=z,
                          monval=mon, mass=mass)


@pytest.mark.parametrize("name", cases.BF_CASE_NAMES)
def test_bf_exact_kernel_equals_reference(name):
    spec, ref, outs, dig = cases.load(name)
    full = cases.bf_input(spec, dig)
    params = {k: spec["params"][k] for k in hbt.BF_KEYS}
    print(f"balance function case {name}, exact form", flush=True)
    table = cases.bf_species_table(spec)
    bf = hbt.BalanceFunction(params, kernel=_lib.HBT_BF_KERNEL_EXACT,
                             species=hbt.read_species_table(table) if table else None)
    rng = hbt.Random(int(spec["params"]["randomSeed"]))
    for first, count in samples.buffers_by_budget(full, int(spec["params"]["event_buffer_size"])):
        bf.calculate_balance_function(full.events(first, count), rng)
        assert bf.info()["kernel"] == _lib.HBT_BF_KERNEL_EXACT
    r = bf.results()
    bf.close()
    assert np.array_equal(r["hist"], ref["hist"])


@pytest.mark.parametrize("bnpts,rap_type,alpha,beta", [(21, 1, 211, 211), (20, 1, 211, -211),
                                                       (21, 0, 321, -211), (300, 1, 211, 211),
                                                       (817, 0, 211, -211)])
def test_bf_fast_equals_oracle(bnpts, rap_type, alpha, beta):
    """Fast form (folded LDS layout for Bnpts <= 242, mod-20 layout above)
    against the oracle, odd and even Bnpts, rapidity and pseudo-rapidity."""
    buf = samples.generate(2201 + bnpts, 40, 300.0, species=PI_K)
    params = dict(particle_alpha=alpha, particle_beta=beta, BpT_min=0.2, BpT_max=3.0,
                  Brap_max=2.0, Bnpts=bnpts, rap_type=rap_type)
    print(f"fast form vs oracle Bnpts={bnpts} rap_type={rap_type}", flush=True)
    r = run_bf(params, buf, 17, _lib.HBT_BF_KERNEL_FAST)
    a = run_oracle(params, buf, 17)
    assert r["info"]["kernel"] == _lib.HBT_BF_KERNEL_FAST
    assert a["hist"].sum() > 0
    assert np.array_equal(r["hist"], a["hist"])
    assert (r["N_b"], r["N_bbar"]) == (a["N_b"], a["N_bbar"])


@pytest.mark.parametrize("bnpts,eta", [(21, False), (20, False), (21, True), (41, False)])
def test_bf_fast_bin_edges(bnpts, eta):
    """Rapidities on a 0.05 grid and azimuths on a pi/40 grid put most pair
    differences on or next to bin edges and many at Delta y = 0: the guard
    band must send those pairs to the FP64 arithmetic."""
    buf = grid_sample(2301 + bnpts, 30, 120.0, 0.05, np.pi / 40, eta=eta)
    params = dict(particle_alpha=211, particle_beta=211, BpT_min=0.2, BpT_max=3.0,
                  Brap_max=2.0, Bnpts=bnpts, rap_type=0 if eta else 1)
    print(f"bin-edge grid Bnpts={bnpts} eta={eta}", flush=True)
    r = run_bf(params, buf, 23, _lib.HBT_BF_KERNEL_FAST)
    a = run_oracle(params, buf, 23)
    assert r["info"]["kernel"] == _lib.HBT_BF_KERNEL_FAST
    assert r["info"]["fallback"] > 0
    assert np.array_equal(r["hist"], a["hist"])


def test_bf_zero_pt_and_infinite_eta():
    """BpT_min = 0 admits particles with pT = 0, whose pseudo-rapidity is
    +-inf or NaN: their pairs are skipped as on the reference's x86 build."""
    buf = samples.generate(2401, 12, 80.0, species=PI_K[:2])
    sel = np.arange(0, buf.n, 7)
    buf.px[sel] = 0.0
    buf.py[sel] = 0.0
    buf.pz[sel[::2]] = 0.0
    params = dict(particle_alpha=211, particle_beta=-211, BpT_min=0.0, BpT_max=3.0,
                  Brap_max=2.0, Bnpts=21, rap_type=0)
    for kern in (_lib.HBT_BF_KERNEL_FAST, _lib.HBT_BF_KERNEL_EXACT):
        r = run_bf(params, buf, 29, kern)
        a = run_oracle(params, buf, 29)
        assert np.array_equal(r["hist"], a["hist"]), kern


def test_bf_fast_equals_exact_large():
    """At a production size (600 events, Poisson(900) over pi+-, K+-) the
    fast form's histograms equal the exact form's bit for bit; a separate
    mixed buffer as well."""
    buf = samples.generate(2501, 600, 900.0, species=PI_K)
    mix = samples.generate(2502, 200, 900.0, species=PI_K)
    params = dict(particle_alpha=211, particle_beta=211, BpT_min=0.2, BpT_max=3.0,
                  Brap_max=2.0, Bnpts=21, rap_type=1)
    print("fast vs exact, 600 events Poisson(900)", flush=True)
    f = run_bf(params, buf, 31, _lib.HBT_BF_KERNEL_FAST)
    e = run_bf(params, buf, 31, _lib.HBT_BF_KERNEL_EXACT)
    assert f["info"]["pairs"] == e["info"]["pairs"] > 1e8
    assert np.array_equal(f["hist"], e["hist"])
    f = run_bf(params, buf, 37, _lib.HBT_BF_KERNEL_FAST, mix=mix)
    e = run_bf(params, buf, 37, _lib.HBT_BF_KERNEL_EXACT, mix=mix)
    assert np.array_equal(f["hist"], e["hist"])


def test_bf_errors_and_empty_buffer():
    """Refused configurations (hbt_bf_create) and calls, and an empty buffer."""
    base = dict(particle_alpha=211, particle_beta=-211, BpT_min=0.2, BpT_max=3.0,
                Brap_max=2.0, Bnpts=21, rap_type=1)
    for bad in (dict(Bnpts=818), dict(Bnpts=1), dict(rap_type=2)):
        with pytest.raises(hbt.HbtError) as ei:
            hbt.BalanceFunction(dict(base, **bad))
        assert ei.value.code == -1, bad  # HBT_ERR_ARG
    with pytest.raises(hbt.HbtError):
        hbt.BalanceFunction(base, kernel=7)
    bf = hbt.BalanceFunction(base)
    empty = samples.generate(2601, 0, 100.0, species=PI_K)
    bf.calculate_balance_function(empty, hbt.Random(5))
    r = bf.results()
    assert r["hist"].sum() == 0 and (r["N_b"], r["N_bbar"]) == (0, 0)
    buf = samples.generate(2602, 3, 50.0, species=PI_K)
    with pytest.raises(hbt.HbtError):  # a separate mixed buffer without events
        bf.calculate_balance_function(buf, hbt.Random(5), mix=empty)
    bf.close()


def test_bf_species_class_errors():
    """A species class without a species table, and a particle whose monval
    is not in the table (the reference's lookup throws), are errors."""
    params = dict(particle_alpha=9998, particle_beta=-9998, BpT_min=0.2, BpT_max=3.0,
                  Brap_max=2.0, Bnpts=21, rap_type=1)
    buf = samples.generate(2701, 4, 60.0, species=PI_K)
    bf = hbt.BalanceFunction(params)
    with pytest.raises(hbt.HbtError):
        bf.calculate_balance_function(buf, hbt.Random(3))
    table = hbt.read_species_table(os.path.join(cases.HERE, "species_table.dat"))
    keep = table["monval"] != 321
    bf.set_species({k: v[keep] for k, v in table.items()})
    with pytest.raises(hbt.HbtError) as ei:
        bf.calculate_balance_function(buf, hbt.Random(3))
    assert "321" in str(ei.value)
    bf.set_species(table)
    bf.calculate_balance_function(buf, hbt.Random(3))
    r = bf.results()
    bf.close()
    o = oracle.BfOracle(params, species=table)
    o.calculate_buffer(buf, oracle.OracleRng(3))
    assert np.array_equal(r["hist"], o.arrays()["hist"])
