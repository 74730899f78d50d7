"""Golden vectors of the balance-function pair histograms from the
reference's own build (oracle/_ref/ref_bf_harness: BalanceFunction and
particleSamples compiled from their sources), run in the build container.

    python tests/golden/make_golden_bf.py        (needs /root/reference)

Per case, under tests/golden/<name>/:
  ref.npz        the eight histograms [8][Bnpts][20] (C_ab, C_abarbbar,
                 C_abbar, C_abarb, then the mixed-event four), N_b, N_bbar,
                 events per buffer
  outputs.json   sha256 and first lines of every file the reference wrote
  params.json    the input description and the parameter keys
  input.sha256   sha256 of the decompressed input
The inputs are regenerated from their seeds by the tests (samples.generate).
"""
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
sys.path.insert(0, HERE)
import make_golden as mg  # noqa: E402

from hadronic_afterburner_toolkit_amd import samples  # noqa: E402

HARNESS = os.path.join(mg.REPO, "oracle", "_ref", "ref_bf_harness")

PIKPM = ((211, 0.13957, 0.25, 3.0), (-211, 0.13957, 0.25, 3.0),
         (321, 0.49368, 0.30, 1.0), (-321, 0.49368, 0.30, 1.0))

CASES = {
    # same species (alpha = beta): the a and b lists coincide
    "bf_pipi": dict(input=(2001, 8, 160.0), params=dict(
        randomSeed=41, particle_alpha=211, particle_beta=211, BpT_min=0.2, BpT_max=3.0,
        Bnpts=21, Brap_max=2.0, rap_type=1, event_buffer_size=100000)),
    # two species, pseudo-rapidity, several buffers (histograms and the random
    # stream carry over from buffer to buffer)
    "bf_kpi_eta": dict(input=(2002, 12, 120.0), params=dict(
        randomSeed=43, particle_alpha=321, particle_beta=211, BpT_min=0.1, BpT_max=2.0,
        Bnpts=15, Brap_max=1.5, rap_type=0, event_buffer_size=500)),
    # a particle and its antiparticle (alpha = -beta)
    "bf_pi_antipi": dict(input=(2003, 6, 200.0), params=dict(
        randomSeed=47, particle_alpha=211, particle_beta=-211, BpT_min=0.3, BpT_max=1.5,
        Bnpts=31, Brap_max=1.0, rap_type=1, event_buffer_size=100000)),
}

# Species classes need the species table: these cases run the reference with
# tests/golden/species_table.dat as its EOS/pdg.dat (particle_decay.cpp:28).
SPECIES_TABLE = os.path.join(HERE, "species_table.dat")
MIXED = PIKPM + ((111, 0.13498, 0.25, 1.5), (311, 0.49761, 0.30, 0.5),
                 (-311, 0.49761, 0.30, 0.5), (2212, 0.93827, 0.35, 0.8),
                 (-2212, 0.93827, 0.35, 0.6), (2112, 0.93957, 0.35, 0.5),
                 (3122, 1.11568, 0.35, 0.4), (-3122, 1.11568, 0.35, 0.3),
                 (7777777, 1.5, 0.3, 0.2))
CLASS_CASES = {
    # the reference's parameters.dat choice: positive against negative hadrons
    "bf_charged": dict(input=(2011, 8, 150.0), species=MIXED, params=dict(
        randomSeed=53, particle_alpha=9998, particle_beta=-9998, BpT_min=0.2, BpT_max=3.0,
        Bnpts=21, Brap_max=2.0, rap_type=1, event_buffer_size=100000)),
    # baryons against negative strangeness, pseudo-rapidity, several buffers
    "bf_baryon_strange": dict(input=(2012, 10, 120.0), species=MIXED, params=dict(
        randomSeed=59, particle_alpha=9997, particle_beta=-9996, BpT_min=0.1, BpT_max=2.5,
        Bnpts=16, Brap_max=1.6, rap_type=0, event_buffer_size=600)),
    # all charged against protons: the antiparticle list of 9999 is that of
    # -9999, an identified id no particle carries
    "bf_allcharged_p": dict(input=(2013, 6, 180.0), species=MIXED, params=dict(
        randomSeed=61, particle_alpha=9999, particle_beta=2212, BpT_min=0.2, BpT_max=3.0,
        Bnpts=25, Brap_max=2.0, rap_type=1, event_buffer_size=100000)),
}


def read_dump(path):
    with open(path, "rb") as f:
        raw = f.read()
    magic, npts, nphi, nbuf = [int(v) for v in np.frombuffer(raw[:16], dtype=np.int32)]
    assert magic == 0x42465231
    nb, nbb = [int(v) for v in np.frombuffer(raw[16:32], dtype=np.int64)]
    off = 32
    n = 8 * npts * nphi
    hist = np.frombuffer(raw[off:off + 8 * n], dtype=np.float64).reshape(8, npts, nphi).copy()
    off += 8 * n
    nev = np.frombuffer(raw[off:off + 4 * nbuf], dtype=np.int32).copy()
    return dict(hist=hist, N_b=np.int64(nb), N_bbar=np.int64(nbb), nev_per_buffer=nev)


def run_case(name, spec):
    work = tempfile.mkdtemp(prefix=f"golden_{name}_")
    res = os.path.join(work, "results")
    os.makedirs(res)
    species = spec.get("species", PIKPM)
    if "species" in spec:
        os.makedirs(os.path.join(work, "EOS"))
        shutil.copy(SPECIES_TABLE, os.path.join(work, "EOS", "pdg.dat"))
    else:
        os.symlink(os.path.join(mg.REF, "EOS"), os.path.join(work, "EOS"))
    shutil.copy(os.path.join(mg.REF, "parameters.dat"), os.path.join(work, "parameters.dat"))
    seed, nev, mult = spec["input"]
    samples.write_mode10(os.path.join(res, "particle_samples.gz"),
                         samples.generate(seed, nev, mult, species=species))
    with gzip.open(os.path.join(res, "particle_samples.gz"), "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    params = dict(mg.BASE)
    params.update(analyze_HBT=0, analyze_balance_function=1)
    params.update(spec["params"])
    dump = os.path.join(work, "dump.bin")
    r = subprocess.run([HARNESS, work, dump] + [f"{k}={v}" for k, v in params.items()],
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{name}: ref_bf_harness failed: {r.stderr[-2000:]}")
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
            outs[fn] = dict(sha256=mg.sha(p), head=head, size=os.path.getsize(p))
    with open(os.path.join(outdir, "outputs.json"), "w") as f:
        json.dump(outs, f, indent=1, sort_keys=True)
    with open(os.path.join(outdir, "params.json"), "w") as f:
        meta = dict(input=list(spec["input"]), species=[list(s) for s in species],
                    params=params)
        if "species" in spec:
            meta["species_table"] = os.path.basename(SPECIES_TABLE)
        json.dump(meta, f, indent=1)
    with open(os.path.join(outdir, "input.sha256"), "w") as f:
        json.dump({"particle_samples.gz": digest}, f, indent=1)
    shutil.rmtree(work)
    print(name, "buffers", list(d["nev_per_buffer"]), "pairs", d["hist"].sum(axis=(1, 2)))


if __name__ == "__main__":
    only = sys.argv[1:]
    for name, spec in {**CASES, **CLASS_CASES}.items():
        if not only or name in only:
            run_case(name, spec)
