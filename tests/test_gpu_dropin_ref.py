"""The drop-in in the reference's own program: oracle/_ref/dropin/
hadronic_afterburner_tools.e is the reference's main.cpp, Analysis.cpp,
particleSamples, ParameterReader and Random compiled from its sources, with
its HBT_correlation replaced by this repository's class (built by
`make -C oracle dropin` where the reference tree is present; the binary and
the data files it reads travel with the tree).  It runs
Analysis::HBTAnalysis (Analysis.cpp:817-835) unchanged: the reference's
reader fills the reference's particleSamples, and the class reads it through
the reference's getters and runs the pairs on the GPU.

Bar (as for the drop-in CLI): identical file names and line counts against
the oracle's files for the same inputs (the oracle's files are byte-identical
to the reference's), every number within 2e-8 relative (one unit in the 8th
printed digit; the GPU adds in another order) and at least 99.9% of lines
byte-identical.  The program reads the reference's parameters.dat and
EOS/pdg.dat, copied beside it by the same make target.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import cases
import oracle

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(REPO, "oracle", "_ref", "dropin")
EXE = os.path.join(DROPIN, "hadronic_afterburner_tools.e")
BASE = dict(read_in_mode=10, particle_monval=211, analyze_flow=0, analyze_HBT=1,
            analyze_balance_function=0, analyze_ebe_yield=0, resonance_feed_down_flag=0,
            resonance_weak_feed_down_flag=0, read_in_real_mixed_events=0, echo_level=0)


@pytest.mark.skipif(not os.path.exists(EXE), reason="the reference's program with the "
                    "drop-in class is built only where the reference tree is present")
@pytest.mark.parametrize("name", ["fixture_3d", "fixture_kphi_qinv", "synth_cfg0",
                                  "synth_mixedfile", "edge_3d"])
def test_reference_program_with_dropin_class(name, tmp_path):
    spec, ref, outs, dig = cases.load(name)
    par = dict(spec["params"])
    work = tmp_path / "work"
    res = work / "results"
    res.mkdir(parents=True)
    (work / "EOS").mkdir()
    shutil.copy(os.path.join(DROPIN, "EOS", "pdg.dat"), work / "EOS" / "pdg.dat")
    shutil.copy(os.path.join(DROPIN, "parameters.dat"), work / "parameters.dat")
    paths = cases.write_inputs(spec, dig, str(res))
    keys = dict(BASE)
    keys.update(par)
    keys.update(hbt_kernel=1, hbt_precision=0)
    print(f"reference program + drop-in class: case {name}", flush=True)
    r = subprocess.run([EXE] + [f"{k}={v}" for k, v in keys.items()], cwd=str(work),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ebs = int(par["event_buffer_size"]) if "event_buffer_size" in par else 100000
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
        for la, lb in zip(a, b):
            total += 1
            if la == lb:
                same += 1
                continue
            va = np.array([float(x) for x in la.split()])
            vb = np.array([float(x) for x in lb.split()])
            ok = np.isclose(va, vb, rtol=2e-8, atol=0) | (np.isnan(va) & np.isnan(vb))
            assert ok.all(), (fn, la, lb)
    assert same >= 0.999 * total, (same, total)
