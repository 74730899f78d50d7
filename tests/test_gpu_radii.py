"""Physics check of the whole chain on the GPU: a Gaussian source with no
momentum-position correlation (x, y, z ~ N(0, R), t = 0, no LCMS boost, so
q_long = q_z) has C(q) = <cos(q.dx / hbar c)> = exp(-R^2 |q|^2 / hbar c^2)
in every direction.  The engine's histograms, the writer's files and the
radius fit must return R_out = R_side = R_long = R and lambda = 1; the bar
(written here) is 3 % on the radii and 0.06 on lambda, where the fit's own
errors on this sample are ~1 % on the radii and ~0.012 on lambda (the CPU
oracle on the same input gives lambda 0.961, radii 4.02 / 4.01 / 4.05 fm at
|q| < 0.1 GeV)."""
import numpy as np
import pytest

from hadronic_afterburner_toolkit_amd import _lib, hbt, radii, samples

pytestmark = pytest.mark.gpu

pytest.importorskip("scipy")


def test_gaussian_source_radii(tmp_path):
    R = 4.0
    buf = samples.generate(4242, 50, 600.0)
    rng = np.random.default_rng(7)
    buf.x = rng.normal(0.0, R, buf.n)
    buf.y = rng.normal(0.0, R, buf.n)
    buf.z = rng.normal(0.0, R, buf.n)
    buf.t = np.zeros(buf.n)
    params = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=2, KT_min=0.2, KT_max=0.7,
                  HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
                  number_of_mixed_events=10, long_comoving_boost=0)
    print(f"gaussian source: {buf.nev} events, {buf.n} particles, R = {R} fm", flush=True)
    g = hbt.HBTCorrelation(params, precision=_lib.HBT_PREC_FP32C, kernel=_lib.HBT_KERNEL_FAST)
    g.calculate_groups([buf], hbt.Random(9))
    hbt.write_outputs(params, vars(g.results()), str(tmp_path))
    g.close()
    (path,) = [p for p in tmp_path.iterdir() if p.name.startswith("HBT_correlation_function_KT_")]
    rows = radii.fit_table(np.loadtxt(path), q_cuts=(0.1,))
    lam, r_out, r_side, r_long = rows[0][[1, 3, 5, 7]]
    print(f"lambda {lam:.4f}  R_out {r_out:.3f}  R_side {r_side:.3f}  R_long {r_long:.3f}")
    assert abs(lam - 1.0) < 0.06
    for r in (r_out, r_side, r_long):
        assert abs(r / R - 1.0) < 0.03
