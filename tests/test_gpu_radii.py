"""Physics check of the whole chain on the GPU: a Gaussian source with no
momentum-position correlation (x, y, z ~ N(0, R), t = 0, no LCMS boost, so
q_long = q_z) has C(q) = <cos(q.dx / hbar c)> = exp(-R^2 |q|^2 / hbar c^2)
in every direction.  The engine's histograms, the writer's files, the event
average with the reference's event-spread error and the radius fit weighted
by it must return R_out = R_side = R_long = R and lambda = 1; the bar
(written here) is 3 % on the radii and 0.06 on lambda.  Physics check only:
statistical, not a parity test."""
import numpy as np
import pytest

from hadronic_afterburner_toolkit_amd import _lib, hbt, radii, samples

pytestmark = pytest.mark.gpu

pytest.importorskip("scipy")


def test_gaussian_source_radii(tmp_path):
    """Six groups of 20 events, each written by the reference's writer into
    its own folder, averaged over the groups with the reference's
    event-spread error, fitted with that error as sigma."""
    R = 4.0
    full = samples.generate(4242, 120, 600.0)
    rng = np.random.default_rng(7)
    full.x = rng.normal(0.0, R, full.n)
    full.y = rng.normal(0.0, R, full.n)
    full.z = rng.normal(0.0, R, full.n)
    full.t = np.zeros(full.n)
    params = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=2, KT_min=0.2, KT_max=0.7,
                  HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
                  number_of_mixed_events=10, long_comoving_boost=0)
    print(f"gaussian source: {full.nev} events, {full.n} particles, R = {R} fm", flush=True)
    folders = []
    for k in range(6):
        g = hbt.HBTCorrelation(params, precision=_lib.HBT_PREC_FP32C,
                               kernel=_lib.HBT_KERNEL_FAST)
        g.calculate_groups([full.events(20 * k, 20)], hbt.Random(9 + k))
        d = tmp_path / f"event_{k}"
        d.mkdir()
        hbt.write_outputs(params, vars(g.results()), str(d))
        g.close()
        folders.append(str(d))
    (name,) = radii.average_folders(folders, str(tmp_path / "avg"))
    table = np.loadtxt(tmp_path / "avg" / name)
    assert table.shape[1] == 7
    assert np.isfinite(table[:, 6]).any()
    rows = radii.fit_table(table, q_cuts=(0.1,))
    lam, r_out, r_side, r_long = rows[0][[1, 3, 5, 7]]
    print(f"lambda {lam:.4f}  R_out {r_out:.3f}  R_side {r_side:.3f}  R_long {r_long:.3f}")
    assert abs(lam - 1.0) < 0.06
    for r in (r_out, r_side, r_long):
        assert abs(r / R - 1.0) < 0.03
