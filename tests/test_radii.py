"""Event averaging and the 3D Gaussian radius fit (host; the step after the
pair path, ebe_scripts/average_event_HBT_correlation_function.py and
ebe_scripts/fit_HBT_radii.py), pinned by what fixes them: the average is the
column-wise mean; the fit returns the parameters a noise-free correlation
function was built from; the correlation of a writer file is num / den_norm
of its bins."""
import os

import numpy as np
import pytest

from hadronic_afterburner_toolkit_amd import radii

pytest.importorskip("scipy")


def grid(nq=41, qmax=0.15):
    q = np.linspace(-qmax, qmax, nq)
    ql, qo, qs = np.meshgrid(q, q, q, indexing="ij")  # [q_long][q_out][q_side]
    return qo, qs, ql


def test_average_is_columnwise_mean(tmp_path):
    rng = np.random.default_rng(1)
    tabs = [rng.normal(size=(27, 5)) for _ in range(4)]
    paths = []
    for k, t in enumerate(tabs):
        p = tmp_path / f"ev{k}" / "HBT_correlation_function_KT_0.2_0.3.dat"
        p.parent.mkdir()
        np.savetxt(p, t, fmt="%.17e")
        paths.append(str(p))
    assert np.array_equal(radii.average_files(paths), sum(tabs) / 4)
    names = radii.average_folders([os.path.dirname(p) for p in paths], str(tmp_path / "avg"))
    assert names == ["HBT_correlation_function_KT_0.2_0.3.dat"]
    back = np.loadtxt(tmp_path / "avg" / names[0])
    assert np.allclose(back, sum(tabs) / 4, rtol=1e-10, atol=0)
    line = (tmp_path / "avg" / names[0]).read_text().splitlines()[0]
    assert "  " in line and line.split()[0].count("e") == 1  # %.10e, two-blank delimiter


@pytest.mark.parametrize("truth", [(1.0, 4.0, 4.0, 4.0, 0.0, 0.0),
                                   (0.7, 5.5, 4.2, 6.1, 1.3, 0.9)])
def test_fit_recovers_the_source_parameters(truth):
    qo, qs, ql = grid()
    c = radii.gaussian_3d((qo, qs, ql), *truth).reshape(qo.shape)
    err = np.full_like(c, 0.01)
    guess = (1.0, 5.0, 5.0, 5.0, 0.5, 0.5) if truth[4] == 0 else radii.GUESS
    p, e, r2 = radii.fit_radii(qo, qs, ql, c, err, 0.1, guess=guess)
    assert r2 > 1 - 1e-12
    assert np.allclose(p[:4], truth[:4], rtol=1e-6, atol=1e-6)
    # the cross terms enter as R^2: only |R_os|, |R_ol| are fixed
    assert np.allclose(np.abs(p[4:]), truth[4:], rtol=1e-4, atol=2e-3)


def test_correlation_of_a_writer_table():
    qo, qs, ql = grid(nq=5)
    rng = np.random.default_rng(2)
    num = rng.uniform(0, 50, qo.size)
    den = rng.uniform(0, 60, qo.size)
    den[3] = 0.0
    table = np.stack([qo.ravel(), qs.ravel(), ql.ravel(), num, den], axis=1)
    a, b, c3, c, err = radii.correlation(table)
    assert np.array_equal(a, qo) and np.array_equal(b, qs) and np.array_equal(c3, ql)
    want = np.where(den > 0, num / np.where(den > 0, den, 1), np.nan).reshape(qo.shape)
    assert np.array_equal(np.isnan(c), np.isnan(want))
    assert np.array_equal(c[~np.isnan(c)], want[~np.isnan(want)])
    assert np.isnan(err.ravel()[3]) and (err.ravel()[den > 0] > 0).all()


def test_fit_table_rows():
    qo, qs, ql = grid()
    truth = (0.9, 4.5, 5.0, 3.5, 0.0, 0.0)
    c = radii.gaussian_3d((qo, qs, ql), *truth)
    table = np.stack([qo.ravel(), qs.ravel(), ql.ravel(), 1e4 * c, np.full(c.size, 1e4)], axis=1)
    rows = radii.fit_table(table, q_cuts=(0.075, 0.1))
    assert rows.shape == (2, 13)
    for r in rows:
        assert np.allclose(r[[1, 3, 5, 7]], truth[:4], rtol=1e-5)
