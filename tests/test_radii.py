"""Event averaging and the 3D Gaussian radius fit (host; the step after the
pair path, ebe_scripts/average_event_HBT_correlation_function.py,
average_event_HBT_correlation_function_h5.py and fit_HBT_radii.py), pinned by
what fixes them: the average is the column-wise mean; identical events have
no spread; one- and two-event cases worked by hand fix C and its error in
both layouts; the fit returns the parameters a noise-free correlation
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
    # the five means, then C and its event-spread error
    assert back.shape == (27, 7)
    assert np.allclose(back[:, :5], sum(tabs) / 4, rtol=1e-10, atol=0)
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


def test_event_average_identical_events_have_no_spread():
    """Identical events: C is the single event's num / den_norm and its error
    is exactly zero (integer columns keep every moment exact)."""
    rng = np.random.default_rng(5)
    t = np.stack([rng.uniform(-0.1, 0.1, 27) for _ in range(3)]
                 + [rng.integers(1, 50, 27).astype(float), rng.integers(1, 60, 27).astype(float)],
                 axis=1)
    out = radii.event_average([t, t, t, t])
    assert out.shape == (27, 7)
    assert np.array_equal(out[:, :5], t)
    assert np.array_equal(out[:, 5], t[:, 3] / t[:, 4])
    assert np.array_equal(out[:, 6], np.zeros(27))


def test_event_average_two_events_by_hand():
    """Two events, one bin, writer layout: num = (10, 14), den_norm = (20, 28).
    C = 12 / 24 = 1/2; the event errors of the means are |10 - 14| / (2 sqrt 2)
    = sqrt 2 and |20 - 28| / (2 sqrt 2) = 2 sqrt 2, so
    C_err = sqrt((sqrt2 / 24)^2 + (12 * 2 sqrt2 / 24^2)^2) = sqrt(2/576 + 2/576) = 1/12."""
    a = np.array([[0.01, 0.02, 0.03, 10.0, 20.0]])
    b = np.array([[0.01, 0.02, 0.03, 14.0, 28.0]])
    out = radii.event_average([a, b])
    assert abs(out[0, 5] - 0.5) < 1e-15
    assert abs(out[0, 6] - 1.0 / 12.0) < 1e-15


def test_event_average_legacy_layout_by_hand():
    """The reference's 6-column layout (npair num den), one bin, two events:
    npair = (4, 4), num = (2, 4), den = (10, 30).  The pair totals are the
    column sums: Npair_num = 4 with no spread, Npair_den = den, so
    C = 3 / (4 / 20 * 20) = 3/4; the numerator's relative spread is
    sqrt(1/9) over sqrt 2 times num/Npair_num = 3/4, i.e. sqrt(2)/8, and the
    denominator's terms cancel (den over its own total): C_err = sqrt(2)/8."""
    a = np.array([[0.0, 0.0, 0.0, 4.0, 2.0, 10.0]])
    b = np.array([[0.0, 0.0, 0.0, 4.0, 4.0, 30.0]])
    out = radii.event_average([a, b])
    assert out.shape == (1, 8)
    assert np.array_equal(out[0, :6], [0.0, 0.0, 0.0, 4.0, 3.0, 20.0])
    assert abs(out[0, 6] - 0.75) < 1e-15
    assert abs(out[0, 7] - np.sqrt(2.0) / 8.0) < 1e-15


def test_averaged_folder_feeds_the_fit(tmp_path):
    """Writer files of several events -> averaged table with C and C_err ->
    fit weighted by C_err: a noise-free source with event-to-event
    fluctuations of the normalization only is fitted back."""
    qo, qs, ql = grid(nq=21)
    truth = (0.8, 4.0, 5.0, 4.5, 0.0, 0.0)
    c = radii.gaussian_3d((qo, qs, ql), *truth)
    folders = []
    for k, scale in enumerate((0.9, 1.0, 1.1, 1.2)):
        d = tmp_path / f"ev{k}"
        d.mkdir()
        den = np.full(c.size, 1e4 * scale)
        np.savetxt(d / "HBT_correlation_function_KT_0.2_0.3.dat",
                   np.stack([qo.ravel(), qs.ravel(), ql.ravel(), c * den, den], axis=1),
                   fmt="%.17e")
        folders.append(str(d))
    radii.average_folders(folders, str(tmp_path / "avg"))
    table = np.loadtxt(tmp_path / "avg" / "HBT_correlation_function_KT_0.2_0.3.dat")
    assert table.shape == (21 ** 3, 7)
    q_out, q_side, q_long, cc, err = radii.correlation(table)
    assert np.allclose(cc.ravel(), c, rtol=1e-9)
    assert (err > 0).all() and np.isfinite(err).all()
    p, e, r2 = radii.fit_radii(q_out, q_side, q_long, cc, err, 0.1,
                               guess=(1.0, 5.0, 5.0, 5.0, 0.5, 0.5))
    assert np.allclose(p[:4], truth[:4], rtol=1e-6)


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
    assert np.isnan(err).all()  # one event has no spread: the fit weighs bins alike


def test_fit_table_rows():
    qo, qs, ql = grid()
    truth = (0.9, 4.5, 5.0, 3.5, 0.0, 0.0)
    c = radii.gaussian_3d((qo, qs, ql), *truth)
    table = np.stack([qo.ravel(), qs.ravel(), ql.ravel(), 1e4 * c, np.full(c.size, 1e4)], axis=1)
    rows = radii.fit_table(table, q_cuts=(0.075, 0.1))
    assert rows.shape == (2, 13)
    for r in rows:
        assert np.allclose(r[[1, 3, 5, 7]], truth[:4], rtol=1e-5)
