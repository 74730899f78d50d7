"""Event averaging of HBT correlation files and the 3D Gaussian radius fit:
the step after the pair path, as the reference's post-processing scripts do it
(ebe_scripts/average_event_HBT_correlation_function.py:30-46,
ebe_scripts/fit_HBT_radii.py:13-101).

Files are the writer's output (HBT_correlation.cpp:726-783): one line per
(q_long, q_out, q_side) bin, q_long outermost and q_side innermost, columns
q_out q_side q_long num den_norm, where num is the sum of cos(q.dx / hbar c)
over the bin's same-event pairs and den_norm the bin's mixed-event pair count
times the K_T bin's numerator/denominator pair ratio.  The correlation
function is C = num / den_norm; averaged over events it is <num> /
<den_norm>, with the error the reference's h5 averager computes from the
event-by-event spread (event_average), which the fit takes as sigma.

    python -m hadronic_afterburner_toolkit_amd.radii average <event folders...> <out folder>
    python -m hadronic_afterburner_toolkit_amd.radii fit <results folder>
"""
from __future__ import annotations

import glob
import os
import sys

import numpy as np

# hbar c of the reference's fit script (fit_HBT_radii.py:10); the pair
# kernels use 0.197327053 GeV fm (parameters.h:4) for the cos phase.
HBARC_FIT = 0.19733
EPS = 1e-15  # fit_HBT_radii.py:11, added to the errors
Q_CUTS = (0.05, 0.075, 0.1, 0.125, 0.15)  # fit_HBT_radii.py:37
GUESS = (1.0, 5.0, 5.0, 5.0, 0.1, 0.1)    # fit_HBT_radii.py:75
RADII_HEADER = ("# q_cut[GeV]  lambda  lambda_err  R_out[fm]  R_out_err[fm]  "
                "R_side[fm]  R_side_err[fm]  R_long [fm]  R_long_err[fm]  "
                "R_os[fm]  R_os_err[fm]  R_ol[fm]  R_ol_err[fm]")


def average_files(paths) -> np.ndarray:
    """Column-wise mean of equally shaped tables, the reference's event
    average (average_event_HBT_correlation_function.py:33-43)."""
    paths = list(paths)
    if not paths:
        raise ValueError("no files to average")
    acc = None
    for p in paths:
        t = np.loadtxt(p)
        acc = t.copy() if acc is None else acc + t
    return acc / len(paths)


def event_average(tables) -> np.ndarray:
    """Event average of one K_T file over events with the correlation
    function and its error from the event-by-event spread, as the
    reference's averager computes them
    (average_event_HBT_correlation_function_h5.py:155-215).

    Legacy 6-column tables (q_out q_side q_long npair num den: pairs in the
    bin, sum of cos, mixed pairs in the bin) give the reference's 8 columns:
    the event means of the six, then C = num / (Npair_num / Npair_den * den)
    and its error, where Npair_* are the tables' column sums (the K_T bin's
    pair totals) and the error combines the event variances of num, den and
    of the totals with their covariances, over sqrt(nev).

    The writer's 5-column tables (q_out q_side q_long num den_norm,
    HBT_correlation.cpp:767-776) already carry the pair-total ratio inside
    den_norm (:728-730, :762-765), so the totals are 1 with no spread and the
    same formula gives C = <num> / <den_norm> with the error from the event
    variances of num and den_norm; returned as 7 columns: the five means, C,
    C_err."""
    tabs = [np.asarray(t, dtype=np.float64) for t in tables]
    if not tabs:
        raise ValueError("no events to average")
    ncol = tabs[0].shape[1]
    if ncol not in (5, 6) or any(t.shape != tabs[0].shape for t in tabs):
        raise ValueError("expected equally shaped 5- or 6-column tables")
    nev = len(tabs)
    legacy = ncol == 6
    ci, di = (4, 5) if legacy else (3, 4)
    rows = tabs[0].shape[0]
    num = np.zeros(rows)
    s_num = np.zeros(rows)
    s_num_np = np.zeros(rows)
    den = np.zeros(rows)
    s_den = np.zeros(rows)
    s_den_np = np.zeros(rows)
    np_num = s_np_num = np_den = s_np_den = 0.0
    avg = np.zeros_like(tabs[0])
    for t in tabs:
        a_num = float(t[:, 3].sum()) if legacy else 1.0
        a_den = float(t[:, 5].sum()) if legacy else 1.0
        num += t[:, ci]
        s_num += t[:, ci] ** 2
        np_num += a_num
        s_np_num += a_num ** 2
        s_num_np += t[:, ci] * a_num
        den += t[:, di]
        s_den += t[:, di] ** 2
        np_den += a_den
        s_np_den += a_den ** 2
        s_den_np += t[:, di] * a_den
        avg += t
    avg /= nev
    num /= nev
    den /= nev
    np_num /= nev
    np_den /= nev
    v_num = s_num / nev - num ** 2
    v_den = s_den / nev - den ** 2
    v_np_num = s_np_num / nev - np_num ** 2
    v_np_den = s_np_den / nev - np_den ** 2
    c_num = s_num_np / nev - num * np_num
    c_den = s_den_np / nev - den * np_den
    with np.errstate(divide="ignore", invalid="ignore"):
        corr = num / (np_num / np_den * den)
        # relative variances; the pair-total terms vanish for 5-column tables
        r_num = v_num / num ** 2 + v_np_num / np_num ** 2 - 2.0 * c_num / (num * np_num)
        r_den = v_den / den ** 2 + v_np_den / np_den ** 2 - 2.0 * c_den / (den * np_den)
        # rounding can leave a zero spread slightly negative
        err_num = num / np_num * np.sqrt(np.maximum(r_num, 0.0)) / np.sqrt(nev)
        err_den = den / np_den * np.sqrt(np.maximum(r_den, 0.0)) / np.sqrt(nev)
        corr_err = np.sqrt((err_num / (den / np_den)) ** 2
                           + ((num / np_num) * err_den / ((den / np_den) ** 2)) ** 2)
    out = np.zeros((rows, ncol + 2))
    out[:, :ncol] = avg
    out[:, ncol], out[:, ncol + 1] = corr, corr_err
    return out


def write_table(path: str, data: np.ndarray) -> None:
    """The averaged table in the reference's format (:45-46)."""
    np.savetxt(path, data, fmt="%.10e", delimiter="  ")


def average_folders(folders, out_folder: str, pattern: str = "HBT*") -> list[str]:
    """For every file matching pattern in the first folder, the event average
    over all folders with C and its event-spread error (event_average),
    written under the same name into out_folder."""
    folders = list(folders)
    os.makedirs(out_folder, exist_ok=True)
    names = sorted(os.path.basename(p) for p in glob.glob(os.path.join(folders[0], pattern)))
    for name in names:
        tabs = [np.loadtxt(os.path.join(f, name)) for f in folders]
        out = event_average(tabs) if tabs[0].ndim == 2 and tabs[0].shape[1] in (5, 6) \
            else sum(tabs) / len(tabs)
        write_table(os.path.join(out_folder, name), out)
    return names


def correlation(table: np.ndarray):
    """(q_out, q_side, q_long, C, C_err) of a table, each shaped (nq, nq, nq)
    as [q_long][q_out][q_side]:
      - an averaged table (7 columns from the writer's files, 8 in the
        reference's layout): C and C_err as event_average computed them;
      - one writer file (5 columns): C = num / den_norm (NaN where den_norm
        is 0); one event carries no spread, so C_err is NaN and the fit
        weighs every bin alike."""
    table = np.asarray(table, dtype=np.float64)
    if table.ndim != 2 or table.shape[1] not in (5, 7, 8):
        raise ValueError("expected a writer table (5 columns, ecoOutput = 0) or an averaged "
                         "table (7 or 8 columns)")
    nq = int(round(table.shape[0] ** (1.0 / 3.0)))
    if nq ** 3 != table.shape[0]:
        raise ValueError(f"{table.shape[0]} lines is not a cube of q bins")
    shp = (nq, nq, nq)
    if table.shape[1] == 5:
        num, den = table[:, 3], table[:, 4]
        with np.errstate(divide="ignore", invalid="ignore"):
            c = np.where(den > 0, num / den, np.nan)
        err = np.full_like(c, np.nan)
    else:
        k = table.shape[1] - 2
        c, err = table[:, k].copy(), table[:, k + 1].copy()
    return (table[:, 0].reshape(shp), table[:, 1].reshape(shp), table[:, 2].reshape(shp),
            c.reshape(shp), err.reshape(shp))


def gaussian_3d(q, lam, r_out, r_side, r_long, r_os, r_ol):
    """lambda exp(-(R_o^2 q_o^2 + R_s^2 q_s^2 + R_l^2 q_l^2 + 2 q_o q_s R_os^2
    + 2 q_o q_l R_ol^2) / hbar c^2), radii in fm, q in GeV (the model of
    fit_HBT_radii.py:13-31, arXiv:1403.4972)."""
    q_out, q_side, q_long = q
    ro, rs, rl = r_out / HBARC_FIT, r_side / HBARC_FIT, r_long / HBARC_FIT
    ros, rol = r_os / HBARC_FIT, r_ol / HBARC_FIT
    return np.ravel(lam * np.exp(-((ro * q_out) ** 2 + (rs * q_side) ** 2 + (rl * q_long) ** 2
                                   + 2.0 * q_out * q_side * ros ** 2
                                   + 2.0 * q_out * q_long * rol ** 2)))


def fit_radii(q_out, q_side, q_long, c, c_err, q_cut: float, q_cut_min: float = 0.0,
              guess=GUESS):
    """Weighted least-squares fit of gaussian_3d to the bins with
    q_cut_min < |q| < q_cut and a finite C (fit_HBT_radii.py:70-88), with
    sigma = C_err + eps; unweighted when no bin has a finite C_err (a single
    event's file).
    Returns (params, errors, R^2) with params (lambda, R_out, R_side,
    R_long, R_os, R_ol)."""
    from scipy.optimize import curve_fit
    qn = np.sqrt(q_out ** 2 + q_side ** 2 + q_long ** 2)
    weighted = bool(np.isfinite(c_err).any())
    sel = (qn > q_cut_min) & (qn < q_cut) & np.isfinite(c)
    if weighted:
        sel &= np.isfinite(c_err)
    if sel.sum() < len(guess):
        raise ValueError("too few bins inside the q cut")
    q = [q_out[sel], q_side[sel], q_long[sel]]
    import warnings
    with warnings.catch_warnings():
        # a cross term fitted at 0 has no gradient (it enters as R^2): its
        # error comes back infinite, as with the reference's script
        warnings.simplefilter("ignore")
        if weighted:  # the averaged error column + eps (fit_HBT_radii.py:61, :74-76)
            params, cov = curve_fit(gaussian_3d, q, np.ravel(c[sel]), p0=list(guess),
                                    sigma=np.ravel(c_err[sel]) + EPS, absolute_sigma=True)
        else:
            params, cov = curve_fit(gaussian_3d, q, np.ravel(c[sel]), p0=list(guess))
    errors = np.sqrt(np.diag(cov))
    resid = c[sel] - gaussian_3d(q, *params)
    r2 = 1.0 - np.var(resid) / np.var(c[sel])
    return params, errors, float(r2)


def fit_table(table, q_cuts=Q_CUTS) -> np.ndarray:
    """One row per q cut: q_cut, then (value, error) of lambda, R_out,
    R_side, R_long, R_os, R_ol, as the reference's radii tables."""
    q_out, q_side, q_long, c, err = correlation(table)
    rows = []
    for qc in q_cuts:
        p, e, _ = fit_radii(q_out, q_side, q_long, c, err, qc)
        rows.append([qc] + [v for pair in zip(p, e) for v in pair])
    return np.array(rows)


def fit_folder(folder: str, q_cuts=Q_CUTS) -> list[str]:
    """HBT_radii_KT_<a>_<b>.dat beside every HBT_correlation_function_KT_<a>_<b>.dat
    of a results folder (text instead of the reference's HDF5 datasets)."""
    written = []
    for path in sorted(glob.glob(os.path.join(folder, "HBT_correlation_function_KT_*.dat"))):
        name = os.path.basename(path)
        if "_Kphi_" in name:
            continue
        out = os.path.join(folder, name.replace("HBT_correlation_function_", "HBT_radii_"))
        np.savetxt(out, fit_table(np.loadtxt(path), q_cuts), fmt="%.8e", delimiter="  ",
                   header=RADII_HEADER[2:])
        written.append(out)
    return written


def main(argv) -> int:
    if len(argv) >= 4 and argv[1] == "average":
        names = average_folders(argv[2:-1], argv[-1])
        print(f"averaged {len(names)} files over {len(argv) - 3} folders into {argv[-1]}")
        return 0
    if len(argv) == 3 and argv[1] == "fit":
        for p in fit_folder(argv[2]):
            print(p)
        return 0
    print(__doc__)
    return 1


if __name__ == "__main__":
    sys.exit(main(sys.argv))
