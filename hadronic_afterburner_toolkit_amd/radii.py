"""Event averaging of HBT correlation files and the 3D Gaussian radius fit:
the step after the pair path, as the reference's post-processing scripts do it
(ebe_scripts/average_event_HBT_correlation_function.py:30-46,
ebe_scripts/fit_HBT_radii.py:13-101).

Files are the writer's output (HBT_correlation.cpp:726-783): one line per
(q_long, q_out, q_side) bin, q_long outermost and q_side innermost, columns
q_out q_side q_long num den_norm, where num is the sum of cos(q.dx / hbar c)
over the bin's same-event pairs and den_norm the bin's mixed-event pair count
times the K_T bin's numerator/denominator pair ratio.  The correlation
function is C = num / den_norm.

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


def write_table(path: str, data: np.ndarray) -> None:
    """The averaged table in the reference's format (:45-46)."""
    np.savetxt(path, data, fmt="%.10e", delimiter="  ")


def average_folders(folders, out_folder: str, pattern: str = "HBT*") -> list[str]:
    """For every file matching pattern in the first folder, the mean over
    all folders, written under the same name into out_folder."""
    folders = list(folders)
    os.makedirs(out_folder, exist_ok=True)
    names = sorted(os.path.basename(p) for p in glob.glob(os.path.join(folders[0], pattern)))
    for name in names:
        write_table(os.path.join(out_folder, name),
                    average_files(os.path.join(f, name) for f in folders))
    return names


def correlation(table: np.ndarray):
    """(q_out, q_side, q_long, C, C_err) of a 5-column table, each shaped
    (nq, nq, nq) as [q_long][q_out][q_side].  C = num / den_norm (NaN where
    den_norm is 0).  The files carry no variance, so C_err is an estimate:
    (1 + |C|) / sqrt(den_norm), the counting error of den_norm pairs whose
    cos terms have unit spread."""
    table = np.asarray(table, dtype=np.float64)
    if table.ndim != 2 or table.shape[1] != 5:
        raise ValueError("expected the writer's 5-column table (ecoOutput = 0)")
    nq = int(round(table.shape[0] ** (1.0 / 3.0)))
    if nq ** 3 != table.shape[0]:
        raise ValueError(f"{table.shape[0]} lines is not a cube of q bins")
    shp = (nq, nq, nq)
    num, den = table[:, 3], table[:, 4]
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(den > 0, num / den, np.nan)
        err = np.where(den > 0, (1.0 + np.abs(c)) / np.sqrt(den), np.nan)
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
    q_cut_min < |q| < q_cut and a finite C (fit_HBT_radii.py:70-88).
    Returns (params, errors, R^2) with params (lambda, R_out, R_side,
    R_long, R_os, R_ol)."""
    from scipy.optimize import curve_fit
    qn = np.sqrt(q_out ** 2 + q_side ** 2 + q_long ** 2)
    sel = (qn > q_cut_min) & (qn < q_cut) & np.isfinite(c) & np.isfinite(c_err)
    if sel.sum() < len(guess):
        raise ValueError("too few bins inside the q cut")
    q = [q_out[sel], q_side[sel], q_long[sel]]
    import warnings
    with warnings.catch_warnings():
        # a cross term fitted at 0 has no gradient (it enters as R^2): its
        # error comes back infinite, as with the reference's script
        warnings.simplefilter("ignore")
        params, cov = curve_fit(gaussian_3d, q, np.ravel(c[sel]), p0=list(guess),
                                sigma=np.ravel(c_err[sel]) + EPS, absolute_sigma=True)
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
