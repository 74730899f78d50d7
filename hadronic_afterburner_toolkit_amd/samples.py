"""Synthetic particle samples for the HBT path and the reference's mode-10 text format.

The source model is the one the HBT workloads are quoted on: a thermal-like
transverse spectrum dN/dpT ~ pT exp(-mT/T) on [0.1, 2] GeV, uniform azimuth,
rapidity uniform in [-1, 1], Gaussian transverse source of width sigma, and
Bjorken-like longitudinal space-time (eta_s = y + N(0, 0.5),
tau = 8 fm/c + Exp(2 fm/c)).

The mode-10 format is what particleSamples::read_in_particle_samples_gzipped
reads (particleSamples.cpp:1247-1286): a line with the particle count of an
event, then one line per particle ``monval mass t x y z E px py pz``.
Values are written with 17 significant digits, so parsing them back gives the
same doubles in every reader.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass

import numpy as np

PION_MASS = 0.13957
KAON_MASS = 0.49368
PION_ID = 211
KAON_ID = 321


@dataclass
class Buffer:
    """Events in structure-of-arrays form: ``ev_off`` has nev+1 entries."""

    ev_off: np.ndarray  # int64
    px: np.ndarray
    py: np.ndarray
    pz: np.ndarray
    E: np.ndarray
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    t: np.ndarray
    monval: np.ndarray | None = None
    mass: np.ndarray | None = None

    @property
    def nev(self) -> int:
        return len(self.ev_off) - 1

    @property
    def n(self) -> int:
        return int(self.ev_off[-1] - self.ev_off[0])

    def soa(self):
        return [self.px, self.py, self.pz, self.E, self.x, self.y, self.z, self.t]

    def events(self, first: int, count: int) -> "Buffer":
        """The sub-buffer of events [first, first+count)."""
        a, b = int(self.ev_off[first]), int(self.ev_off[first + count])
        sl = slice(a, b)
        return Buffer(
            ev_off=(self.ev_off[first:first + count + 1] - a).astype(np.int64),
            px=self.px[sl], py=self.py[sl], pz=self.pz[sl], E=self.E[sl],
            x=self.x[sl], y=self.y[sl], z=self.z[sl], t=self.t[sl],
            monval=None if self.monval is None else self.monval[sl],
            mass=None if self.mass is None else self.mass[sl],
        )

    def select(self, monval: int) -> "Buffer":
        """Species filter by exact Monte-Carlo id (particleSamples.cpp:1329-1346)."""
        keep = self.monval == monval
        counts = np.add.reduceat(keep.astype(np.int64), self.ev_off[:-1]) if self.n else \
            np.zeros(self.nev, np.int64)
        # reduceat misbehaves on empty segments: fix them up
        lens = np.diff(self.ev_off)
        counts = np.where(lens == 0, 0, counts)
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        return Buffer(ev_off=off, px=self.px[keep], py=self.py[keep], pz=self.pz[keep],
                      E=self.E[keep], x=self.x[keep], y=self.y[keep], z=self.z[keep],
                      t=self.t[keep], monval=self.monval[keep], mass=self.mass[keep])


def _sample_pt(rng: np.random.Generator, n: int, mass: float, T: float,
               lo: float = 0.1, hi: float = 2.0) -> np.ndarray:
    """Rejection sampling of dN/dpT ~ pT exp(-mT/T) on [lo, hi]."""
    grid = np.linspace(lo, hi, 2001)
    fmax = float(np.max(grid * np.exp(-np.sqrt(grid * grid + mass * mass) / T))) * 1.01
    out = np.empty(0)
    while out.size < n:
        m = max(2 * (n - out.size), 64)
        pt = rng.uniform(lo, hi, m)
        f = pt * np.exp(-np.sqrt(pt * pt + mass * mass) / T)
        acc = rng.uniform(0.0, fmax, m) < f
        out = np.concatenate([out, pt[acc]])
    return out[:n]


def generate(seed: int, nev: int, mult, species=((PION_ID, PION_MASS, 0.25, 1.0),),
             sigma: float = 4.0) -> Buffer:
    """Generate ``nev`` events.

    ``mult`` is the mean of a Poisson multiplicity (float) or a callable
    ``mult(rng, nev) -> counts``.  ``species`` lists (monval, mass, T,
    weight); each particle's species is drawn with the given weights.
    """
    rng = np.random.default_rng(seed)
    if callable(mult):
        counts = np.asarray(mult(rng, nev), dtype=np.int64)
    else:
        counts = rng.poisson(mult, nev).astype(np.int64)
    n = int(counts.sum())
    w = np.array([s[3] for s in species], dtype=float)
    sp = rng.choice(len(species), size=n, p=w / w.sum()) if len(species) > 1 else \
        np.zeros(n, dtype=np.int64)
    monval = np.array([species[k][0] for k in range(len(species))])[sp]
    mass = np.array([species[k][1] for k in range(len(species))])[sp]
    pt = np.empty(n)
    for k, (_, m, T, _) in enumerate(species):
        sel = sp == k
        pt[sel] = _sample_pt(rng, int(sel.sum()), m, T)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    yr = rng.uniform(-1.0, 1.0, n)
    mt = np.sqrt(pt * pt + mass * mass)
    E = mt * np.cosh(yr)
    pz = mt * np.sinh(yr)
    px = pt * np.cos(phi)
    py = pt * np.sin(phi)
    xs = rng.normal(0.0, sigma, n)
    ys = rng.normal(0.0, sigma, n)
    eta = yr + rng.normal(0.0, 0.5, n)
    tau = 8.0 + rng.exponential(2.0, n)
    t = tau * np.cosh(eta)
    z = tau * np.sinh(eta)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return Buffer(ev_off=off, px=px, py=py, pz=pz, E=E, x=xs, y=ys, z=z, t=t,
                  monval=monval.astype(np.int64), mass=mass)


def log_uniform_mult(lo: float, hi: float):
    """Multiplicities log-uniform on [lo, hi] (central-to-peripheral spread)."""
    def f(rng, nev):
        return np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), nev))).astype(np.int64)
    return f


def write_mode10(path: str, buf: Buffer) -> None:
    """Write events in the reference's mode-10 text format (gzip)."""
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wt") as f:
        for e in range(buf.nev):
            a, b = int(buf.ev_off[e]), int(buf.ev_off[e + 1])
            f.write(f"{b - a}\n")
            for i in range(a, b):
                f.write("%d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n" % (
                    buf.monval[i], buf.mass[i], buf.t[i], buf.x[i], buf.y[i], buf.z[i],
                    buf.E[i], buf.px[i], buf.py[i], buf.pz[i]))


def read_mode10_all(path: str) -> Buffer:
    """Read every event of a mode-10 file (no species filter, no buffering)."""
    opener = gzip.open if path.endswith(".gz") else open
    counts, rows = [], []
    with opener(path, "rt") as f:
        lines = f.read().split("\n")
    k = 0
    while k < len(lines) and lines[k].strip():
        n = int(lines[k].split()[0])
        counts.append(n)
        rows.extend(lines[k + 1:k + 1 + n])
        k += 1 + n
    arr = np.array([r.split() for r in rows], dtype=object) if rows else np.zeros((0, 10))
    num = arr[:, 1:].astype(float) if rows else np.zeros((0, 9))
    mon = arr[:, 0].astype(np.int64) if rows else np.zeros(0, np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return Buffer(ev_off=off, mass=num[:, 0], t=num[:, 1], x=num[:, 2], y=num[:, 3],
                  z=num[:, 4], E=num[:, 5], px=num[:, 6], py=num[:, 7], pz=num[:, 8],
                  monval=mon)


def buffers_by_budget(full: Buffer, event_buffer_size: int):
    """Split events like the reference's reader: whole events are taken while
    the running particle total (all species) is below event_buffer_size
    (particleSamples.cpp:1256-1284).  The end of the file is only noticed by
    a read that finds nothing, so when a buffer fills up exactly with the
    last event, one more, empty buffer follows (the reference's
    ``while (!end_of_file())`` loop, Analysis.cpp:821).  Returns a list of
    (first, count)."""
    out, e, nev = [], 0, full.nev
    lens = np.diff(full.ev_off)
    while True:
        first, tot = e, 0
        while tot < event_buffer_size and e < nev:
            tot += int(lens[e])
            e += 1
        out.append((first, e - first))
        if tot < event_buffer_size:  # the read hit the end of the file
            break
    return out


def read_buffers(path: str, event_buffer_size: int, monval: int):
    """Buffers of a mode-10 file, filtered to one species, as the reference's
    HBTAnalysis loop sees them (Analysis.cpp:821-833)."""
    full = read_mode10_all(path)
    return [full.events(f, c).select(monval) for f, c in buffers_by_budget(full,
                                                                          event_buffer_size)]
