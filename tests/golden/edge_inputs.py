"""Hand-placed particle pairs on the decision edges of the HBT pair cascade
(HBT_correlation.cpp:311-429 same-event, :574-678 mixed).

Every designed pair has its pair momentum K along a coordinate axis, so that
cos(phi_K) and sin(phi_K) are exactly +-1 and 0 and the LCMS components are
plain coordinate differences:

* K along +x (K_y = 0 exactly): K_perp = sqrt(K_x^2) = K_x,
  q_out = q_x * 1 + q_y * 0 = q_x, q_side = q_y * 1 - q_x * 0 = q_y;
* p_z1 = -p_z2: K_z = 0, beta = 0, M_t = sqrt(K_E^2) = K_E, gamma = 1, so
  q_long = 1 * (q_z - 0 * q_E) = q_z = 2 p_z1.

The values hit exactly:

* q_out, q_side, q_long equal to lo = (q_min - dq/2) + 1e-8 and
  hi = (q_max + dq/2) - 1e-8 (same-event keeps both, mixed drops hi:
  ``>`` vs ``>=``, HBT_correlation.cpp:364 vs :613), their neighbours one ulp
  inside and outside, and interior bin edges q_base + k dq with their ulp
  neighbours (the truncation of (q - q_base)/dq decides the bin);
* K_perp^2 equal to KT_min^2 and KT_max^2 (K_perp = KT_max gives the
  K_T index n_KT - 1, accumulated but never written), one ulp outside both,
  and the interior K_T edges;
* (K_phi case) pair axes at angle 0, pi/2, pi, 3 pi/2 from an event plane
  psi_ref = 0 exactly, and one just below 0 that the wrap loop turns into
  2 pi (K_phi index n_Kphi: the reference skips the pair,
  HBT_correlation.cpp:408-423).  psi_ref = atan2(sum sin 2phi, sum cos 2phi)/2
  (:233-249) is exactly 0 because every particle is followed by its mirror
  image (p_x, -p_y): the running sum of sin 2phi returns to exactly 0 after
  each mirror pair, and sum cos 2phi > 0.

q_out = q_x = p_x1 - p_x2 is exact only when p_x1 = p_x2 + v is representable:
``_exact_split`` searches p_x2 until fl(p_x2 + v) - p_x2 == v (Sterbenz's
lemma makes the subtraction itself exact) and checks it.

The builders return a Buffer of two events: event 0 holds particle 1 of every
designed pair, event 1 particle 2.  As one buffer they form every designed
pair as a same-event pair in the reference's orientation (particle 1 earlier
in gather order); mixing event 0 with event 1 under the rotation
(cos, sin) = (1, 0) forms them again as mixed pairs.
"""
from __future__ import annotations

import math

import numpy as np

from hadronic_afterburner_toolkit_amd.samples import PION_ID, PION_MASS, Buffer

# the parameter set of the 3D edge case (the bench's grid)
EDGE_PARAMS = dict(qnpts=41, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
                   HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
                   long_comoving_boost=1)
KPHI_PARAMS = dict(EDGE_PARAMS, azimuthal_flag=1, n_Kphi=4)


def derived(p):
    """The cascade's constants in the reference's expression order
    (HBT_correlation.cpp:28, :48, :318-319, :363-369)."""
    dq = (p["q_max"] - p["q_min"]) / (p["qnpts"] - 1)
    base = p["q_min"] - dq / 2.
    return dict(dq=dq, base=base, lo=p["q_min"] - dq / 2. + 1e-8,
                hi=p["q_max"] + dq / 2. - 1e-8,
                dKT=(p["KT_max"] - p["KT_min"]) / (p["n_KT"] - 1))


def _up(v):
    return math.nextafter(v, math.inf)


def _down(v):
    return math.nextafter(v, -math.inf)


def edge_values(p):
    """The q values placed exactly: window edges, their one-ulp neighbours,
    interior bin edges and their neighbours."""
    d = derived(p)
    vals = [d["lo"], _down(d["lo"]), _up(d["lo"]), d["hi"], _down(d["hi"]), _up(d["hi"])]
    for k in (1, 2, 20, 21, p["qnpts"] - 1):
        e = d["base"] + k * d["dq"]
        vals += [e, _down(e), _up(e)]
    return vals


def _exact_split(v, kx_lo=0.25, kx_hi=0.6):
    """(x1, x2) with x1 - x2 == v exactly and K_x = (x1 + x2)/2 inside
    [kx_lo, kx_hi], or None: with |p_x| >= 0.125 a difference is a multiple
    of 2^-55, so small interior edges (finer bits) have no exact split and
    are placed through q_side and q_long only (v/2 is always exact)."""
    for k in range(4000):
        x2 = 0.16 + 0.0001 * k
        x1 = x2 + v
        if x1 - x2 == v and x1 > 0 and kx_lo <= 0.5 * (x1 + x2) <= kx_hi:
            return x1, x2
    return None


def _energy(px, py, pz):
    return math.sqrt(PION_MASS * PION_MASS + px * px + py * py + pz * pz)


def _buffer(pairs, space=None):
    """Two events from designed pairs: particle 1 of each pair in event 0,
    particle 2 in event 1.  Each particle is (px, py, pz); space-time points
    come from a fixed generator so that the numerators are not trivial."""
    rng = np.random.default_rng(4242 if space is None else space)
    rows = [[a for a, _ in pairs], [b for _, b in pairs]]
    n = 2 * len(pairs)
    cols = {k: np.empty(n) for k in ("px", "py", "pz", "E", "x", "y", "z", "t")}
    k = 0
    for ev in rows:
        for px, py, pz in ev:
            cols["px"][k], cols["py"][k], cols["pz"][k] = px, py, pz
            cols["E"][k] = _energy(px, py, pz)
            k += 1
    cols["x"][:] = rng.normal(0.0, 4.0, n)
    cols["y"][:] = rng.normal(0.0, 4.0, n)
    cols["z"][:] = rng.normal(0.0, 2.0, n)
    cols["t"][:] = 9.0 + rng.exponential(2.0, n)
    off = np.array([0, len(pairs), n], np.int64)
    return Buffer(ev_off=off, monval=np.full(n, PION_ID, np.int64),
                  mass=np.full(n, PION_MASS), **cols)


def edge_pairs_3d(p=EDGE_PARAMS):
    """Designed pairs of the 3D case: list of ((px, py, pz), (px, py, pz))."""
    pairs = []
    d = derived(p)
    for v in edge_values(p):
        split = _exact_split(v)                       # q_out = v
        if split is None and v in (d["lo"], d["hi"]):
            raise ValueError(f"no exact split for the window edge {v!r}")
        if split is not None:
            pairs.append(((split[0], 0.0, 0.01), (split[1], 0.0, -0.01)))
        pairs.append(((0.43, v / 2, 0.013), (0.43, -v / 2, -0.013)))  # q_side = v
        pairs.append(((0.45, 0.0, v / 2), (0.45, 0.0, -v / 2)))      # q_long = v
    kts = [p["KT_min"], _down(p["KT_min"]), p["KT_max"], _up(p["KT_max"])]
    for k in range(1, p["n_KT"] - 1):
        e = p["KT_min"] + k * d["dKT"]
        kts += [e, _down(e), _up(e)]
    for kx in kts:                                    # K_perp = kx, q = (0, 0, 0.02)
        pairs.append(((kx, 0.0, 0.01), (kx, 0.0, -0.01)))
    return pairs


def edge_buffer_3d(p=EDGE_PARAMS):
    return _buffer(edge_pairs_3d(p))


def edge_pairs_kphi():
    """Designed pairs of the K_phi case, before mirroring."""
    tiny = 1e-300
    return [
        ((0.42, 0.03, 0.01), (0.40, -0.03, -0.01)),    # K along +x: angle 0
        ((-0.42, 0.03, 0.01), (-0.40, -0.03, -0.01)),  # K along -x: angle pi
        ((0.03, 0.42, 0.01), (-0.03, 0.40, -0.01)),    # K along +y: pi/2
        ((0.25, 0.0, 0.02), (0.27, -tiny, -0.02)),     # angle just below 0: 2 pi after the wrap
        ((0.30, 0.02, 0.0), (0.32, -0.02, 0.0)),
        ((-0.30, 0.0, 0.01), (-0.32, 0.0, -0.01)),     # K_y = 0 exactly, K_x < 0
    ]


def edge_buffer_kphi():
    """Same layout as edge_buffer_3d, with each particle followed by its
    mirror image (px, -py, pz) inside its event, so psi_ref = 0 exactly."""
    pairs = []
    for a, b in edge_pairs_kphi():
        am = (a[0], -a[1], a[2])
        bm = (b[0], -b[1], b[2])
        pairs += [(a, b), (am, bm)]
    return _buffer(pairs)


def build(kind):
    if kind == "3d":
        return edge_buffer_3d()
    if kind == "kphi":
        return edge_buffer_kphi()
    raise ValueError(kind)
