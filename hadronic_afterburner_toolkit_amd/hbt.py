"""Python host side of the HBT engine: a thin mirror of the reference's
HBT_correlation interface (src/HBT_correlation.h:63-85) over the C ABI.

``HBTCorrelation(params)`` takes the reference's parameter keys;
``calculate_HBT_correlation_function(buffer)`` runs one particle buffer
(same-event pairs of the whole buffer, then mixed pairs with the mixing plan
drawn from the shared random stream), histograms accumulating across calls;
``results()`` returns the raw histograms and counters as numpy arrays.  The
text writers live in the C++ drop-in class (csrc/host).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (HbtConfig, HbtError, HbtRequestInfo, HbtResult, HbtServerConfig,
                   HbtServerStats, HbtTuning)

DEFAULTS = dict(qnpts=31, q_min=-0.15, q_max=0.15, n_KT=5, KT_min=0.15, KT_max=0.55,
                n_Kphi=48, azimuthal_flag=0, invariant_radius_flag=0,
                long_comoving_boost=1, HBTrap_min=-0.5, HBTrap_max=0.5,
                needed_number_of_pairs=30000000, number_of_mixed_events=0)


def _f64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _u64p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def config_from_params(params: dict, precision: int = _lib.HBT_PREC_FP64,
                       kernel: int = _lib.HBT_KERNEL_FAST, device: int = 0) -> HbtConfig:
    p = dict(DEFAULTS)
    p.update({k: v for k, v in (params or {}).items() if k in DEFAULTS})
    cfg = HbtConfig()
    for name in ("qnpts", "n_KT", "n_Kphi", "azimuthal_flag", "invariant_radius_flag",
                 "long_comoving_boost", "number_of_mixed_events"):
        setattr(cfg, name, int(p[name]))
    for name in ("q_min", "q_max", "KT_min", "KT_max", "HBTrap_min", "HBTrap_max",
                 "needed_number_of_pairs"):
        setattr(cfg, name, float(p[name]))
    cfg.precision, cfg.kernel, cfg.device = precision, kernel, device
    # engine key: K_phi guard band in 1e-12 rad (0: the default 1e-12 rad)
    cfg.kphi_guard_prad = int((params or {}).get("hbt_kphi_guard_prad", 0))
    # engine key: 0 device-scope atomics (default), 1 slab binning where it
    # applies, 2 the faster of the two as hbt_autotune measures it
    cfg.binning = int((params or {}).get("hbt_binning", 0))
    return cfg


def write_outputs(params: dict, arrays: dict, path: str, eco_output: int = 0, what: int = 0):
    """The reference's text writers (HBT_correlation.cpp:694-855) on raw
    results: arrays holds the hbt_result fields (numpy)."""
    lib = _lib.load()
    cfg = config_from_params(params)
    keep = {}
    res = HbtResult()
    for name, typ in HbtResult._fields_:
        a = arrays.get(name)
        if a is None:
            continue
        dt = np.uint64 if typ is _lib._U64P else np.float64
        keep[name] = np.ascontiguousarray(a, dtype=dt)
        setattr(res, name, _u64p(keep[name]) if dt == np.uint64 else _f64p(keep[name]))
    rc = lib.hbt_write_outputs(C.byref(cfg), C.byref(res), path.encode(), int(eco_output),
                               int(what))
    if rc:
        raise HbtError(rc, "hbt_write_outputs")


def flow_event_plane_angle(px, py, n_order: int) -> float:
    lib = _lib.load()
    px = np.ascontiguousarray(px, np.float64)
    py = np.ascontiguousarray(py, np.float64)
    return float(lib.hbt_flow_event_plane_angle(len(px), _f64p(px), _f64p(py), int(n_order)))


def read_buffers(filename: str, event_buffer_size: int, particle_monval: int,
                 read_in_mode: int = 10, species_table: str | None = None):
    """The mode-10 (or mode-9 binary) reader of the engine library, buffer by
    buffer (the reference's read loop, Analysis.cpp:821-833).  A species
    class as particle_monval needs species_table, the path of a PDG table
    file (the reference reads EOS/pdg.dat)."""
    from .samples import Buffer
    lib = _lib.load()
    rd = C.c_void_p()
    rc = lib.hbt_reader_open_mode(filename.encode(), int(read_in_mode), int(event_buffer_size),
                                  int(particle_monval), C.byref(rd))
    if rc:
        raise HbtError(rc, f"hbt_reader_open {filename}")
    out = []
    try:
        if species_table is not None:
            t = C.c_void_p()
            rc = lib.hbt_species_read(species_table.encode(), C.byref(t))
            if rc:
                raise HbtError(rc, f"hbt_species_read {species_table}")
            rc = lib.hbt_reader_set_species(rd, t)
            lib.hbt_species_free(t)
            if rc:
                raise HbtError(rc, "hbt_reader_set_species")
        while not lib.hbt_reader_eof(rd):
            nev = lib.hbt_reader_next(rd)
            if nev < 0:
                bad = lib.hbt_reader_bad_monval(rd)
                raise HbtError(nev, f"hbt_reader_next {filename}" +
                               (f": monval {bad} is not in the species table" if bad else ""))
            off_p = C.POINTER(C.c_int64)()
            soa = (_lib._F64P * 8)()
            mass = _lib._F64P()
            total = C.c_int64()
            lib.hbt_reader_buffer(rd, C.byref(off_p), soa, C.byref(mass), C.byref(total))
            off = np.ctypeslib.as_array(off_p, shape=(nev + 1,)).copy()
            n = int(off[-1])
            cols = [np.ctypeslib.as_array(soa[a], shape=(n,)).copy() if n else np.zeros(0)
                    for a in range(8)]
            mon_p = C.POINTER(C.c_int32)()
            lib.hbt_reader_monval(rd, C.byref(mon_p))
            mon = np.ctypeslib.as_array(mon_p, shape=(n,)).copy() if n else np.zeros(0, np.int32)
            m = np.ctypeslib.as_array(mass, shape=(n,)).copy() if n else np.zeros(0)
            out.append((Buffer(off, *cols, monval=mon, mass=m), int(total.value)))
    finally:
        lib.hbt_reader_close(rd)
    return out


class Random:
    """RandomUtil::Random (Random.h:11-25): std::mt19937 with the reference's
    uniform_real_distribution<double>(0,1) and uniform_int_distribution<int>()."""

    def __init__(self, seed: int):
        self._lib = _lib.load()
        h = C.c_void_p()
        rc = self._lib.hbt_rng_create(int(seed), C.byref(h))
        if rc:
            raise HbtError(rc, "hbt_rng_create")
        self.h = h

    def rand_int_uniform(self) -> int:
        return int(self._lib.hbt_rng_int(self.h))

    def rand_uniform(self) -> float:
        return float(self._lib.hbt_rng_uniform(self.h))

    def draw_plan(self, nev: int, mixed_nev: int, number_of_mixed_events: int = 0):
        M = number_of_mixed_events if number_of_mixed_events > 0 else mixed_nev // 2 + 1
        ids = np.zeros(max(nev * M, 1), dtype=np.int32)
        cs = np.zeros(max(2 * nev * M, 2), dtype=np.float64)
        r = self._lib.hbt_rng_draw_plan(self.h, nev, mixed_nev, number_of_mixed_events,
                                        ids.ctypes.data_as(C.POINTER(C.c_int32)), _f64p(cs))
        if r < 0:
            raise HbtError(r, "hbt_rng_draw_plan")
        return r, ids[:nev * M], cs[:2 * nev * M]

    def __del__(self):
        try:
            self._lib.hbt_rng_destroy(self.h)
        except Exception:
            pass


@dataclass
class Results:
    num_count: np.ndarray
    den: np.ndarray
    q_out_sum: np.ndarray
    q_side_sum: np.ndarray
    q_long_sum: np.ndarray
    num: np.ndarray
    qinv_count: np.ndarray
    qinv_den: np.ndarray
    qinv_sum: np.ndarray
    qinv_num: np.ndarray
    cnt_num_kt: np.ndarray
    cnt_den_kt: np.ndarray
    cnt_num_qinv: np.ndarray
    cnt_den_qinv: np.ndarray
    cnt_num_ktphi: np.ndarray
    cnt_den_ktphi: np.ndarray


def pack_groups(bufs):
    """Group offsets, event offsets and the 8 contiguous columns of several
    buffers, as hbt_upload_groups takes them."""
    grp = np.zeros(len(bufs) + 1, np.int32)
    offs, cols = [np.zeros(1, np.int64)], [[] for _ in range(8)]
    base_p = 0
    for g, b in enumerate(bufs):
        grp[g + 1] = grp[g] + b.nev
        offs.append(np.asarray(b.ev_off[1:], np.int64) - int(b.ev_off[0]) + base_p)
        base_p += b.n
        a0 = int(b.ev_off[0])
        for k, a in enumerate(b.soa()):
            cols[k].append(np.asarray(a[a0:a0 + b.n], np.float64))
    off = np.ascontiguousarray(np.concatenate(offs))
    arrs = [np.ascontiguousarray(np.concatenate(c)) if c else np.zeros(0) for c in cols]
    return grp, off, arrs


class HBTCorrelation:
    """GPU HBT correlation accumulator with the reference's parameter keys."""

    def __init__(self, params: dict | None = None, precision: int = _lib.HBT_PREC_FP64,
                 kernel: int = _lib.HBT_KERNEL_FAST, device: int = 0):
        self.params, cfg = _make_config(params, precision, kernel, device)
        self._lib = _lib.load()
        h = C.c_void_p()
        rc = self._lib.hbt_create(C.byref(cfg), C.byref(h))
        if rc:
            raise HbtError(rc, self._lib.hbt_last_error(None).decode())
        self.h = h
        self.nphi = cfg.n_Kphi if cfg.azimuthal_flag else 1
        self._keep = []

    def _check(self, rc: int, what: str):
        if rc:
            raise HbtError(rc, f"{what}: {self._lib.hbt_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self._lib.hbt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        self._check(self._lib.hbt_reset(self.h), "hbt_reset")

    def upload(self, buf, slot: int = 0):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in buf.soa()]
        off = np.ascontiguousarray(buf.ev_off, dtype=np.int64)
        ptrs = (C.POINTER(C.c_double) * 8)(*[_f64p(a) for a in arrs])
        self._check(self._lib.hbt_upload(self.h, slot, buf.nev,
                                         off.ctypes.data_as(C.POINTER(C.c_int64)), ptrs),
                    "hbt_upload")

    def upload_groups(self, bufs, slot: int = 0):
        """Makes several buffers (groups) resident at once (batched form)."""
        self.upload_packed(pack_groups(bufs), slot)

    def upload_packed(self, packed, slot: int = 0):
        """Uploads groups already packed by pack_groups (no Python-side copies)."""
        grp, off, arrs = packed
        ptrs = (C.POINTER(C.c_double) * 8)(*[_f64p(a) for a in arrs])
        self._check(self._lib.hbt_upload_groups(
            self.h, slot, len(grp) - 1, grp.ctypes.data_as(C.POINTER(C.c_int32)), int(grp[-1]),
            off.ctypes.data_as(C.POINTER(C.c_int64)), ptrs), "hbt_upload_groups")

    def prepare_batch(self, M_per_group, ids, cs):
        Mg = np.ascontiguousarray(M_per_group, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        cs = np.ascontiguousarray(cs, dtype=np.float64)
        if ids.size == 0:
            ids = np.zeros(1, np.int32)
            cs = np.zeros(2)
        self._check(self._lib.hbt_prepare_batch(self.h, Mg.ctypes.data_as(C.POINTER(C.c_int32)),
                                                ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                                _f64p(cs)), "hbt_prepare_batch")

    def launch_batch(self):
        self._check(self._lib.hbt_launch_batch(self.h), "hbt_launch_batch")

    def sync(self) -> float:
        ms = C.c_double()
        self._check(self._lib.hbt_sync(self.h, C.byref(ms)), "hbt_sync")
        return ms.value

    def batch_info(self):
        a, b, n = C.c_uint64(), C.c_uint64(), C.c_int32()
        self._check(self._lib.hbt_batch_info(self.h, C.byref(a), C.byref(b), C.byref(n)),
                    "hbt_batch_info")
        return a.value, b.value, n.value

    def batch_graph(self) -> bool:
        """True when hbt_launch_batch replays the captured hipGraph."""
        v = C.c_int32()
        self._check(self._lib.hbt_batch_graph(self.h, C.byref(v)), "hbt_batch_graph")
        return bool(v.value)

    def batch_queues(self) -> list:
        """First item of each work queue of the prepared batch, and the end
        (one queue per XCD at most, every item of a group in one queue)."""
        n = C.c_int32()
        q = (C.c_int32 * 9)()
        self._check(self._lib.hbt_batch_queues(self.h, C.byref(n), q), "hbt_batch_queues")
        return [q[k] for k in range(n.value + 1)]

    def slab_info(self) -> dict:
        """Slab binning of the last fast launch: on, same-event pairs added
        directly (region full, or stored values rounding across a bin edge),
        region capacity in records."""
        on = C.c_int32()
        d, cap = np.zeros(1, np.uint64), np.zeros(1, np.uint64)
        self._check(self._lib.hbt_slab_info(self.h, C.byref(on), _u64p(d), _u64p(cap)),
                    "hbt_slab_info")
        return {"on": bool(on.value), "direct": int(d[0]), "capacity": int(cap[0])}

    def tuning(self) -> dict:
        t = HbtTuning()
        self._check(self._lib.hbt_get_tuning(self.h, C.byref(t)), "hbt_get_tuning")
        return {k: getattr(t, k) for k, _ in HbtTuning._fields_}

    def set_tuning(self, R=2, jchunk=1024, kchunk=8, blocks_per_cu=8):
        t = HbtTuning(R, jchunk, kchunk, blocks_per_cu, 0.0)
        self._check(self._lib.hbt_set_tuning(self.h, C.byref(t)), "hbt_set_tuning")

    def autotune(self) -> dict:
        t = HbtTuning()
        self._check(self._lib.hbt_autotune(self.h, C.byref(t)), "hbt_autotune")
        return {k: getattr(t, k) for k, _ in HbtTuning._fields_}

    def calculate_groups(self, bufs, rng: "Random", mixed_bufs=None):
        """Several buffers in one batch: same result as calling
        calculate_HBT_correlation_function on each in turn."""
        self.upload_groups(bufs, 0)
        if mixed_bufs is not None:
            self.upload_groups(mixed_bufs, 1)
        self._check(self._lib.hbt_calculate_buffer(self.h, rng.h), "hbt_calculate_buffer")

    def set_psi_ref(self, psi: float):
        self._check(self._lib.hbt_set_psi_ref(self.h, float(psi)), "hbt_set_psi_ref")

    def run_same(self):
        self._check(self._lib.hbt_run_same(self.h), "hbt_run_same")

    def run_mixed(self, M: int, ids: np.ndarray, cs: np.ndarray):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        cs = np.ascontiguousarray(cs, dtype=np.float64)
        self._check(self._lib.hbt_run_mixed(self.h, int(M),
                                            ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                            _f64p(cs)), "hbt_run_mixed")

    def calculate_HBT_correlation_function(self, buf, rng: Random, mixed_buf=None):
        """HBT_correlation::calculate_HBT_correlation_function (HBT_correlation.cpp:177-218)."""
        self.upload(buf, 0)
        if mixed_buf is not None:
            self.upload(mixed_buf, 1)
        self._check(self._lib.hbt_calculate_buffer(self.h, rng.h), "hbt_calculate_buffer")

    def kernel_ms(self):
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.hbt_last_kernel_ms(self.h, C.byref(a), C.byref(b)), "ms")
        return a.value, b.value

    def slot_info(self, slot: int = 0):
        nev, n = C.c_int32(), C.c_int64()
        self._check(self._lib.hbt_slot_info(self.h, slot, C.byref(nev), C.byref(n)), "slot")
        return nev.value, n.value

    def stage_counts(self):
        s = np.zeros(5, np.uint64)
        m = np.zeros(5, np.uint64)
        self._check(self._lib.hbt_stage_counts(self.h, _u64p(s), _u64p(m)), "stage")
        return s, m

    def deferred_pairs(self) -> int:
        """Pairs decided on the host so far (K_phi within the guard band of a
        bin edge)."""
        n = np.zeros(1, np.uint64)
        self._check(self._lib.hbt_deferred_pairs(self.h, _u64p(n)), "hbt_deferred_pairs")
        return int(n[0])

    def results(self) -> Results:
        r, res = _alloc_results(self.params, self.nphi)
        self._check(self._lib.hbt_read(self.h, C.byref(res)), "hbt_read")
        return r


def fold_kphi_mod_pi(r: Results, params: dict) -> Results:
    """K_phi bins of Phi_K mod pi: the reference bins Delta phi on [0, 2 pi)
    (HBT_correlation.cpp:408-416); bin k and bin k + n_Kphi/2 cover the same
    pair axis, so their sums are the n_Kphi/2 bins on [0, pi).  Exact: the
    folding follows the reference's own indexing."""
    p = dict(DEFAULTS)
    p.update(params)
    nkt, nphi, q3 = int(p["n_KT"]), int(p["n_Kphi"]), int(p["qnpts"]) ** 3
    if not int(p["azimuthal_flag"]) or nphi % 2:
        raise ValueError("folding needs azimuthal_flag=1 and an even n_Kphi")
    h = nphi // 2

    def fold(a, per):
        x = a.reshape(nkt, nphi, per)
        return (x[:, :h] + x[:, h:]).reshape(-1)

    out = {}
    for name in Results.__dataclass_fields__:
        a = getattr(r, name)
        if name in ("num_count", "den", "q_out_sum", "q_side_sum", "q_long_sum", "num"):
            out[name] = fold(a, q3)
        elif name in ("cnt_num_ktphi", "cnt_den_ktphi"):
            out[name] = fold(a, 1)
        else:
            out[name] = a.copy()
    return Results(**out)


def _make_config(params, precision, kernel, device):
    p = dict(DEFAULTS)
    if params:
        p.update({k: v for k, v in params.items() if k in DEFAULTS})
    cfg = HbtConfig()
    cfg.qnpts = int(p["qnpts"])
    cfg.q_min = float(p["q_min"])
    cfg.q_max = float(p["q_max"])
    cfg.n_KT = int(p["n_KT"])
    cfg.KT_min = float(p["KT_min"])
    cfg.KT_max = float(p["KT_max"])
    cfg.n_Kphi = int(p["n_Kphi"])
    cfg.azimuthal_flag = int(p["azimuthal_flag"])
    cfg.invariant_radius_flag = int(p["invariant_radius_flag"])
    cfg.long_comoving_boost = int(p["long_comoving_boost"])
    cfg.HBTrap_min = float(p["HBTrap_min"])
    cfg.HBTrap_max = float(p["HBTrap_max"])
    cfg.needed_number_of_pairs = float(p["needed_number_of_pairs"])
    cfg.number_of_mixed_events = int(p["number_of_mixed_events"])
    cfg.precision = precision
    cfg.kernel = kernel
    cfg.device = device
    # engine key: K_phi guard band in 1e-12 rad (0: the default 1e-12 rad)
    cfg.kphi_guard_prad = int((params or {}).get("hbt_kphi_guard_prad", 0))
    # engine key: 0 device-scope atomics (default), 1 slab binning where it
    # applies, 2 the faster of the two as hbt_autotune measures it
    cfg.binning = int((params or {}).get("hbt_binning", 0))
    return p, cfg


def _alloc_results(p, nphi):
    """Host arrays of one result and the hbt_result pointing at them."""
    q = int(p["qnpts"])
    nkt = int(p["n_KT"])
    nb = nkt * nphi * q ** 3
    nq = nkt * q
    r = Results(
        num_count=np.zeros(nb, np.uint64), den=np.zeros(nb, np.uint64),
        q_out_sum=np.zeros(nb), q_side_sum=np.zeros(nb), q_long_sum=np.zeros(nb),
        num=np.zeros(nb), qinv_count=np.zeros(nq, np.uint64),
        qinv_den=np.zeros(nq, np.uint64), qinv_sum=np.zeros(nq), qinv_num=np.zeros(nq),
        cnt_num_kt=np.zeros(nkt, np.uint64), cnt_den_kt=np.zeros(nkt, np.uint64),
        cnt_num_qinv=np.zeros(nkt, np.uint64), cnt_den_qinv=np.zeros(nkt, np.uint64),
        cnt_num_ktphi=np.zeros(nkt * nphi, np.uint64),
        cnt_den_ktphi=np.zeros(nkt * nphi, np.uint64))
    res = HbtResult()
    for name, _ in HbtResult._fields_:
        a = getattr(r, name)
        setattr(res, name, _u64p(a) if a.dtype == np.uint64 else _f64p(a))
    return r, res


class HBTServer:
    """Request server (hbt_server_* of include/hbt_hip.h): independent
    analyses submitted as they arrive, admitted by HBM footprint, batched
    into shared launches of the fast kernel.  One configuration (the
    reference's HBT keys) for all requests."""

    def __init__(self, params: dict | None = None, precision: int = _lib.HBT_PREC_FP64,
                 hbm_bytes: int = 16 << 30, max_resident: int = 16,
                 max_step_items: int = 1 << 18, step_pairs: float = 2e9,
                 tuning=(2, 1024, 4, 8), device: int = 0, part_particles: int = 0):
        self.params, cfg = _make_config(params, precision, _lib.HBT_KERNEL_FAST, device)
        self.nphi = cfg.n_Kphi if cfg.azimuthal_flag else 1
        self._lib = _lib.load()
        sc = HbtServerConfig()
        sc.hbm_bytes = int(hbm_bytes)
        sc.max_resident = int(max_resident)
        sc.max_step_items = int(max_step_items)
        sc.step_pairs = float(step_pairs)
        sc.tune = HbtTuning(*[int(v) for v in tuning], 0.0)
        sc.part_particles = int(part_particles)  # 0: the server's default (1e6)
        h = C.c_void_p()
        rc = self._lib.hbt_server_create(C.byref(cfg), C.byref(sc), C.byref(h))
        if rc:
            raise HbtError(rc, self._lib.hbt_server_last_error(None).decode())
        self.h = h

    def _check(self, rc: int, what: str):
        if rc:
            raise HbtError(rc, f"{what}: {self._lib.hbt_server_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self._lib.hbt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, buf, events_per_group: int, seed: int) -> int:
        """Queues one analysis of the events of buf (a samples.Buffer)."""
        a0 = int(buf.ev_off[0])
        arrs = [np.ascontiguousarray(a[a0:a0 + buf.n], dtype=np.float64) for a in buf.soa()]
        off = np.ascontiguousarray(np.asarray(buf.ev_off, np.int64) - a0)
        ptrs = (C.POINTER(C.c_double) * 8)(*[_f64p(a) for a in arrs])
        rid = C.c_int64()
        self._check(self._lib.hbt_server_submit(self.h, int(buf.nev),
                                                off.ctypes.data_as(C.POINTER(C.c_int64)), ptrs,
                                                int(events_per_group), int(seed),
                                                C.byref(rid)), "hbt_server_submit")
        return rid.value

    def info(self, rid: int) -> dict:
        i = HbtRequestInfo()
        self._check(self._lib.hbt_server_info(self.h, int(rid), C.byref(i)), "hbt_server_info")
        d = {k: getattr(i, k) for k, _ in HbtRequestInfo._fields_}
        d["state"] = _lib.HBT_REQ_STATES.get(d["state"], d["state"])
        return d

    def wait(self, rid: int, timeout_s: float = -1.0) -> bool:
        rc = self._lib.hbt_server_wait(self.h, int(rid), float(timeout_s))
        if rc == -4:  # HBT_ERR_STATE: timed out
            return False
        self._check(rc, "hbt_server_wait")
        return True

    def result(self, rid: int) -> Results:
        r, res = _alloc_results(self.params, self.nphi)
        self._check(self._lib.hbt_server_result(self.h, int(rid), C.byref(res)),
                    "hbt_server_result")
        return r

    def release(self, rid: int):
        self._check(self._lib.hbt_server_release(self.h, int(rid)), "hbt_server_release")

    def stats(self) -> dict:
        st = HbtServerStats()
        self._check(self._lib.hbt_server_get_stats(self.h, C.byref(st)), "hbt_server_get_stats")
        return {k: getattr(st, k) for k, _ in HbtServerStats._fields_ if k != "reserved"}


BF_KEYS = ("particle_alpha", "particle_beta", "BpT_min", "BpT_max", "Brap_max", "Bnpts",
           "rap_type")


def bf_config_from_params(params: dict, device: int = 0,
                          kernel: int = _lib.HBT_BF_KERNEL_FAST) -> "_lib.HbtBfConfig":
    return _lib.HbtBfConfig(int(params["particle_alpha"]), int(params["particle_beta"]),
                            float(params["BpT_min"]), float(params["BpT_max"]),
                            float(params["Brap_max"]), int(params["Bnpts"]),
                            int(params["rap_type"]), int(device), int(kernel))


def read_species_table(path: str) -> dict:
    """Charge, baryon number and strangeness per monval of a PDG table file in
    the reference's format (EOS/pdg.dat, particle_decay.cpp:23-129), with the
    anti-baryon entries the reference adds; int32 arrays keyed monval,
    charge, baryon, strange."""
    lib = _lib.load()
    t = C.c_void_p()
    rc = lib.hbt_species_read(path.encode(), C.byref(t))
    if rc:
        raise HbtError(rc, f"hbt_species_read {path}")
    try:
        n = lib.hbt_species_count(t)
        out = {k: np.zeros(n, np.int32) for k in ("monval", "charge", "baryon", "strange")}
        v = [C.c_int32() for _ in range(4)]
        for i in range(n):
            rc = lib.hbt_species_get(t, i, *[C.byref(x) for x in v])
            if rc:
                raise HbtError(rc, "hbt_species_get")
            for key, x in zip(("monval", "charge", "baryon", "strange"), v):
                out[key][i] = x.value
    finally:
        lib.hbt_species_free(t)
    return out


class BalanceFunction:
    """BalanceFunction (BalanceFunction.h) on the GPU: calculate per buffer,
    histograms and counts, the reference's output files.  species: a table
    from read_species_table, needed when particle_alpha or particle_beta is
    a species class (9999, +-9998, +-9997, +-9996)."""

    def __init__(self, params: dict, device: int = 0, kernel: int = _lib.HBT_BF_KERNEL_FAST,
                 species: dict | None = None):
        self._lib = _lib.load()
        self.params = {k: params[k] for k in BF_KEYS}
        self.cfg = bf_config_from_params(params, device, kernel)
        h = C.c_void_p()
        rc = self._lib.hbt_bf_create(C.byref(self.cfg), C.byref(h))
        if rc:
            raise HbtError(rc, "hbt_bf_create: " + self._lib.hbt_bf_last_error(None).decode())
        self.h = h
        if species is not None:
            self.set_species(species)

    def set_species(self, species: dict):
        cols = [np.ascontiguousarray(species[k], dtype=np.int32)
                for k in ("monval", "charge", "baryon", "strange")]
        self._check(self._lib.hbt_bf_set_species(
            self.h, len(cols[0]), *[c.ctypes.data_as(C.POINTER(C.c_int32)) for c in cols]),
            "hbt_bf_set_species")

    def _check(self, rc, what):
        if rc:
            raise HbtError(rc, f"{what}: {self._lib.hbt_bf_last_error(self.h).decode()}")

    @staticmethod
    def _buf_args(buf):
        off = np.ascontiguousarray(buf.ev_off, dtype=np.int64)
        mon = np.ascontiguousarray(buf.monval, dtype=np.int32)
        cols = [np.ascontiguousarray(getattr(buf, k), dtype=np.float64)
                for k in ("mass", "px", "py", "pz", "E")]
        ptrs = (_lib._F64P * 5)(*[c.ctypes.data_as(_lib._F64P) for c in cols])
        return off, mon, cols, ptrs

    def calculate_balance_function(self, buf, rng: "Random", mix=None):
        """One buffer (all species; mass and monval needed), as
        BalanceFunction::calculate_balance_function."""
        off, mon, cols, ptrs = self._buf_args(buf)
        if mix is not None:
            moff, mmon, mcols, mptrs = self._buf_args(mix)
            margs = (len(moff) - 1, moff.ctypes.data_as(C.POINTER(C.c_int64)),
                     mmon.ctypes.data_as(C.POINTER(C.c_int32)), mptrs)
        else:
            margs = (0, None, None, None)
        self._check(self._lib.hbt_bf_calculate(
            self.h, len(off) - 1, off.ctypes.data_as(C.POINTER(C.c_int64)),
            mon.ctypes.data_as(C.POINTER(C.c_int32)), ptrs, *margs, rng.h), "hbt_bf_calculate")

    def results(self) -> dict:
        n = 8 * int(self.params["Bnpts"]) * 20
        hist = np.zeros(n)
        nb, nbb = C.c_int64(), C.c_int64()
        self._check(self._lib.hbt_bf_read(self.h, hist.ctypes.data_as(_lib._F64P), C.byref(nb),
                                          C.byref(nbb)), "hbt_bf_read")
        return dict(hist=hist.reshape(8, int(self.params["Bnpts"]), 20), N_b=int(nb.value),
                    N_bbar=int(nbb.value))

    def info(self) -> dict:
        """Timing and counts of the last calculate_balance_function call."""
        i = _lib.HbtBfInfo()
        self._check(self._lib.hbt_bf_last_info(self.h, C.byref(i)), "hbt_bf_last_info")
        return {f: getattr(i, f) for f, _ in _lib.HbtBfInfo._fields_}

    def output_balance_function(self, path: str):
        hist = np.ascontiguousarray(self.results()["hist"].ravel())
        self._check(self._lib.hbt_bf_write_outputs(C.byref(self.cfg),
                                                   hist.ctypes.data_as(_lib._F64P),
                                                   path.encode()), "hbt_bf_write_outputs")

    def reset(self):
        self._check(self._lib.hbt_bf_reset(self.h), "hbt_bf_reset")

    def close(self):
        if getattr(self, "h", None):
            self._lib.hbt_bf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_bf_outputs(params: dict, hist, path: str):
    """The reference's balance-function files from histograms (host only)."""
    lib = _lib.load()
    cfg = bf_config_from_params(params)
    h = np.ascontiguousarray(np.asarray(hist, dtype=np.float64).ravel())
    rc = lib.hbt_bf_write_outputs(C.byref(cfg), h.ctypes.data_as(_lib._F64P), path.encode())
    if rc:
        raise HbtError(rc, "hbt_bf_write_outputs")
