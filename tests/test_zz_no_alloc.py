"""hbt_calculate_noalloc, the allocation-free entry point, against the entry
point it varies: hbt_upload + hbt_run_same + hbt_run_mixed of a context with
the plain FP64 kernel, on the same buffer and mixing plan, through the C ABI.

Two calls in a row on fresh histograms, the second with a smaller buffer (the
scratch array keeps the first call's contents).  Bars: every count (bin
counts, denominators) bit-exact; every FP64 sum within 1e-12 under
err = |S - S_ref| / max(|S_ref|, sum|.|_bin), the floor sum|cos| / sum|q|
of each bin from the oracle.  The caller's device arrays and stream come
from the HIP runtime the engine library is linked with.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from hadronic_afterburner_toolkit_amd import _lib, hbt, samples

pytestmark = pytest.mark.gpu

PARAMS = dict(qnpts=21, q_min=-0.15, q_max=0.15, n_KT=6, KT_min=0.2, KT_max=0.7,
              HBTrap_min=-1.0, HBTrap_max=1.0, needed_number_of_pairs=1e18,
              long_comoving_boost=1)
SUMS = ((1, "q_out_sum", "abs_q_out"), (2, "q_side_sum", "abs_q_side"),
        (3, "q_long_sum", "abs_q_long"), (4, "num", "abs_num"))


def existing_entry(buf, M, ids, cs):
    g = hbt.HBTCorrelation(PARAMS, precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_PLAIN)
    g.upload(buf)
    g.run_same()
    g.run_mixed(M, ids, cs)
    r = g.results()
    g.close()
    return r


class Hip:
    """The caller's side of the entry point: device arrays and a stream from
    the HIP runtime the engine library runs on (loaded with it)."""

    def __init__(self):
        _lib.load()
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]
        self.ptrs = []

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(1, a.nbytes)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def to_host(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def stream(self):
        s = C.c_void_p()
        assert self.h.hipStreamCreate(C.byref(s)) == 0
        return s

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def noalloc_entry(hip, buf, M, ids, cs, stream):
    lib = _lib.load()
    cfg = hbt.config_from_params(PARAMS, precision=_lib.HBT_PREC_FP64,
                                 kernel=_lib.HBT_KERNEL_PLAIN)
    cols = [hip.to_device(np.asarray(a, np.float64)) for a in buf.soa()]
    off = hip.to_device(np.asarray(buf.ev_off, np.int64))
    d_ids = hip.to_device(np.asarray(ids, np.int32))
    d_cs = hip.to_device(np.asarray(cs, np.float64))
    nbins = PARAMS["n_KT"] * PARAMS["qnpts"] ** 3
    zeros = np.zeros(nbins * 8)
    bins = hip.to_device(zeros)
    soa = (C.c_void_p * 8)(*[c.value for c in cols])
    lens = np.diff(buf.ev_off)
    rc = lib.hbt_calculate_noalloc(C.byref(cfg), buf.nev, int(buf.n), int(lens.max()),
                                   off, soa, M, d_ids, d_cs, bins, stream)
    assert rc == 0, lib.hbt_noalloc_last_error()
    assert hip.h.hipStreamSynchronize(stream) == 0
    out = hip.to_host(bins, zeros).reshape(nbins, 8)
    hip.free()
    return out


def check(bins, r, floors):
    assert np.array_equal(bins[:, 0].astype(np.uint64), r.num_count)
    assert np.all(bins[:, 0] == np.floor(bins[:, 0]))
    assert np.array_equal(bins[:, 5].view(np.uint64), r.den)
    assert int(r.num_count.sum()) > 0 and int(r.den.sum()) > 0
    for col, key, fl in SUMS:
        ref = getattr(r, key)
        d = np.maximum(np.abs(ref), floors[fl])
        d = np.where(d > 0, d, 1.0)
        err = float((np.abs(bins[:, col] - ref) / d).max())
        assert err <= 1e-12, (key, err)


def test_noalloc_equals_existing_entry_point(oracle_mod):
    full = samples.generate(777, 14, 220.0)
    hip = Hip()
    stream = hip.stream()
    for k, (first, count) in enumerate(((0, 14), (3, 5))):
        buf = full.events(first, count)
        M, ids, cs = hbt.Random(40 + k).draw_plan(buf.nev, buf.nev, 0)
        print(f"no-alloc call {k}: {buf.nev} events, {buf.n} particles, M={M}", flush=True)
        r = existing_entry(buf, M, ids, cs)
        bins = noalloc_entry(hip, buf, M, ids, cs, stream)
        o = oracle_mod.Oracle(PARAMS)
        o.combine_same(buf)
        o.run_mixed_plan(buf, M, ids, cs)
        check(bins, r, o.arrays())
    hip.h.hipStreamDestroy(stream)
