"""The C ABI library: it loads, exports every entry point include/hbt_hip.h
declares, and reports errors cleanly (CPU only; no compute calls)."""
import ctypes as C
import os
import re

import pytest

from hadronic_afterburner_toolkit_amd import _lib, hbt

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "hbt_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hbt_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "hbt_create" in names and "hbt_run_same" in names and "hbt_run_mixed" in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_covers_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_create_without_gpu_or_bad_args_fails_cleanly():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.hbt_create(None, C.byref(h)) == -1
    assert lib.hbt_last_error(None)
    cfg = _lib.HbtConfig()
    cfg.qnpts = 1  # invalid
    assert lib.hbt_create(C.byref(cfg), C.byref(h)) == -1
    assert b"qnpts" in lib.hbt_last_error(None)


def test_binning_must_be_atomics_slab_or_autotuned():
    """hbt_config.binning is validated before any device is touched: 0, 1, 2
    get past it (to the missing device here), anything else is HBT_ERR_ARG."""
    from hadronic_afterburner_toolkit_amd import hbt
    lib = _lib.load()
    h = C.c_void_p()
    for b, want in ((3, _lib.HBT_ERR_ARG), (-1, _lib.HBT_ERR_ARG)):
        _, cfg = hbt._make_config({"hbt_binning": b}, _lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 0)
        assert lib.hbt_create(C.byref(cfg), C.byref(h)) == want
        assert b"binning" in lib.hbt_last_error(None)
    for b in (0, 1, 2):
        _, cfg = hbt._make_config({"hbt_binning": b}, _lib.HBT_PREC_FP32C, _lib.HBT_KERNEL_FAST, 0)
        rc = lib.hbt_create(C.byref(cfg), C.byref(h))
        assert rc != _lib.HBT_ERR_ARG
        if rc == _lib.HBT_OK:  # a device is visible: the context is real
            lib.hbt_destroy(h)


def test_rng_works_without_gpu():
    r = hbt.Random(0)
    assert r.rand_int_uniform() == 1178568022


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="missing"):
        _lib.load()


def test_server_create_without_gpu_or_bad_args_fails_cleanly():
    lib = _lib.load()
    h = C.c_void_p()
    assert lib.hbt_server_create(None, None, C.byref(h)) == -1
    cfg = _lib.HbtConfig()
    cfg.qnpts, cfg.q_min, cfg.q_max, cfg.n_KT, cfg.KT_min, cfg.KT_max = 11, -0.1, 0.1, 3, 0.2, 0.6
    cfg.kernel = _lib.HBT_KERNEL_PLAIN
    assert lib.hbt_server_create(C.byref(cfg), None, C.byref(h)) == -1
    assert b"fast kernel" in lib.hbt_server_last_error(None)
    cfg.kernel = _lib.HBT_KERNEL_FAST
    sc = _lib.HbtServerConfig()
    sc.hbm_bytes, sc.max_resident, sc.max_step_items, sc.step_pairs = 1 << 30, 4, 1024, 1e9
    sc.tune = _lib.HbtTuning(3, 1024, 4, 8, 0.0)  # R must be 1, 2 or 4
    assert lib.hbt_server_create(C.byref(cfg), C.byref(sc), C.byref(h)) == -1
    sc.tune = _lib.HbtTuning(2, 1024, 4, 8, 0.0)
    rc = lib.hbt_server_create(C.byref(cfg), C.byref(sc), C.byref(h))
    if rc == 0:  # a GPU is visible: nothing more to check here
        lib.hbt_server_destroy(h)
    else:
        assert rc == -6 and h.value is None


@pytest.mark.skipif(not os.path.isdir("/root/reference/src"),
                    reason="needs the reference's sources (build container only)")
def test_dropin_class_links_into_the_reference_program():
    """The drop-in HBT_correlation compiles against the reference's own
    particleSamples, ParameterReader and Random headers and links into the
    reference's program (main.cpp, Analysis.cpp, ...) in place of the
    reference's class (oracle/Makefile target dropin)."""
    import subprocess
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "dropin"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    exe = os.path.join(REPO, "oracle", "_ref", "dropin", "hadronic_afterburner_tools.e")
    syms = subprocess.run(["nm", "-C", exe], capture_output=True, text=True).stdout
    # this repository's class (its gather through the reference's getters),
    # the reference's Analysis driving it, and the C ABI it calls
    assert "HBT_correlation::gather(bool" in syms
    assert "Analysis::HBTAnalysis()" in syms
    assert "particleSamples::read_in_particle_samples_and_filter()" in syms
    assert " U hbt_run_same" in syms


def test_noalloc_rejects_sizes_beyond_its_scratch():
    """hbt_calculate_noalloc checks its limits on the host, before any device
    work: more events or particles than its fixed scratch array holds, a
    configuration it does not cover, a binding pair cap."""
    lib = _lib.load()
    cfg = hbt.config_from_params(dict(qnpts=21, n_KT=4, KT_min=0.2, KT_max=0.8,
                                      needed_number_of_pairs=1e18),
                                 precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_PLAIN)
    dummy = C.c_void_p(1 << 20)  # never dereferenced: every call below fails its checks
    soa = (C.c_void_p * 8)(*([dummy] * 8))

    def call(c, nev, n, mx):
        return lib.hbt_calculate_noalloc(C.byref(c), nev, n, mx, dummy, soa, 1, dummy, dummy,
                                         dummy, None)

    assert call(cfg, (1 << 14) + 1, 100, 10) == _lib.HBT_ERR_CAPACITY
    assert b"EVENTS" in lib.hbt_noalloc_last_error()
    assert call(cfg, 10, (1 << 18) + 1, 10) == _lib.HBT_ERR_CAPACITY
    assert b"PARTICLES" in lib.hbt_noalloc_last_error()
    assert call(cfg, 10, 100, 101) == _lib.HBT_ERR_ARG
    capped = hbt.config_from_params(dict(qnpts=21, needed_number_of_pairs=3e7),
                                    precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_PLAIN)
    assert call(capped, 10, 100, 10) == _lib.HBT_ERR_ARG
    fast = hbt.config_from_params(dict(qnpts=21, needed_number_of_pairs=1e18),
                                  precision=_lib.HBT_PREC_FP64, kernel=_lib.HBT_KERNEL_FAST)
    assert call(fast, 10, 100, 10) == _lib.HBT_ERR_ARG
