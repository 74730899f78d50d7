"""ctypes binding of the C ABI in include/hbt_hip.h (libhbt_hip.so).

The library is built in-tree by ``build.py``.  Loading fails loudly when it
is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhbt_hip.so")

HBT_PREC_FP64 = 0
HBT_PREC_FP32C = 1
HBT_KERNEL_PLAIN = 0
HBT_KERNEL_FAST = 1

ERRORS = {0: "HBT_OK", -1: "HBT_ERR_ARG", -2: "HBT_ERR_HIP", -3: "HBT_ERR_NOMEM",
          -4: "HBT_ERR_STATE", -5: "HBT_ERR_CAPACITY", -6: "HBT_ERR_NODEVICE"}
HBT_OK, HBT_ERR_ARG, HBT_ERR_HIP, HBT_ERR_NOMEM = 0, -1, -2, -3
HBT_ERR_STATE, HBT_ERR_CAPACITY, HBT_ERR_NODEVICE = -4, -5, -6


class HbtConfig(C.Structure):
    _fields_ = [
        ("qnpts", C.c_int32), ("q_min", C.c_double), ("q_max", C.c_double),
        ("n_KT", C.c_int32), ("KT_min", C.c_double), ("KT_max", C.c_double),
        ("n_Kphi", C.c_int32), ("azimuthal_flag", C.c_int32),
        ("invariant_radius_flag", C.c_int32), ("long_comoving_boost", C.c_int32),
        ("HBTrap_min", C.c_double), ("HBTrap_max", C.c_double),
        ("needed_number_of_pairs", C.c_double), ("number_of_mixed_events", C.c_int32),
        ("precision", C.c_int32), ("kernel", C.c_int32), ("device", C.c_int32),
        ("kphi_guard_prad", C.c_int32), ("binning", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class HbtTuning(C.Structure):
    _fields_ = [("R", C.c_int32), ("jchunk", C.c_int32), ("kchunk", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("ms", C.c_double)]


_U64P = C.POINTER(C.c_uint64)
_F64P = C.POINTER(C.c_double)


class HbtResult(C.Structure):
    _fields_ = [
        ("num_count", _U64P), ("den", _U64P), ("q_out_sum", _F64P), ("q_side_sum", _F64P),
        ("q_long_sum", _F64P), ("num", _F64P), ("qinv_count", _U64P), ("qinv_den", _U64P),
        ("qinv_sum", _F64P), ("qinv_num", _F64P), ("cnt_num_kt", _U64P),
        ("cnt_den_kt", _U64P), ("cnt_num_qinv", _U64P), ("cnt_den_qinv", _U64P),
        ("cnt_num_ktphi", _U64P), ("cnt_den_ktphi", _U64P),
    ]


class HbtBfConfig(C.Structure):
    _fields_ = [("particle_alpha", C.c_int32), ("particle_beta", C.c_int32),
                ("BpT_min", C.c_double), ("BpT_max", C.c_double), ("Brap_max", C.c_double),
                ("Bnpts", C.c_int32), ("rap_type", C.c_int32), ("device", C.c_int32),
                ("kernel", C.c_int32)]


HBT_BF_KERNEL_FAST = 0
HBT_BF_KERNEL_EXACT = 1


class HbtBfInfo(C.Structure):
    _fields_ = [("host_ms", C.c_double), ("prep_ms", C.c_double), ("kernel_ms", C.c_double),
                ("pairs", C.c_uint64), ("fallback", C.c_uint64), ("kernel", C.c_int32),
                ("launches", C.c_int32)]


class HbtServerConfig(C.Structure):
    _fields_ = [("hbm_bytes", C.c_int64), ("max_resident", C.c_int32),
                ("max_step_items", C.c_int32), ("step_pairs", C.c_double),
                ("tune", HbtTuning), ("part_particles", C.c_int32), ("reserved", C.c_int32 * 7)]


HBT_REQ_STATES = {0: "queued", 1: "loading", 2: "active", 3: "draining", 4: "done",
                  5: "failed"}


class HbtRequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("nev", C.c_int32), ("n_kept", C.c_int64),
                ("enum_same", C.c_uint64), ("enum_mixed", C.c_uint64),
                ("footprint", C.c_int64), ("t_submit", C.c_double), ("t_queued", C.c_double),
                ("t_admit", C.c_double), ("t_active", C.c_double),
                ("t_first_launch", C.c_double), ("t_last_launch", C.c_double),
                ("t_done", C.c_double)]


class HbtServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64),
                ("launches", C.c_int64), ("items", C.c_int64), ("pairs", C.c_double),
                ("arena_bytes", C.c_int64), ("arena_used", C.c_int64),
                ("arena_peak", C.c_int64), ("resident", C.c_int32),
                ("resident_peak", C.c_int32), ("queued", C.c_int32), ("deferred", C.c_int32),
                ("busy_s", C.c_double)]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("hbt_create", C.c_int, [C.POINTER(HbtConfig), C.POINTER(C.c_void_p)]),
    ("hbt_destroy", C.c_int, [C.c_void_p]),
    ("hbt_last_error", C.c_char_p, [C.c_void_p]),
    ("hbt_reset", C.c_int, [C.c_void_p]),
    ("hbt_upload", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int64),
                             C.POINTER(_F64P)]),
    ("hbt_set_psi_ref", C.c_int, [C.c_void_p, C.c_double]),
    ("hbt_run_same", C.c_int, [C.c_void_p]),
    ("hbt_run_mixed", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), _F64P]),
    ("hbt_read", C.c_int, [C.c_void_p, C.POINTER(HbtResult)]),
    ("hbt_stage_counts", C.c_int, [C.c_void_p, _U64P, _U64P]),
    ("hbt_deferred_pairs", C.c_int, [C.c_void_p, _U64P]),
    ("hbt_rng_create", C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    ("hbt_rng_destroy", C.c_int, [C.c_void_p]),
    ("hbt_rng_int", C.c_int32, [C.c_void_p]),
    ("hbt_rng_uniform", C.c_double, [C.c_void_p]),
    ("hbt_rng_draw_plan", C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_int32), _F64P]),
    ("hbt_calculate_buffer", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hbt_last_kernel_ms", C.c_int, [C.c_void_p, _F64P, _F64P]),
    ("hbt_slot_info", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                C.POINTER(C.c_int64)]),
    ("hbt_upload_groups", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                    C.c_int32, C.POINTER(C.c_int64), C.POINTER(_F64P)]),
    ("hbt_prepare_batch", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    _F64P]),
    ("hbt_launch_batch", C.c_int, [C.c_void_p]),
    ("hbt_sync", C.c_int, [C.c_void_p, _F64P]),
    ("hbt_batch_info", C.c_int, [C.c_void_p, _U64P, _U64P, C.POINTER(C.c_int32)]),
    ("hbt_batch_graph", C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    ("hbt_batch_queues", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("hbt_slab_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int32), _U64P, _U64P]),
    ("hbt_calculate_noalloc", C.c_int, [C.POINTER(HbtConfig), C.c_int32, C.c_int64, C.c_int64,
                                        C.c_void_p,
                                        C.POINTER(C.c_void_p), C.c_int32, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
    ("hbt_noalloc_last_error", C.c_char_p, []),
    ("hbt_get_tuning", C.c_int, [C.c_void_p, C.POINTER(HbtTuning)]),
    ("hbt_set_tuning", C.c_int, [C.c_void_p, C.POINTER(HbtTuning)]),
    ("hbt_autotune", C.c_int, [C.c_void_p, C.POINTER(HbtTuning)]),
    ("hbt_write_outputs", C.c_int, [C.POINTER(HbtConfig), C.POINTER(HbtResult), C.c_char_p,
                                    C.c_int32, C.c_int32]),
    ("hbt_flow_event_plane_angle", C.c_double, [C.c_int32, _F64P, _F64P, C.c_int32]),
    ("hbt_reader_open", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("hbt_reader_open_mode", C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                       C.POINTER(C.c_void_p)]),
    ("hbt_reader_next", C.c_int32, [C.c_void_p]),
    ("hbt_reader_eof", C.c_int32, [C.c_void_p]),
    ("hbt_reader_buffer", C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_int64)),
                                    C.POINTER(_F64P), C.POINTER(_F64P),
                                    C.POINTER(C.c_int64)]),
    ("hbt_reader_close", C.c_int, [C.c_void_p]),
    ("hbt_reader_set_species", C.c_int, [C.c_void_p, C.c_void_p]),
    ("hbt_reader_bad_monval", C.c_int64, [C.c_void_p]),
    ("hbt_reader_monval", C.c_int, [C.c_void_p, C.POINTER(C.POINTER(C.c_int32))]),
    ("hbt_server_create", C.c_int, [C.POINTER(HbtConfig), C.POINTER(HbtServerConfig),
                                    C.POINTER(C.c_void_p)]),
    ("hbt_server_destroy", C.c_int, [C.c_void_p]),
    ("hbt_server_last_error", C.c_char_p, [C.c_void_p]),
    ("hbt_server_submit", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                    C.POINTER(_F64P), C.c_int32, C.c_int32,
                                    C.POINTER(C.c_int64)]),
    ("hbt_server_info", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(HbtRequestInfo)]),
    ("hbt_server_wait", C.c_int, [C.c_void_p, C.c_int64, C.c_double]),
    ("hbt_server_result", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(HbtResult)]),
    ("hbt_server_release", C.c_int, [C.c_void_p, C.c_int64]),
    ("hbt_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(HbtServerStats)]),
    ("hbt_bf_create", C.c_int, [C.POINTER(HbtBfConfig), C.POINTER(C.c_void_p)]),
    ("hbt_bf_destroy", C.c_int, [C.c_void_p]),
    ("hbt_bf_last_error", C.c_char_p, [C.c_void_p]),
    ("hbt_bf_reset", C.c_int, [C.c_void_p]),
    ("hbt_bf_calculate", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int32), C.POINTER(_F64P), C.c_int32,
                                   C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                   C.POINTER(_F64P), C.c_void_p]),
    ("hbt_bf_calculate_planned", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int32), C.POINTER(_F64P), C.c_int32,
                                           C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                           C.POINTER(_F64P), C.POINTER(C.c_int32), _F64P]),
    ("hbt_bf_read", C.c_int, [C.c_void_p, _F64P, C.POINTER(C.c_int64),
                              C.POINTER(C.c_int64)]),
    ("hbt_bf_write_outputs", C.c_int, [C.POINTER(HbtBfConfig), _F64P, C.c_char_p]),
    ("hbt_bf_last_info", C.c_int, [C.c_void_p, C.POINTER(HbtBfInfo)]),
    ("hbt_bf_set_species", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
    ("hbt_species_read", C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    ("hbt_species_count", C.c_int32, [C.c_void_p]),
    ("hbt_species_get", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                  C.POINTER(C.c_int32)]),
    ("hbt_species_free", C.c_int, [C.c_void_p]),
]

_lib = None


def load() -> C.CDLL:
    """Load libhbt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
            "g.build()'` (there is no fallback implementation)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class HbtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code
