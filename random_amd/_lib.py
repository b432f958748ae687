"""ctypes binding of ``libktm.so`` (the C ABI declared in ``include/ktm.h``).

The library is built in-tree by ``random_amd/csrc/Makefile``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libktm.so")

# status codes (ktm_status)
KTM_OK = 0
KTM_ERR_ARG = -1
KTM_ERR_IO = -2
KTM_ERR_FORMAT = -3
KTM_ERR_VELOCITY = -4
KTM_ERR_NOMEM = -5
KTM_ERR_DEVICE = -6
KTM_ERR_UNSUPPORTED = -7
STATUS_NAMES = {
    0: "KTM_OK", -1: "KTM_ERR_ARG", -2: "KTM_ERR_IO", -3: "KTM_ERR_FORMAT",
    -4: "KTM_ERR_VELOCITY", -5: "KTM_ERR_NOMEM", -6: "KTM_ERR_DEVICE",
    -7: "KTM_ERR_UNSUPPORTED",
}

PREP_REFERENCE, PREP_CLOSED = 0, 1
TRACE_F32, TRACE_BF16, TRACE_F16 = 0, 1, 2
ACCUM_PLAIN, ACCUM_KAHAN = 0, 1
INDEX_FAST = 0x100          # OR'ed into accum: relaxed traveltime index (CLOSED prep)
SEGY_IBM, SEGY_IEEE = 1, 5          # SEG-Y sample format codes


class KtmError(RuntimeError):
    """A non-zero ktm_status returned through the C ABI."""

    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what} failed: {STATUS_NAMES.get(status, status)}")


class KtmConfig(C.Structure):
    """ktm_config: the macros of Kirchhoff.c:17-65 as fields."""

    _fields_ = [
        ("start_iline", C.c_int32), ("end_iline", C.c_int32),
        ("start_xline", C.c_int32), ("end_xline", C.c_int32),
        ("inc_iline", C.c_double), ("inc_xline", C.c_double),
        ("dh", C.c_double), ("nh", C.c_int32), ("reserved0", C.c_int32),
        ("oh", C.c_double), ("ox", C.c_double), ("oy", C.c_double),
        ("angle_e_to_n", C.c_double), ("iline_distance", C.c_double),
        ("xline_distance", C.c_double), ("aperture", C.c_double),
        ("anti_alias_cutoff", C.c_double),
        ("nvel_ilines", C.c_int32), ("nvel_xlines", C.c_int32),
        ("dvel_ilines", C.c_int32), ("dvel_xlines", C.c_int32),
        ("ovel_ilines", C.c_int32), ("ovel_xlines", C.c_int32),
        ("ntblock", C.c_int32), ("jobs_x", C.c_int32), ("jobs_y", C.c_int32),
        ("image_origin_zero", C.c_int32), ("vel_lower_clamp", C.c_int32),
        ("reserved1", C.c_int32), ("scale_xy", C.c_double),
    ]


class KtmSetup(C.Structure):
    _fields_ = [
        ("ns", C.c_int32), ("nsm", C.c_int32), ("dt", C.c_float), ("dtm", C.c_float),
        ("trange_max", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("no_job", C.c_int32),
        ("nxp", C.c_int32), ("nyp", C.c_int32), ("nh", C.c_int32),
        ("sinA", C.c_float), ("cosA", C.c_float), ("reserved", C.c_int32),
    ]


class KtmStats(C.Structure):
    _fields_ = [("contributions", C.c_uint64), ("migrate_ms", C.c_double),
                ("total_ms", C.c_double)]


class KtmTraces(C.Structure):
    _fields_ = [("samples", C.c_void_p), ("coords", C.c_void_p), ("ntr", C.c_int64),
                ("ns", C.c_int32), ("dt", C.c_float)]


class KtmVelocity(C.Structure):
    _fields_ = [("v", C.c_void_p), ("nsm", C.c_int32), ("dtm", C.c_float)]


class KtmImage(C.Structure):
    _fields_ = [("km", C.c_void_p), ("fold", C.c_void_p), ("image", C.c_void_p)]


class KtmEomConfig(C.Structure):
    _fields_ = [("ox", C.c_double), ("oy", C.c_double), ("angle_e_to_n", C.c_double),
                ("oxcsp", C.c_double), ("oycsp", C.c_double), ("excsp", C.c_double),
                ("eycsp", C.c_double), ("scale_xy", C.c_double), ("dcsp_x", C.c_double),
                ("dcsp_y", C.c_double), ("dh", C.c_double), ("aperture", C.c_double),
                ("low_pass_cutoff", C.c_double), ("nh", C.c_int32), ("oh", C.c_int32),
                ("jobs_x", C.c_int32), ("jobs_y", C.c_int32), ("ntblock", C.c_int32),
                ("reserved", C.c_int32), ("velocity_start", C.c_double),
                ("velocity_end", C.c_double)]


class KtmEomSetup(C.Structure):
    _fields_ = [("ns", C.c_int32), ("dt", C.c_float), ("Nx", C.c_int32), ("Ny", C.c_int32),
                ("nx", C.c_int32), ("ny", C.c_int32), ("no_job", C.c_int32), ("nxp", C.c_int32),
                ("nyp", C.c_int32), ("nh", C.c_int32), ("sinA", C.c_float), ("cosA", C.c_float),
                ("start_x", C.c_float), ("start_y", C.c_float)]


GEOM_FIELDS = ("sxt", "syt", "gxt", "gyt", "cmpx", "cmpy", "off", "scale")

_vp, _i64, _i32, _f32 = C.c_void_p, C.c_int64, C.c_int, C.c_float
_PROTOS = {
    "ktm_config_default": (None, [C.POINTER(KtmConfig)]),
    "ktm_setup_init": (_i32, [C.POINTER(KtmConfig), _i32, _f32, _i32, _f32, C.POINTER(KtmSetup), _vp]),
    "ktm_plan_create": (_i32, [C.POINTER(_vp), C.POINTER(KtmConfig), _i32, _f32, _i32, _f32, _vp]),
    "ktm_plan_destroy": (None, [_vp]),
    "ktm_plan_setup": (_i32, [_vp, C.POINTER(KtmSetup)]),
    "ktm_plan_trace_row": (_i64, [_vp]),
    "ktm_plan_image_size": (_i64, [_vp]),
    "ktm_geometry": (_i32, [_vp, _vp, _i64, _vp, _vp]),
    "ktm_prep": (_i32, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    "ktm_image_reset": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "ktm_migrate_block": (_i32, [_vp, _vp, _i32, _i32, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "ktm_migrate_block_static": (_i32, [_vp, _vp, _i32, _i32, _vp, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "ktm_poststack": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "ktm_migrate": (_i32, [C.POINTER(KtmConfig), C.POINTER(KtmTraces), C.POINTER(KtmVelocity),
                           C.POINTER(KtmImage), C.POINTER(KtmStats)]),
    "ktm_ibm_to_float": (None, [_vp, _i64, _vp]),
    "ktm_segy_count": (_i32, [C.c_char_p, C.POINTER(_i32), C.POINTER(_f32), C.POINTER(_i32)]),
    "ktm_segy_read": (_i32, [C.c_char_p, _i64, _i64, _i32, _vp, _vp, _vp]),
    "ktm_segy_write": (_i32, [C.c_char_p, _vp, _vp, _vp, _i64, _i32, _i32]),
    "ktm_segy_read_velocity": (_i32, [C.c_char_p, C.POINTER(KtmConfig), _vp, C.POINTER(_i32),
                                      C.POINTER(_f32), C.POINTER(_i64)]),
    "ktm_segy_file_header": (None, [C.POINTER(KtmConfig), _i32, _f32, _i32, _vp]),
    "ktm_segy_image_trace_header": (None, [C.POINTER(KtmConfig), _i32, _f32, _i32, _i32, _i32, _i32, _vp]),
    "ktm_globals_init": (_i32, [C.POINTER(KtmConfig), C.c_char_p, C.c_char_p, _vp, _i32, _f32, C.c_char_p]),
    "kirchhoff": (None, [_i32]),
    "ktm_last_status": (_i32, []),
    "ktm_last_stats": (_i32, [C.POINTER(KtmStats)]),
    "ktm_globals_free": (None, []),
    "ktm_device_sync": (_i32, []),
    "ktm_synth_traces": (_i32, [_vp, _vp, _i64, _vp, _i32, _f32, _f32, _f32, _f32, _vp, _vp]),
    "ktm_synth_noise": (_i32, [_vp, _i64, _f32, C.c_uint64, _vp]),
    "ktm_selftest_sqrt": (_i32, [_vp, _i64, _vp, _vp]),
    "ktm_gate_threshold": (C.c_float, [C.c_double]),
    "ktm_demigrate": (_i32, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "ktm_demigrate_index": (_i32, [_vp, _vp, _vp, _i64, _vp, _i32, _vp]),
    "ktm_poststack_adjoint": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "ktm_ingest": (_i32, [_vp, _vp, _i64, _i32, _i32, _i32, _vp, _vp, _vp, _vp]),
    "ktm_segy_decode": (_i32, [_vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    "ktm_eom_config_default": (None, [C.POINTER(KtmEomConfig)]),
    "ktm_eom_setup_init": (_i32, [C.POINTER(KtmEomConfig), _i32, _f32, C.POINTER(KtmEomSetup),
                                  _vp, _vp]),
    "ktm_eom_plan_create": (_i32, [C.POINTER(_vp), C.POINTER(KtmEomConfig), _i32, _f32, _vp]),
    "ktm_eom_plan_destroy": (_i32, [_vp]),
    "ktm_eom_plan_setup": (_i32, [_vp, C.POINTER(KtmEomSetup)]),
    "ktm_eom_image_size": (_i64, [_vp]),
    "ktm_eom_reset": (_i32, [_vp, _vp, _vp, _vp]),
    "ktm_eom_block": (_i32, [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "ktm_eom_post": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "ktm_probe_lds": (_i32, [_vp, _i32, _i32, _vp]),
    "ktm_probe_valu": (_i32, [_i32, _vp, _i32, _i32, _vp]),
    "ktm_probe_valu_ops": (_i32, [_i32]),
    "ktm_eom_job_size": (_i64, [_vp]),
    "ktm_eom_job_reset": (_i32, [_vp, _vp, _vp, _vp]),
    "ktm_eom_job_block": (_i32, [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp]),
    "ktm_eom_job_post": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "ktm_eom_job_records": (_i32, [_vp, _vp, _i32, _vp, C.POINTER(_i64), _vp]),
    "ktm_combine_script": (_i32, [C.c_char_p, C.c_char_p, _i32]),
    "ktm_combine_files": (_i32, [C.c_char_p, C.c_char_p, _i32]),
    "ktm_eom_globals_init": (_i32, [C.POINTER(KtmEomConfig), C.c_char_p, C.c_char_p]),
    "csp": (None, [_i32]),
    "ktm_eom_last_status": (_i32, []),
    "ktm_eom_globals_free": (None, []),
    "ktm_segy_encode": (_i32, [_vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    "ktm_image_records": (_i32, [_vp, _vp, _i32, _vp, _vp]),
    "ktm_device_alloc": (_i32, [C.POINTER(_vp), C.c_size_t]),
    "ktm_device_free": (_i32, [_vp]),
    "ktm_copy_to_device": (_i32, [_vp, _vp, C.c_size_t]),
    "ktm_copy_to_host": (_i32, [_vp, _vp, C.c_size_t]),
    "ktm_vgrad_create": (_i32, [C.POINTER(_vp), C.POINTER(KtmConfig), _i32, _f32, _i32, _i32, _i32]),
    "ktm_vgrad_destroy": (_i32, [_vp]),
    "ktm_vgrad_bind": (_i32, [_vp, _vp, _i64, _vp, _i64, _vp]),
    "ktm_vgrad_velocity": (_i32, [_vp, _vp, _vp, _vp]),
    "ktm_vgrad_eval": (_i32, [_vp, _vp, C.POINTER(C.c_double), _vp, _vp, _vp]),
    "ktm_vgrad_pairs": (_i64, [_vp]),
    "ktm_selftest_traveltime": (_i32, [_vp, _i64, _f32, _i32, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load libktm.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                " (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            if hasattr(L, name):
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != KTM_OK:
        raise KtmError(status, what)
