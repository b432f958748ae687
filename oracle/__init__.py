"""CPU oracle of the Kirchhoff migration path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper of ``oracle/build/libktm_oracle.so`` (built from
``oracle/ktm_oracle.c`` by ``oracle/Makefile``), a from-scratch C restatement
of ``kirchhoff()`` in Kirchhoff.c with the reference's exact float/double
promotions.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may use it, and only as the checker or the timed CPU
baseline -- never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "libktm_oracle.so")
LIB_REFFLAGS_PATH = os.path.join(_HERE, "build", "libktm_oracle_refflags.so")
LIB_NATIVE_PATH = os.path.join(_HERE, "build", "libktm_oracle_native.so")

FP32, FP64_DATA, FP64_FULL = 0, 1, 2


class OrcParams(C.Structure):
    _fields_ = [
        ("start_iline", C.c_int32), ("end_iline", C.c_int32),
        ("start_xline", C.c_int32), ("end_xline", C.c_int32),
        ("inc_iline", C.c_double), ("inc_xline", C.c_double),
        ("dh", C.c_double), ("nh", C.c_int32), ("pad0", C.c_int32),
        ("oh", C.c_double), ("ox", C.c_double), ("oy", C.c_double),
        ("angle_e_to_n", C.c_double), ("iline_distance", C.c_double),
        ("xline_distance", C.c_double), ("aperture", C.c_double),
        ("anti_alias_cutoff", C.c_double),
        ("nvel_ilines", C.c_int32), ("nvel_xlines", C.c_int32),
        ("dvel_ilines", C.c_int32), ("dvel_xlines", C.c_int32),
        ("ovel_ilines", C.c_int32), ("ovel_xlines", C.c_int32),
        ("ntblock", C.c_int32), ("jobs_x", C.c_int32), ("jobs_y", C.c_int32),
        ("image_origin_zero", C.c_int32), ("vel_lower_clamp", C.c_int32),
        ("pad1", C.c_int32), ("scale_xy", C.c_double),
    ]


class OrcGrid(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("no_job", C.c_int32), ("sinA", C.c_float), ("cosA", C.c_float)]


_vp, _i64, _i32, _f32 = C.c_void_p, C.c_int64, C.c_int, C.c_float
_PROTOS = {
    "orc_grid_init": (None, [C.POINTER(OrcParams), C.POINTER(OrcGrid)]),
    "orc_ibm_block": (None, [_vp, _i64, _vp]),
    "orc_alias_filter": (_i32, [C.POINTER(OrcParams), _i32, _f32, _vp]),
    "orc_trace_prep_fp32": (None, [_vp, _i32, _f32, _i32, _vp]),
    "orc_trace_prep_fp64": (None, [_vp, _i32, _f32, _i32, _vp]),
    "orc_trace_geometry": (_i32, [C.POINTER(OrcParams), C.POINTER(OrcGrid), _vp, _vp]),
    "orc_migrate_job": (_i64, [C.POINTER(OrcParams), _i32, _f32, _i32, _f32, _i32, _vp, _vp,
                               _i64, _vp, _i32, _vp, _vp]),
    "orc_migrate_all": (_i64, [C.POINTER(OrcParams), _i32, _f32, _i32, _f32, _vp, _vp, _i64,
                               _vp, _i32, _vp, _vp, _i32]),
    "orc_poststack_fp32": (None, [_i32, _f32, _vp, _vp, _vp, _i64, _vp]),
    "orc_poststack_fp64": (None, [_i32, _f32, _vp, _vp, _vp, _i64, _vp]),
    "orc_segy_trace_header": (None, [C.POINTER(OrcParams), C.POINTER(OrcGrid), _i32, _f32, _i32,
                                     _i32, _i32, _i32, _i32, _i32, _vp]),
    "orc_segy_job_headers": (None, [C.POINTER(OrcParams), C.POINTER(OrcGrid), _i32, _f32, _i32,
                                    _vp]),
    "orc_segy_file_header": (None, [C.POINTER(OrcParams), C.POINTER(OrcGrid), _i32, _f32, _i32, _vp]),
    "orc_count_size": (_i32, [_vp, _i64, C.POINTER(_i32), C.POINTER(_f32), C.POINTER(_i32)]),
    "orc_vgrad": (C.c_double, [_i32, C.c_double, C.c_double, _i32, _i32, _vp, _f32, _vp, _vp, _vp, _i64,
                               _i32, _i32, _vp, _vp, _vp]),
    "orc_vgrad_velocity": (None, [_i32, _i32, _i32, _i32, _vp, _vp]),
}

_libs: dict = {}


def build_native() -> str:
    """Compile the -O3 -march=native build for the CPU this process runs on
    (make -C oracle native); returns its path."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, capture_output=True)
    return LIB_NATIVE_PATH


def lib(refflags: bool = False, native: bool = False) -> C.CDLL:
    path = LIB_REFFLAGS_PATH if refflags else (LIB_NATIVE_PATH if native else LIB_PATH)
    if path not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        L = C.CDLL(path)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[path] = L
    return _libs[path]


def _vpn(a) -> C.c_void_p:
    return a.ctypes.data_as(C.c_void_p)


def params(cfg) -> OrcParams:
    """OrcParams from a random_amd Config (or any object with the same fields)."""
    p = OrcParams()
    for name, _ in OrcParams._fields_:
        if name.startswith("pad"):
            continue
        setattr(p, name, getattr(cfg, name))
    return p


def grid(cfg) -> OrcGrid:
    g = OrcGrid()
    p = params(cfg)
    lib().orc_grid_init(C.byref(p), C.byref(g))
    return g


def ibm_to_float(raw: np.ndarray) -> np.ndarray:
    """raw: uint8 bytes in file order (multiple of 4) -> float32."""
    raw = np.ascontiguousarray(raw, np.uint8)
    out = np.empty(raw.size // 4, np.float32)
    lib().orc_ibm_block(_vpn(raw), out.size, _vpn(out))
    return out


def alias_filter(cfg, ns: int, dt: float):
    aa = np.zeros(ns, np.float32)
    p = params(cfg)
    tm = lib().orc_alias_filter(C.byref(p), ns, dt, _vpn(aa))
    return tm, aa


def trace_prep(tr: np.ndarray, dt: float, tm: int, fp64: bool = False) -> np.ndarray:
    tr = np.ascontiguousarray(tr, np.float32)
    ns = tr.size
    if fp64:
        P = np.zeros(ns + 2 * tm, np.float64)
        lib().orc_trace_prep_fp64(_vpn(tr), ns, dt, tm, _vpn(P))
    else:
        P = np.zeros(ns + 2 * tm, np.float32)
        lib().orc_trace_prep_fp32(_vpn(tr), ns, dt, tm, _vpn(P))
    return P


def filtered_trace(P: np.ndarray, tm: int, ns: int) -> np.ndarray:
    """F[T] = -2 P[T+tm] + P[T+tm-2] + P[T+tm+2] (Kirchhoff.c:1392), evaluated in
    double and rounded to the dtype of P."""
    Pd = P.astype(np.float64)
    T = np.arange(ns)
    F = -2.0 * Pd[T + tm] + Pd[T + tm - 2] + Pd[T + tm + 2]
    return F.astype(P.dtype)


def trace_geometry(cfg, coords: np.ndarray):
    """(geo float32 [ntr][6], off int32 [ntr]) for header words [ntr][4]."""
    p, g = params(cfg), grid(cfg)
    coords = np.ascontiguousarray(coords, np.int32)
    geo = np.zeros((coords.shape[0], 6), np.float32)
    off = np.zeros(coords.shape[0], np.int32)
    for i in range(coords.shape[0]):
        off[i] = lib().orc_trace_geometry(C.byref(p), C.byref(g), _vpn(coords[i]), _vpn(geo[i]))
    return geo, off


def migrate(cfg, ns, dt, traces, coords, vel, mode=FP32, nthreads=None, dtm=None,
            refflags=False, native=False):
    """All jobs into the padded image [nx*jobs_x][ny*jobs_y][nh][ns].
    Returns (km, fold, contributions)."""
    p, g = params(cfg), grid(cfg)
    traces = np.ascontiguousarray(traces, np.float32)
    coords = np.ascontiguousarray(coords, np.int32)
    vel = np.ascontiguousarray(vel, np.float32)
    shape = (g.nx * cfg.jobs_x, g.ny * cfg.jobs_y, cfg.nh, ns)
    dtype = np.float32 if mode == FP32 else np.float64
    km = np.zeros(shape, dtype)
    fold = np.zeros(shape, dtype)
    nthreads = nthreads or os.cpu_count() or 1
    n = lib(refflags, native).orc_migrate_all(C.byref(p), ns, dt, vel.shape[2], dt if dtm is None else dtm,
                                      _vpn(traces), _vpn(coords), traces.shape[0], _vpn(vel), mode,
                                      _vpn(km), _vpn(fold), nthreads)
    if n < 0:
        raise ValueError("oracle: velocity index below the model")
    return km, fold, int(n)


def migrate_job(cfg, ns, dt, job, traces, coords, vel, mode=FP32, dtm=None):
    p, g = params(cfg), grid(cfg)
    traces = np.ascontiguousarray(traces, np.float32)
    coords = np.ascontiguousarray(coords, np.int32)
    vel = np.ascontiguousarray(vel, np.float32)
    shape = (g.nx, g.ny, cfg.nh, ns)
    dtype = np.float32 if mode == FP32 else np.float64
    km = np.zeros(shape, dtype)
    fold = np.zeros(shape, dtype)
    n = lib().orc_migrate_job(C.byref(p), ns, dt, vel.shape[2], dt if dtm is None else dtm, job,
                              _vpn(traces), _vpn(coords), traces.shape[0], _vpn(vel), mode,
                              _vpn(km), _vpn(fold))
    if n < 0:
        raise ValueError("oracle: bad job or velocity index")
    return km, fold, int(n)


def poststack(km, fold, dt, anti_alias):
    """Kirchhoff.c:1423-1479 per trace (last axis); float32 or float64 inputs."""
    ns = km.shape[-1]
    aa = np.ascontiguousarray(anti_alias, np.float32)
    if km.dtype == np.float32:
        k = np.ascontiguousarray(km.reshape(-1, ns))
        f = np.ascontiguousarray(fold.reshape(-1, ns), np.float32)
        out = np.zeros_like(k)
        lib().orc_poststack_fp32(ns, dt, _vpn(aa), _vpn(k), _vpn(f), k.shape[0], _vpn(out))
    else:
        k = np.ascontiguousarray(km.reshape(-1, ns), np.float64)
        f = np.ascontiguousarray(fold.reshape(-1, ns), np.float64)
        out = np.zeros_like(k)
        lib().orc_poststack_fp64(ns, dt, _vpn(aa), _vpn(k), _vpn(f), k.shape[0], _vpn(out))
    return out.reshape(km.shape)


def segy_trace_header(cfg, ns, dt, job, xc, yc, ih, tracl, mess) -> bytes:
    """One image trace header; mess is 0 for the first record of a job file
    and 9999 for every later one (the reference's carried local)."""
    p, g = params(cfg), grid(cfg)
    out = np.zeros(240, np.uint8)
    lib().orc_segy_trace_header(C.byref(p), C.byref(g), ns, dt, job, xc, yc, ih, tracl, mess,
                                _vpn(out))
    return out.tobytes()


def segy_job_headers(cfg, ns, dt, job) -> np.ndarray:
    """[nx*ny*nh][240] uint8: the headers of image<job>.su in write order."""
    p, g = params(cfg), grid(cfg)
    out = np.zeros((g.nx * g.ny * cfg.nh, 240), np.uint8)
    lib().orc_segy_job_headers(C.byref(p), C.byref(g), ns, dt, job, _vpn(out))
    return out


def segy_file_header(cfg, ns, dt, fix_count=0) -> bytes:
    p, g = params(cfg), grid(cfg)
    out = np.zeros(3600, np.uint8)
    lib().orc_segy_file_header(C.byref(p), C.byref(g), ns, dt, fix_count, _vpn(out))
    return out.tobytes()


def count_size(head: bytes, length: int):
    b = np.frombuffer(head[:3720].ljust(3720, b"\0"), np.uint8).copy()
    ns, dt, nt = C.c_int(), C.c_float(), C.c_int()
    rc = lib().orc_count_size(_vpn(b), length, C.byref(ns), C.byref(dt), C.byref(nt))
    if rc != 0:
        raise ValueError("file too short")
    return ns.value, dt.value, nt.value


def column_coords(cfg, n: int, axis: int = 0) -> np.ndarray:
    """Image column coordinates IX (axis 0) or IY (axis 1) as float32,
    Kirchhoff.c:1346-1347 (origin dropped when image_origin_zero)."""
    start = 0.0 if cfg.image_origin_zero else float(cfg.start_xline if axis else cfg.start_iline)
    inc = cfg.inc_xline if axis else cfg.inc_iline
    d = cfg.xline_distance if axis else cfg.iline_distance
    return np.array([(start + c * inc) * d for c in range(n)], np.float64).astype(np.float32)


def filtered_traces_fp64(traces: np.ndarray, dt: float, tm: int = 2) -> np.ndarray:
    """F[ntr][ns] of every trace in FP64 (the double integral of
    Kirchhoff.c:1312-1328 and the taps of 1392 with every float widened)."""
    ns = traces.shape[1]
    return np.stack([filtered_trace(trace_prep(tr, dt, tm, fp64=True), tm, ns) for tr in traces])


def vgrad_velocity(ns: int, nxp: int, V: np.ndarray) -> np.ndarray:
    V = np.ascontiguousarray(V, np.float64)
    v = np.zeros((nxp, ns), np.float64)
    lib().orc_vgrad_velocity(ns, nxp, V.shape[0], V.shape[1], _vpn(V), _vpn(v))
    return v


def vgrad(cfg, ns, dt, F, coords, V, cig=False, grad=True):
    """(J, dJdV or None, cig or None) of the differentiable offset-class
    migration (FP64).  F: [ntr][ns] filtered traces; V: [nt_nodes][nx_nodes]."""
    g = grid(cfg)
    nxp = g.nx * cfg.jobs_x
    if g.ny * cfg.jobs_y != 1:
        raise ValueError("vgrad: lines only")
    geo, off = trace_geometry(cfg, coords)
    colx = column_coords(cfg, nxp, 0)
    coly = column_coords(cfg, 1, 1)[0]
    F = np.ascontiguousarray(F, np.float64)
    V = np.ascontiguousarray(V, np.float64)
    c = np.zeros((nxp, cfg.nh, ns), np.float64) if cig else None
    d = np.zeros(V.shape, np.float64) if grad else None
    J = lib().orc_vgrad(ns, float(np.float32(dt)), float(cfg.aperture), cfg.nh, nxp, _vpn(colx),
                        float(coly), _vpn(F), _vpn(geo), _vpn(off), F.shape[0], V.shape[0],
                        V.shape[1], _vpn(V), None if c is None else _vpn(c),
                        None if d is None else _vpn(d))
    return J, d, c


# ---------------------------------------------------------------- EOM ----
class OrcEomParams(C.Structure):
    _fields_ = [("ox", C.c_double), ("oy", C.c_double), ("angle_e_to_n", C.c_double),
                ("oxcsp", C.c_double), ("oycsp", C.c_double), ("excsp", C.c_double),
                ("eycsp", C.c_double), ("scale_xy", C.c_double), ("dcsp_x", C.c_double),
                ("dcsp_y", C.c_double), ("dh", C.c_double), ("aperture", C.c_double),
                ("low_pass_cutoff", C.c_double), ("nh", C.c_int32), ("oh", C.c_int32),
                ("jobs_x", C.c_int32), ("jobs_y", C.c_int32)]


class OrcEomGrid(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("no_job", C.c_int32), ("sinA", C.c_float), ("cosA", C.c_float),
                ("start_x", C.c_float), ("start_y", C.c_float)]


_PROTOS.update({
    "orc_eom_grid_init": (None, [C.POINTER(OrcEomParams), C.POINTER(OrcEomGrid)]),
    "orc_eom_filters": (None, [_i32, _f32, C.c_double, _vp, _vp]),
    "orc_eom_job": (_i64, [C.POINTER(OrcEomParams), _i32, _f32, _i32, _vp, _vp, _i64, _vp, _vp,
                           _vp]),
    "orc_eom_post": (None, [_i32, _f32, _vp, _vp, _vp, _vp, _i64, _vp]),
    "orc_eom_trace_header": (None, [C.POINTER(OrcEomParams), _i32, _f32, _i32, _i32, _i32, _i32,
                                    _i32, _i32, _vp]),
    "orc_eom_job_records": (_i64, [C.POINTER(OrcEomParams), _i32, _f32, _i32, _vp, _vp]),
    "orc_eom_file_header": (None, [_i32, _f32, _vp]),
})
_libs.clear()


def eom_params(cfg) -> OrcEomParams:
    p = OrcEomParams()
    for name, _ in OrcEomParams._fields_:
        setattr(p, name, getattr(cfg, name))
    return p


def eom_grid(cfg) -> OrcEomGrid:
    g = OrcEomGrid()
    p = eom_params(cfg)
    lib().orc_eom_grid_init(C.byref(p), C.byref(g))
    return g


def eom_filters(cfg, ns: int, dt: float):
    lp = np.zeros(ns, np.float32)
    rr = np.zeros(ns, np.float32)
    lib().orc_eom_filters(ns, dt, float(cfg.low_pass_cutoff), _vpn(lp), _vpn(rr))
    return lp, rr


def eom_job(cfg, ns, dt, job, traces, coords, vel, native=False):
    """(csp, fold, evaluated) of one job slab [nx][ny][nh][ns]."""
    p, g = eom_params(cfg), eom_grid(cfg)
    traces = np.ascontiguousarray(traces, np.float32)
    coords = np.ascontiguousarray(coords, np.int32)
    vel = np.ascontiguousarray(vel, np.float32)
    shape = (g.nx, g.ny, cfg.nh, ns)
    csp = np.zeros(shape, np.float32)
    fold = np.zeros(shape, np.float32)
    n = lib(native=native).orc_eom_job(C.byref(p), ns, dt, job, _vpn(traces), _vpn(coords),
                                       traces.shape[0], _vpn(vel), _vpn(csp), _vpn(fold))
    if n < 0:
        raise ValueError("oracle: bad job")
    return csp, fold, int(n)


def eom_all(cfg, ns, dt, traces, coords, vel, nthreads=1, native=False):
    """All jobs into the padded gather volume [nx*jobs_x][ny*jobs_y][nh][ns],
    the jobs on nthreads threads (the reference's one rank per slab; ctypes
    releases the GIL during each job)."""
    from concurrent.futures import ThreadPoolExecutor
    g = eom_grid(cfg)
    out_c = np.zeros((g.nx * cfg.jobs_x, g.ny * cfg.jobs_y, cfg.nh, ns), np.float32)
    out_f = np.zeros_like(out_c)
    with ThreadPoolExecutor(max_workers=max(1, nthreads)) as ex:
        res = list(ex.map(lambda job: eom_job(cfg, ns, dt, job, traces, coords, vel, native),
                          range(g.no_job)))
    total = 0
    for job, (c, f, n) in enumerate(res):
        xs, ys = job // cfg.jobs_y, job % cfg.jobs_y
        out_c[xs * g.nx:(xs + 1) * g.nx, ys * g.ny:(ys + 1) * g.ny] = c
        out_f[xs * g.nx:(xs + 1) * g.nx, ys * g.ny:(ys + 1) * g.ny] = f
        total += n
    return out_c, out_f, total


def eom_post(csp, fold, dt, lowpass, rootrho):
    ns = csp.shape[-1]
    c = np.ascontiguousarray(csp.reshape(-1, ns), np.float32)
    f = np.ascontiguousarray(fold.reshape(-1, ns), np.float32)
    out = np.zeros_like(c)
    lib().orc_eom_post(ns, dt, _vpn(np.ascontiguousarray(lowpass, np.float32)),
                       _vpn(np.ascontiguousarray(rootrho, np.float32)), _vpn(c), _vpn(f),
                       c.shape[0], _vpn(out))
    return out.reshape(csp.shape)


def eom_trace_header(cfg, ns, dt, job, xc, yc, ih, it, mess) -> bytes:
    p = eom_params(cfg)
    out = np.zeros(240, np.uint8)
    lib().orc_eom_trace_header(C.byref(p), ns, dt, job, xc, yc, ih, it, mess, _vpn(out))
    return out.tobytes()


def eom_job_records(cfg, ns, dt, job, post) -> bytes:
    """The bytes of image<job>.su from the job's post-filtered gathers
    [nx][ny][nh][ns] (only the records with it < Nx*Ny*nh + 1)."""
    p = eom_params(cfg)
    post = np.ascontiguousarray(post, np.float32)
    out = np.zeros(post.size // ns * (240 + 4 * ns), np.uint8)
    n = lib().orc_eom_job_records(C.byref(p), ns, dt, job, _vpn(post), _vpn(out))
    if n < 0:
        raise ValueError("oracle: bad job")
    return out[:n * (240 + 4 * ns)].tobytes()


def eom_file_header(ns, dt) -> bytes:
    out = np.zeros(3600, np.uint8)
    lib().orc_eom_file_header(ns, dt, _vpn(out))
    return out.tobytes()
