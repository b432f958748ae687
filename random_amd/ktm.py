"""Host-side mirror of the reference's migration interface.

``Config`` carries the macro block of Kirchhoff.c:17-65 (same names as the
C struct ``ktm_config``).  ``Plan`` owns the device-resident setup of one
configuration (velocity model, traveltime and column tables; what main()
computes before its job loop, Kirchhoff.c:263-331) and drives the kernels on
torch tensors: ``prep`` (Kirchhoff.c:1296-1328), ``migrate_block`` (the
image loop, Kirchhoff.c:1332-1404) and ``poststack`` (Kirchhoff.c:1423-1479).
``migrate_host`` is the one-call host-array path (``ktm_migrate``).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import (ACCUM_KAHAN, ACCUM_PLAIN, INDEX_FAST, PREP_REFERENCE, SEGY_IBM, TRACE_F32, KtmConfig,
                   KtmImage, KtmSetup, KtmStats, KtmTraces, KtmVelocity, check, lib)


@dataclass
class Config:
    """The configuration macros of Kirchhoff.c:17-65 (reference values by default)."""

    start_iline: int = 1500
    end_iline: int = 1600
    start_xline: int = 2500
    end_xline: int = 2600
    inc_iline: float = 1.0
    inc_xline: float = 1.0
    dh: float = 100.0
    nh: int = 50
    oh: float = 200.0
    ox: float = 444479.29
    oy: float = 478003.98
    angle_e_to_n: float = -10.3690
    iline_distance: float = 25.0
    xline_distance: float = 12.5
    aperture: float = 3000.0
    anti_alias_cutoff: float = 50.0
    nvel_ilines: int = 472
    nvel_xlines: int = 643
    dvel_ilines: int = 4
    dvel_xlines: int = 8
    ovel_ilines: int = 852
    ovel_xlines: int = 752
    ntblock: int = 50000
    jobs_x: int = 36
    jobs_y: int = 1
    image_origin_zero: int = 0   # Kirchhoff_GHTest1.c:1346-1348,1365-1366
    vel_lower_clamp: int = 0     # Kirchhoff_GHTest1.c:1395-1399
    scale_xy: float = 0.0        # Kirchhoff_GHTest1.c:56 (0 = Kirchhoff.c form)

    def to_c(self) -> KtmConfig:
        c = KtmConfig()
        for f in dataclasses.fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)


def setup_info(cfg: Config, ns: int, dt: float, nsm: int | None = None, dtm: float | None = None):
    """ktm_setup_init: (setup dict, anti_alias float32[ns]) without a device."""
    nsm = ns if nsm is None else nsm
    dtm = dt if dtm is None else dtm
    s = KtmSetup()
    aa = np.zeros(ns, np.float32)
    c = cfg.to_c()
    check(lib().ktm_setup_init(C.byref(c), ns, dt, nsm, dtm, C.byref(s),
                               aa.ctypes.data_as(C.c_void_p)), "ktm_setup_init")
    return {f: getattr(s, f) for f, _ in KtmSetup._fields_ if f != "reserved"}, aa


def _ptr(t) -> int | None:
    return None if t is None else t.data_ptr()


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class Plan:
    """Device-resident setup of one configuration (``ktm_plan``).

    Every call is issued on torch's current stream; the plan's workspaces
    (16-bit decode buffer, demigration sort keys) are reused without stream
    ordering, so all calls on one plan must run on a single stream."""

    def __init__(self, cfg: Config, ns: int, dt: float, vel: np.ndarray, dtm: float | None = None):
        import torch  # noqa: F401  (device context comes from torch)
        vel = np.ascontiguousarray(vel, dtype=np.float32)
        if vel.ndim != 3 or vel.shape[0] != cfg.nvel_ilines or vel.shape[1] != cfg.nvel_xlines:
            raise ValueError(f"velocity must be [nvel_ilines][nvel_xlines][nsm], got {vel.shape}")
        self.cfg = cfg
        self.ns, self.dt = int(ns), float(np.float32(dt))
        self.nsm = int(vel.shape[2])
        self.dtm = self.dt if dtm is None else float(np.float32(dtm))
        self._c = cfg.to_c()
        h = C.c_void_p()
        check(lib().ktm_plan_create(C.byref(h), C.byref(self._c), self.ns, self.dt, self.nsm,
                                    self.dtm, vel.ctypes.data_as(C.c_void_p)), "ktm_plan_create")
        self._h = h
        s = KtmSetup()
        check(lib().ktm_plan_setup(self._h, C.byref(s)), "ktm_plan_setup")
        self.setup = {f: getattr(s, f) for f, _ in KtmSetup._fields_ if f != "reserved"}
        self.row = int(lib().ktm_plan_trace_row(self._h))
        self.nxp, self.nyp, self.nh = self.setup["nxp"], self.setup["nyp"], self.setup["nh"]
        self.image_shape = (self.nxp, self.nyp, self.nh, self.ns)

    # -- lifetime ------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().ktm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers -------------------------------------------------------------
    def new_image(self, kahan: bool = False, device="cuda"):
        """km = 0, fold = 1 (and comp = 0) on the device (Kirchhoff.c:1262-1274)."""
        import torch
        km = torch.empty(self.image_shape, dtype=torch.float32, device=device)
        fold = torch.empty_like(km)
        comp = torch.empty_like(km) if kahan else None
        self.reset(km, fold, comp)
        return km, fold, comp

    def reset(self, km, fold, comp=None) -> None:
        check(lib().ktm_image_reset(self._h, _ptr(km), _ptr(fold), _ptr(comp), _stream()),
              "ktm_image_reset")

    # -- kernels -------------------------------------------------------------
    def geometry(self, coords):
        import torch
        ntr = coords.shape[0]
        geom = torch.empty((ntr, 8), dtype=torch.float32, device=coords.device)
        check(lib().ktm_geometry(self._h, _ptr(coords), ntr, _ptr(geom), _stream()), "ktm_geometry")
        return geom

    def prep(self, traces, coords, mode: int = PREP_REFERENCE, dtype: int = TRACE_F32):
        """Prepared traces + geometry for ``migrate_block`` (Kirchhoff.c:1296-1328)."""
        import torch
        ntr = traces.shape[0]
        assert traces.dtype == torch.float32 and traces.is_contiguous() and traces.shape[1] == self.ns
        assert coords.dtype == torch.int32 and coords.is_contiguous() and coords.shape == (ntr, 4)
        et = torch.float32 if dtype == TRACE_F32 else torch.int16
        prepared = torch.empty((ntr, self.row), dtype=et, device=traces.device)
        geom = torch.empty((ntr, 8), dtype=torch.float32, device=traces.device)
        check(lib().ktm_prep(self._h, _ptr(traces), _ptr(coords), ntr, mode, dtype, _ptr(prepared),
                             _ptr(geom), _stream()), "ktm_prep")
        return prepared, geom

    def ingest(self, raw, ntr: int, fmt: int = SEGY_IBM, mode: int = PREP_REFERENCE,
               dtype: int = TRACE_F32):
        """Prepared traces, geometry and header words from ``ntr`` raw SEG-Y
        trace records (uint8 device tensor, 240 + 4 ns bytes each, as stored
        in the file), decoded on the device (``ktm_ingest``)."""
        import torch
        rec = 240 + 4 * self.ns
        assert raw.dtype == torch.uint8 and raw.is_contiguous() and raw.numel() >= ntr * rec
        et = torch.float32 if dtype == TRACE_F32 else torch.int16
        prepared = torch.empty((ntr, self.row), dtype=et, device=raw.device)
        geom = torch.empty((ntr, 8), dtype=torch.float32, device=raw.device)
        coords = torch.empty((ntr, 4), dtype=torch.int32, device=raw.device)
        check(lib().ktm_ingest(self._h, _ptr(raw), ntr, fmt, mode, dtype, _ptr(prepared),
                               _ptr(geom), _ptr(coords), _stream()), "ktm_ingest")
        return prepared, geom, coords

    def migrate_block(self, prepared, geom, km, fold, comp=None, *, dtype: int = TRACE_F32,
                      mode: int = PREP_REFERENCE, accum: int = ACCUM_PLAIN, job: int = -1,
                      counter=None) -> None:
        """Accumulate one block of prepared traces into (km, fold[, comp])."""
        ntr = geom.shape[0]
        check(lib().ktm_migrate_block(self._h, _ptr(prepared), dtype, mode, _ptr(geom), ntr, accum,
                                      job, _ptr(km), _ptr(fold), _ptr(comp), _ptr(counter),
                                      _stream()), "ktm_migrate_block")

    def poststack(self, km, fold):
        """Derivative, fold scaling and anti-alias low-pass (Kirchhoff.c:1423-1479)."""
        import torch
        out = torch.empty_like(km)
        check(lib().ktm_poststack(self._h, _ptr(km), _ptr(fold), _ptr(out), _stream()),
              "ktm_poststack")
        return out

    # -- adjoint pair (CLOSED prep) ------------------------------------------
    def migrate_lin(self, traces, coords, *, kahan: bool = False, fast: bool = False):
        """km = L^T d: the linear migration operator (CLOSED filter) on a zeroed
        image; returns (km, fold).  ``fast`` takes the relaxed traveltime
        index (INDEX_FAST) instead of the reference's."""
        from ._lib import PREP_CLOSED
        prepared, geom = self.prep(traces, coords, PREP_CLOSED, TRACE_F32)
        km, fold, comp = self.new_image(kahan=kahan, device=traces.device)
        accum = (ACCUM_KAHAN if kahan else ACCUM_PLAIN) | (INDEX_FAST if fast else 0)
        self.migrate_block(prepared, geom, km, fold, comp, mode=PREP_CLOSED, accum=accum)
        return km, fold

    def demigrate(self, km, geom, *, fast: bool = False):
        """d = L km: the exact transpose of ``migrate_lin`` with the same
        ``fast`` (modelling)."""
        import torch
        ntr = geom.shape[0]
        out = torch.empty((ntr, self.ns), dtype=torch.float32, device=km.device)
        check(lib().ktm_demigrate_index(self._h, _ptr(km.contiguous()), _ptr(geom), ntr,
                                        _ptr(out), INDEX_FAST if fast else 0, _stream()),
              "ktm_demigrate_index")
        return out

    def poststack_adjoint(self, image, fold):
        """km = S^T image, the transpose of ``poststack`` for a fixed fold."""
        import torch
        out = torch.empty_like(image)
        check(lib().ktm_poststack_adjoint(self._h, _ptr(image.contiguous()), _ptr(fold), _ptr(out),
                                          _stream()), "ktm_poststack_adjoint")
        return out

    def migrate(self, traces, coords, *, block: int | None = None, mode: int = PREP_REFERENCE,
                dtype: int = TRACE_F32, accum: int = ACCUM_PLAIN, job: int = -1):
        """All traces in blocks of ``ntblock``: returns (km, fold, contributions)."""
        import torch
        block = block or self.cfg.ntblock
        km, fold, comp = self.new_image(kahan=(accum & ~INDEX_FAST) == ACCUM_KAHAN,
                                        device=traces.device)
        counter = torch.zeros(1, dtype=torch.int64, device=traces.device)
        for i0 in range(0, traces.shape[0], block):
            prepared, geom = self.prep(traces[i0:i0 + block], coords[i0:i0 + block], mode, dtype)
            self.migrate_block(prepared, geom, km, fold, comp, dtype=dtype, mode=mode, accum=accum,
                               job=job, counter=counter)
        return km, fold, int(counter.item())


def segy_decode(raw, ntr: int, ns: int, fmt: int = SEGY_IBM):
    """(traces float32 [ntr][ns], header words int32 [ntr][4]) of raw SEG-Y
    trace records on the device (``ktm_segy_decode``)."""
    import torch
    traces = torch.empty((ntr, ns), dtype=torch.float32, device=raw.device)
    coords = torch.empty((ntr, 4), dtype=torch.int32, device=raw.device)
    check(lib().ktm_segy_decode(_ptr(raw), ntr, ns, fmt, _ptr(traces), _ptr(coords), _stream()),
          "ktm_segy_decode")
    return traces, coords


def migrate_host(cfg: Config, ns: int, dt: float, traces: np.ndarray, coords: np.ndarray,
                 vel: np.ndarray, dtm: float | None = None):
    """``ktm_migrate`` on host arrays: returns (km, fold, image, stats dict)."""
    traces = np.ascontiguousarray(traces, np.float32)
    coords = np.ascontiguousarray(coords, np.int32)
    vel = np.ascontiguousarray(vel, np.float32)
    info, _ = setup_info(cfg, ns, dt, vel.shape[2], dtm)
    shape = (info["nxp"], info["nyp"], info["nh"], ns)
    km = np.empty(shape, np.float32)
    fold = np.empty(shape, np.float32)
    image = np.empty(shape, np.float32)
    st = KtmStats()
    c = cfg.to_c()
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    tr = KtmTraces(vp(traces), vp(coords), traces.shape[0], ns, dt)
    vm = KtmVelocity(vp(vel), vel.shape[2], dt if dtm is None else dtm)
    out = KtmImage(vp(km), vp(fold), vp(image))
    check(lib().ktm_migrate(C.byref(c), C.byref(tr), C.byref(vm), C.byref(out), C.byref(st)),
          "ktm_migrate")
    return km, fold, image, {"contributions": st.contributions, "migrate_ms": st.migrate_ms,
                             "total_ms": st.total_ms}


__all__ = ["Config", "Plan", "migrate_host", "segy_decode", "setup_info", "ACCUM_KAHAN", "_lib"]
