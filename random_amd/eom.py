"""Equivalent-offset migration (EOM.c, ``csp(job_id)``): common-scatter-point
gathers on the same engine.

``EomConfig`` carries the macro block of EOM.c:60-99 (same names as the C
struct ``ktm_eom_config``).  For every trace and every CSP within the aperture,
each output sample t >= talpha is binned by the equivalent offset
he^2 = x^2 + h^2 - 4 x^2 h^2 / (t dt v)^2 (EOM.c:1285-1286) into two adjacent
offset bins with linear weights and a time taper (EOM.c:1287-1306); the
gathers are then scaled by t dt / fold and filtered by the low-pass and
(1+i) sqrt(rho) spectra (EOM.c:1386-1413).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass

import numpy as np

from ._lib import KtmEomConfig, KtmEomSetup, check, lib


@dataclass
class EomConfig:
    """The configuration macros of EOM.c:60-99 (reference values by default)."""

    ox: float = 789868.0             # EOM.c:66
    oy: float = 265844.0             # EOM.c:67
    angle_e_to_n: float = 0.3        # EOM.c:68
    oxcsp: float = 789868.0          # EOM.c:73
    oycsp: float = 265844.0          # EOM.c:74
    excsp: float = 819945.0          # EOM.c:75
    eycsp: float = 265999.0          # EOM.c:76
    scale_xy: float = 0.001          # EOM.c:77
    dcsp_x: float = 12.5             # EOM.c:79
    dcsp_y: float = 0.0              # EOM.c:80
    dh: float = 6.25                 # EOM.c:81
    aperture: float = 6000.0         # EOM.c:85
    low_pass_cutoff: float = 50.0    # EOM.c:86
    nh: int = 1000                   # EOM.c:82
    oh: int = 0                      # EOM.c:99 (an integer macro)
    jobs_x: int = 36                 # EOM.c:91
    jobs_y: int = 1                  # EOM.c:92
    ntblock: int = 20000             # EOM.c:89
    velocity_start: float = 3000.0   # EOM.c:61
    velocity_end: float = 3000.0     # EOM.c:62

    def replace(self, **kw) -> "EomConfig":
        return dataclasses.replace(self, **kw)

    def to_c(self) -> KtmEomConfig:
        c = KtmEomConfig()
        for f in dataclasses.fields(self):
            setattr(c, f.name, getattr(self, f.name))
        return c

    def velocity(self, ns: int):
        """vel_t of EOM.c:871-875: velocityStart + (End - Start) * (i/ns) with
        the integer division of the source, i.e. velocityStart for every
        sample (i < ns)."""
        return np.array([self.velocity_start + (self.velocity_end - self.velocity_start) * (i // ns)
                         for i in range(ns)], np.float32)


def eom_setup(cfg: EomConfig, ns: int, dt: float):
    """(setup dict, lowpass float32[ns], rootrho float32[ns]) without a device."""
    s = KtmEomSetup()
    lp = np.zeros(ns, np.float32)
    rr = np.zeros(ns, np.float32)
    c = cfg.to_c()
    check(lib().ktm_eom_setup_init(C.byref(c), ns, dt, C.byref(s), lp.ctypes.data_as(C.c_void_p),
                                   rr.ctypes.data_as(C.c_void_p)), "ktm_eom_setup_init")
    return {f: getattr(s, f) for f, _ in KtmEomSetup._fields_}, lp, rr


class EomPlan:
    """Device-resident setup of one EOM configuration (``ktm_eom_plan``)."""

    def __init__(self, cfg: EomConfig, ns: int, dt: float, vel=None):
        import torch  # noqa: F401  (device context comes from torch)
        self.cfg = cfg
        self.ns, self.dt = int(ns), float(np.float32(dt))
        vel = cfg.velocity(ns) if vel is None else np.ascontiguousarray(vel, np.float32)
        if vel.shape != (self.ns,):
            raise ValueError("vel must have ns samples")
        self._c = cfg.to_c()
        h = C.c_void_p()
        check(lib().ktm_eom_plan_create(C.byref(h), C.byref(self._c), self.ns, self.dt,
                                        vel.ctypes.data_as(C.c_void_p)), "ktm_eom_plan_create")
        self._h = h
        s = KtmEomSetup()
        check(lib().ktm_eom_plan_setup(self._h, C.byref(s)), "ktm_eom_plan_setup")
        self.setup = {f: getattr(s, f) for f, _ in KtmEomSetup._fields_}
        self.shape = (s.nxp, s.nyp, s.nh, self.ns)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().ktm_eom_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _stream():
        import torch
        return torch.cuda.current_stream().cuda_stream

    def new_gathers(self, device="cuda"):
        """csp = 0, fold = 1 on the device (EOM.c:1202-1211)."""
        import torch
        csp = torch.empty(self.shape, dtype=torch.float32, device=device)
        fold = torch.empty_like(csp)
        check(lib().ktm_eom_reset(self._h, csp.data_ptr(), fold.data_ptr(), self._stream()),
              "ktm_eom_reset")
        return csp, fold

    def block(self, traces, coords, csp, fold, job: int = -1, counter=None) -> None:
        """Bin one block of traces (device tensors) into the gathers."""
        import torch
        assert traces.dtype == torch.float32 and traces.is_contiguous() and traces.shape[1] == self.ns
        assert coords.dtype == torch.int32 and coords.is_contiguous() and coords.shape == (traces.shape[0], 4)
        check(lib().ktm_eom_block(self._h, traces.data_ptr(), coords.data_ptr(), traces.shape[0],
                                  job, csp.data_ptr(), fold.data_ptr(),
                                  None if counter is None else counter.data_ptr(), self._stream()),
              "ktm_eom_block")

    def migrate(self, traces, coords, block: int | None = None, job: int = -1):
        """All traces in blocks of ntblock: (csp, fold, evaluated samples)."""
        import torch
        block = block or self.cfg.ntblock
        csp, fold = self.new_gathers(traces.device)
        counter = torch.zeros(1, dtype=torch.int64, device=traces.device)
        for i0 in range(0, traces.shape[0], block):
            self.block(traces[i0:i0 + block], coords[i0:i0 + block], csp, fold, job, counter)
        return csp, fold, int(counter.item())

    def post(self, csp, fold):
        """The scaled, low-passed, sqrt-rho filtered gathers (EOM.c:1386-1413)."""
        import torch
        out = torch.empty_like(csp)
        check(lib().ktm_eom_post(self._h, csp.data_ptr(), fold.data_ptr(), out.data_ptr(),
                                 self._stream()), "ktm_eom_post")
        return out

    # ---- one job's slab only, [nx][ny][nh][ns] (csp(job_id)'s arrays) ----
    @property
    def job_shape(self):
        return (self.setup["nx"], self.setup["ny"], self.setup["nh"], self.ns)

    def new_job_gathers(self, device="cuda"):
        import torch
        csp = torch.empty(self.job_shape, dtype=torch.float32, device=device)
        fold = torch.empty_like(csp)
        check(lib().ktm_eom_job_reset(self._h, csp.data_ptr(), fold.data_ptr(), self._stream()),
              "ktm_eom_job_reset")
        return csp, fold

    def job_migrate(self, traces, coords, job: int, block: int | None = None):
        """(csp, fold, evaluated samples) of one job's slab."""
        import torch
        block = block or self.cfg.ntblock
        csp, fold = self.new_job_gathers(traces.device)
        counter = torch.zeros(1, dtype=torch.int64, device=traces.device)
        for i0 in range(0, traces.shape[0], block):
            t, c = traces[i0:i0 + block], coords[i0:i0 + block]
            assert t.dtype == torch.float32 and t.is_contiguous() and t.shape[1] == self.ns
            assert c.dtype == torch.int32 and c.is_contiguous() and c.shape == (t.shape[0], 4)
            check(lib().ktm_eom_job_block(self._h, t.data_ptr(), c.data_ptr(), t.shape[0], job,
                                          csp.data_ptr(), fold.data_ptr(), counter.data_ptr(),
                                          self._stream()), "ktm_eom_job_block")
        return csp, fold, int(counter.item())

    def job_post(self, csp, fold):
        import torch
        assert tuple(csp.shape) == self.job_shape and tuple(fold.shape) == self.job_shape
        out = torch.empty_like(csp)
        check(lib().ktm_eom_job_post(self._h, csp.data_ptr(), fold.data_ptr(), out.data_ptr(),
                                     self._stream()), "ktm_eom_job_post")
        return out

    def job_records(self, post, job: int) -> bytes:
        """The bytes of image<job>.su from the job's post-filtered gathers."""
        import torch
        assert tuple(post.shape) == self.job_shape and post.dtype == torch.float32
        rec = 240 + 4 * self.ns
        nx, ny, nh, _ = self.job_shape
        buf = torch.empty(nx * ny * nh * rec, dtype=torch.uint8, device=post.device)
        n = C.c_int64()
        check(lib().ktm_eom_job_records(self._h, post.data_ptr(), job, buf.data_ptr(), C.byref(n),
                                        self._stream()), "ktm_eom_job_records")
        return buf[:n.value * rec].cpu().numpy().tobytes()
