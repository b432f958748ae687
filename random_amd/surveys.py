"""Seeded synthetic surveys of the benchmark configurations.

Every survey is a deterministic function of (configuration, seed): shot and
receiver header coordinates (integers, as stored in SEG-Y words 18-21), point
scatterers, an RMS velocity model on the image grid, and the configuration
macros.  Traces are sums of Ricker wavelets on the double-square-root
diffraction curve of each scatterer (the traveltime the migration inverts,
Kirchhoff.c:1381-1383) plus Gaussian noise; ``traces_numpy`` builds them on
the CPU for small surveys and ``random_amd.synth`` on the GPU for large ones.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .ktm import Config

DT = np.float32(0.004)
RICKER_F0 = 25.0


@dataclass
class Survey:
    name: str
    cfg: Config
    ns: int
    dt: float
    coords: np.ndarray          # int32 [ntr][4]: sx, sy, gx, gy header words
    scatterers: np.ndarray      # float32 [nscat][4]: x, y, t0, amplitude (grid coordinates)
    vel: np.ndarray             # float32 [nvel_il][nvel_xl][nsm]
    v0: float = 1500.0          # v(t, x) = v0 + vt*t + vx*x at a scatterer
    vt: float = 0.0
    vx: float = 0.0
    snr_db: float = 10.0
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def ntr(self) -> int:
        return int(self.coords.shape[0])

    def grid_coords(self) -> np.ndarray:
        """Source/receiver positions in the migration grid frame (float64),
        for angle 0: (word - origin/scale) * scale, as Kirchhoff_GHTest1.c:1302-1315
        (scale 1 reproduces Kirchhoff.c:1297-1304)."""
        c = self.coords.astype(np.float64)
        s = self.cfg.scale_xy if self.cfg.scale_xy > 0 else 1.0
        out = np.empty_like(c)
        out[:, 0::2] = (np.trunc(c[:, 0::2] - self.cfg.ox / s)) * s
        out[:, 1::2] = (np.trunc(c[:, 1::2] - self.cfg.oy / s)) * s
        return out


def _velocity_grid(cfg: Config, ns: int, dt: float, v0: float, vt: float, vx: float) -> np.ndarray:
    """v(t, IX) = v0 + vt * t + vx * IX on the image grid (velocity grid == image grid)."""
    t = np.arange(ns, dtype=np.float64) * float(dt)
    ix = (cfg.start_iline + np.arange(cfg.nvel_ilines) * cfg.inc_iline) * cfg.iline_distance
    v = v0 + vt * t[None, None, :] + vx * ix[:, None, None]
    return np.ascontiguousarray(np.broadcast_to(v, (cfg.nvel_ilines, cfg.nvel_xlines, ns)),
                                dtype=np.float32)


def line2d(nshots=40, nrecv_side=60, shot_dx=50.0, recv_dx=12.5, ns=1001, ncols=800,
           col_dx=6.25, aperture=2000.0, nscat=20, seed=20240001, jobs_x=1,
           nh=1, oh=0.0, dh=2000.0, x0=1522.0, t0_range=(0.3, 3.5)) -> Survey:
    """2-D split-spread line (by default 40 shots x 120 receivers, 1001 samples): shots every
    ``shot_dx`` m, ``nrecv_side`` receivers each side at ``recv_dx`` m,
    ``ncols`` image columns at ``col_dx`` m starting at 0, v = 1500 + 400 t.
    Coordinates are stored in centimetres with scale_xy = 0.01
    (Kirchhoff_GHTest1.c:56) so that the 12.5 m receiver spacing is exact."""
    rng = np.random.default_rng(seed)
    s = x0 + shot_dx * np.arange(nshots)
    j = np.concatenate([-np.arange(nrecv_side, 0, -1), np.arange(1, nrecv_side + 1)])
    sx = np.repeat(s, j.size)
    gx = sx + recv_dx * np.tile(j, nshots)
    coords = np.zeros((sx.size, 4), np.int32)
    coords[:, 0] = np.round(sx * 100.0).astype(np.int32)
    coords[:, 2] = np.round(gx * 100.0).astype(np.int32)
    cfg = Config(start_iline=0, end_iline=ncols, start_xline=0, end_xline=0, inc_iline=1.0,
                 inc_xline=1.0, dh=dh, nh=nh, oh=oh, ox=0.0, oy=0.0, angle_e_to_n=0.0,
                 iline_distance=col_dx, xline_distance=col_dx, aperture=aperture,
                 anti_alias_cutoff=50.0, nvel_ilines=ncols, nvel_xlines=1, dvel_ilines=1,
                 dvel_xlines=1, ovel_ilines=0, ovel_xlines=0, ntblock=50000, jobs_x=jobs_x,
                 jobs_y=1, scale_xy=0.01)
    cmp = 0.5 * (sx + gx)
    scat = np.zeros((nscat, 4), np.float32)
    scat[:, 0] = rng.uniform(cmp.min(), cmp.max(), nscat)
    scat[:, 1] = 0.0
    scat[:, 2] = rng.uniform(t0_range[0], t0_range[1], nscat)
    scat[:, 3] = rng.uniform(0.5, 1.0, nscat)
    vel = _velocity_grid(cfg, ns, DT, 1500.0, 400.0, 0.0)
    return Survey("line2d", cfg, ns, float(DT), coords, scat, vel, 1500.0, 400.0, 0.0, 10.0, seed)


def area3d(ntr=1_000_000, ns=1500, ncols=400, col_dx=25.0, extent=10_000.0, max_offset=4000.0,
           aperture=2500.0, nscat=200, seed=20240011, jobs_x=1, jobs_y=1) -> Survey:
    """3-D survey (by default 1e6 traces, 1500 samples, 400 x 400 image): shots uniform over an
    ``extent`` square, offsets U(0, max_offset), azimuths U(0, 2 pi); integer
    metre coordinates shifted by +5000 m with ox = oy = 5000 (header words stay
    positive); image ncols x ncols at col_dx from 0; v = 1500 + 350 t + 0.02 IX."""
    rng = np.random.default_rng(seed)
    s = rng.uniform(0.0, extent, (ntr, 2))
    off = rng.uniform(0.0, max_offset, ntr)
    az = rng.uniform(0.0, 2.0 * np.pi, ntr)
    g = s + off[:, None] * np.stack([np.cos(az), np.sin(az)], axis=1)
    shift = 5000.0
    coords = np.empty((ntr, 4), np.int32)
    coords[:, 0:2] = np.round(s + shift).astype(np.int32)
    coords[:, 2:4] = np.round(g + shift).astype(np.int32)
    cfg = Config(start_iline=0, end_iline=ncols, start_xline=0, end_xline=ncols, inc_iline=1.0,
                 inc_xline=1.0, dh=max_offset * 2.5, nh=1, oh=0.0, ox=shift, oy=shift,
                 angle_e_to_n=0.0, iline_distance=col_dx, xline_distance=col_dx,
                 aperture=aperture, anti_alias_cutoff=50.0, nvel_ilines=ncols, nvel_xlines=ncols,
                 dvel_ilines=1, dvel_xlines=1, ovel_ilines=0, ovel_xlines=0, ntblock=50000,
                 jobs_x=jobs_x, jobs_y=jobs_y, scale_xy=0.0)
    scat = np.zeros((nscat, 4), np.float32)
    scat[:, 0] = rng.uniform(0.0, ncols * col_dx, nscat)
    scat[:, 1] = rng.uniform(0.0, ncols * col_dx, nscat)
    scat[:, 2] = rng.uniform(0.3, 3.5, nscat)
    scat[:, 3] = rng.uniform(0.5, 1.0, nscat)
    vel = _velocity_grid(cfg, ns, DT, 1500.0, 350.0, 0.02)
    return Survey("area3d", cfg, ns, float(DT), coords, scat, vel, 1500.0, 350.0, 0.02, 10.0, seed)


def ricker(tau: np.ndarray, f0: float = RICKER_F0) -> np.ndarray:
    a = (np.pi * f0 * tau) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def diffraction_times(sv: Survey, rows=slice(None)) -> np.ndarray:
    """DSR traveltime of every (trace, scatterer): sqrt((t0/2)^2 + |s-x|^2/v^2) +
    sqrt((t0/2)^2 + |g-x|^2/v^2), v = v0 + vt t0 + vx x at the scatterer."""
    g = sv.grid_coords()[rows]
    x, y, t0 = (sv.scatterers[:, k].astype(np.float64) for k in range(3))
    v = sv.v0 + sv.vt * t0 + sv.vx * x
    ds2 = (g[:, 0:1] - x) ** 2 + (g[:, 1:2] - y) ** 2
    dg2 = (g[:, 2:3] - x) ** 2 + (g[:, 3:4] - y) ** 2
    h = (0.5 * t0) ** 2
    return np.sqrt(h + ds2 / v ** 2) + np.sqrt(h + dg2 / v ** 2)


def traces_numpy(sv: Survey, noise: bool = True, chunk: int = 512) -> np.ndarray:
    """Signal (+ seeded Gaussian noise at sv.snr_db) on the CPU, float32 [ntr][ns]."""
    t = np.arange(sv.ns, dtype=np.float64) * sv.dt
    out = np.zeros((sv.ntr, sv.ns), np.float32)
    amp = sv.scatterers[:, 3].astype(np.float64)
    for i0 in range(0, sv.ntr, chunk):
        tau = diffraction_times(sv, slice(i0, i0 + chunk))
        w = ricker(t[None, None, :] - tau[:, :, None])
        out[i0:i0 + chunk] = np.einsum("ijt,j->it", w, amp).astype(np.float32)
    if noise:
        p = float(np.mean(out.astype(np.float64) ** 2))
        sigma = np.sqrt(p / 10.0 ** (sv.snr_db / 10.0))
        rng = np.random.default_rng(sv.seed ^ 0x5EED)
        out += (sigma * rng.standard_normal(out.shape)).astype(np.float32)
    return out


def line2d_lsm(seed=20240021) -> Survey:
    """The adjoint + LSM line: 200 shots x 240 receivers, 1501
    samples, 1600 image columns; 50 point scatterers for the reflectivity."""
    return line2d(nshots=200, nrecv_side=120, ns=1501, ncols=1600, nscat=50, seed=seed)


def line2d_vgrad(seed=20240041, **kw) -> Survey:
    """The velocity-gradient line: 600 shots x 240 split-spread receivers (144,000
    traces), 2001 samples, 2000 image columns at 16 m, 64 offset classes of
    23.5 m (class = floor((offset - 6.25) / 23.5), offsets 12.5..1500 m),
    aperture 2000 m, 50 scatterers."""
    args = dict(nshots=600, nrecv_side=120, ns=2001, ncols=2000, col_dx=16.0, aperture=2000.0,
                nscat=50, nh=64, oh=6.25, dh=23.5, x0=1500.0, seed=seed)
    args.update(kw)
    return line2d(**args)


def line2d_eom(seed=20240001):
    """The 40-shot line (40 shots x 120 receivers, 1001 samples) with
    an equivalent-offset migration set-up: 800 CSPs at 6.25 m from 0 (the KTM
    image columns), 160 offset bins of 12.5 m, aperture 2000 m, and the
    constant velocity of EOM.c's own model (2000 m/s).  Returns (survey,
    EomConfig)."""
    from .eom import EomConfig
    sv = line2d(seed=seed)
    cfg = EomConfig(ox=0.0, oy=0.0, angle_e_to_n=0.0, oxcsp=0.0, oycsp=0.0, excsp=5000.0,
                    eycsp=0.0, scale_xy=0.01, dcsp_x=6.25, dcsp_y=0.0, dh=12.5, nh=160,
                    aperture=2000.0, jobs_x=1, jobs_y=1, ntblock=50000,
                    velocity_start=2000.0, velocity_end=2000.0)
    return sv, cfg


def point_reflectivity(sv: Survey, image_shape) -> np.ndarray:
    """Image-domain model with one unit-amplitude spike (times the scatterer
    amplitude) at each scatterer's (column, t0 sample)."""
    m = np.zeros(image_shape, np.float32)
    for x, y, t0, a in sv.scatterers:
        X = int(round(x / sv.cfg.iline_distance)) - sv.cfg.start_iline
        Y = int(round(y / sv.cfg.xline_distance)) - sv.cfg.start_xline if image_shape[1] > 1 else 0
        T = int(round(t0 / sv.dt))
        if 0 <= X < image_shape[0] and 0 <= Y < image_shape[1] and 0 <= T < image_shape[3]:
            m[X, Y, 0, T] += a
    return m


def subset(sv: Survey, idx) -> Survey:
    """The same survey restricted to traces ``idx`` (order kept)."""
    import dataclasses
    return dataclasses.replace(sv, coords=np.ascontiguousarray(sv.coords[idx]))
