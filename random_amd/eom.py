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

import dataclasses
from dataclasses import dataclass


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

    def velocity(self, ns: int):
        """vel_t of EOM.c:871-875: velocityStart + (End - Start) * (i/ns) with
        the integer division of the source, i.e. velocityStart for every
        sample (i < ns)."""
        import numpy as np
        return np.array([self.velocity_start + (self.velocity_end - self.velocity_start) * (i // ns)
                         for i in range(ns)], np.float32)
