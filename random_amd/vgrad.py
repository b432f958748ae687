"""Velocity gradient of a gather-flatness objective (``ktm_vgrad_*``).

A differentiable variant of the migration loop of Kirchhoff.c:1332-1404 for
velocity analysis on a line: the RMS velocity is a uniform cubic B-spline
on an (nt_nodes x nx_nodes) grid of (time, inline) nodes, the traveltime is
not floored and the CLOSED filtered trace is read by Catmull-Rom cubic
interpolation, so the image is C1 in the node velocities.  The migration
writes one common-image gather of nh offset classes per column and

    J(V) = sum_{X,t} sum_{h<nh-1} (I[X][h+1][t] - I[X][h][t])^2

measures how far the gathers are from flat.  ``VelocityGradient`` evaluates
J and dJ/dV on the GPU: one workgroup owns (column, time chunk), streams the
offset classes through registers and recomputes the traveltimes for the
backward pass instead of storing them.  Binding copies the traces into a
table the plan holds, binned by (offset class, CMP) and zero-padded.  The reference has no such operator
(its nearest-node velocity, Kirchhoff.c:1357-1358,1376, and floor(dsr/dt),
1384, have zero derivative almost everywhere).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import PREP_CLOSED, check, lib
from .ktm import Plan


def linear_nodes(ns: int, dt: float, nt_nodes: int, nx_nodes: int, v0: float, vt: float,
                 scale: float = 1.0) -> np.ndarray:
    """Node velocities whose spline is exactly scale * (v0 + vt * t): cubic
    B-splines reproduce linear functions when the nodes sample them at the
    node centres t_i = (i - 1) dtn."""
    dtn = (ns - 1) / (nt_nodes - 3) * dt
    ti = (np.arange(nt_nodes) - 1.0) * dtn
    V = scale * (v0 + vt * ti)
    return np.ascontiguousarray(np.repeat(V[:, None], nx_nodes, axis=1), np.float32)


class VelocityGradient:
    """J(V) and dJ/dV for one bound block of traces (device-resident)."""

    def __init__(self, plan: Plan, traces, coords, nt_nodes: int = 100, nx_nodes: int = 200,
                 chunk: int = 250):
        import torch
        self.plan = plan
        self.nt_nodes, self.nx_nodes, self.chunk = int(nt_nodes), int(nx_nodes), int(chunk)
        self.prepared, self.geom = plan.prep(traces, coords, PREP_CLOSED)
        h = C.c_void_p()
        c = plan.cfg.to_c()
        check(lib().ktm_vgrad_create(C.byref(h), C.byref(c), plan.ns, plan.dt, self.nt_nodes,
                                     self.nx_nodes, self.chunk), "ktm_vgrad_create")
        self._h = h
        stream = torch.cuda.current_stream().cuda_stream
        check(lib().ktm_vgrad_bind(self._h, self.prepared.data_ptr(), plan.row,
                                   self.geom.data_ptr(), traces.shape[0], stream), "ktm_vgrad_bind")
        self.cig_shape = (plan.nxp, plan.nh, plan.ns)
        #: (trace, column) pairs in aperture; a pass evaluates pairs * ns samples
        self.pairs = int(lib().ktm_vgrad_pairs(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().ktm_vgrad_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _nodes(self, V):
        import torch
        if not isinstance(V, torch.Tensor):
            V = torch.as_tensor(np.asarray(V, np.float32))
        V = V.to(device="cuda", dtype=torch.float32).contiguous()
        if tuple(V.shape) != (self.nt_nodes, self.nx_nodes):
            raise ValueError(f"V must be [{self.nt_nodes}][{self.nx_nodes}], got {tuple(V.shape)}")
        return V

    def velocity(self, V):
        """v(t, X) on the image grid, [nxp][ns]."""
        import torch
        V = self._nodes(V)
        v = torch.empty((self.plan.nxp, self.plan.ns), dtype=torch.float32, device="cuda")
        check(lib().ktm_vgrad_velocity(self._h, V.data_ptr(), v.data_ptr(),
                                       torch.cuda.current_stream().cuda_stream),
              "ktm_vgrad_velocity")
        return v

    def __call__(self, V, grad: bool = True, cig: bool = False):
        """(J as a Python float, dJ/dV float64 tensor or None, gathers or None)."""
        import torch
        V = self._nodes(V)
        J = C.c_double()
        g = torch.empty((self.nt_nodes, self.nx_nodes), dtype=torch.float64,
                        device="cuda") if grad else None
        I = torch.empty(self.cig_shape, dtype=torch.float32, device="cuda") if cig else None
        check(lib().ktm_vgrad_eval(self._h, V.data_ptr(), C.byref(J),
                                   None if g is None else g.data_ptr(),
                                   None if I is None else I.data_ptr(),
                                   torch.cuda.current_stream().cuda_stream), "ktm_vgrad_eval")
        return J.value, g, I


def central_differences(vg: VelocityGradient, V: np.ndarray, nodes, h: float = 1.0) -> np.ndarray:
    """(J(V + h e_k) - J(V - h e_k)) / 2h at each node k = (i, j)."""
    out = []
    V = np.asarray(V, np.float32)
    for (i, j) in nodes:
        Vp = V.copy()
        Vp[i, j] += h
        Vm = V.copy()
        Vm[i, j] -= h
        Jp, _, _ = vg(Vp, grad=False)
        Jm, _, _ = vg(Vm, grad=False)
        out.append((Jp - Jm) / (float(Vp[i, j]) - float(Vm[i, j])))
    return np.array(out)
