"""Adjoint-pair wrappers: autograd, dot-product test and least-squares
migration (CGLS) on the GPU kernels.

L^T (migration, ``Plan.migrate_lin``) maps traces to the raw image, L
(demigration, ``Plan.demigrate``) maps an image to traces; with
``poststack=True`` the chain includes the derivative / fold / low-pass of
Kirchhoff.c:1457-1479 (``Plan.poststack`` and its transpose).  Scalars of
the iterations (dot products, step lengths, residual norms) are FP64.
"""
from __future__ import annotations

import torch

from .ktm import Plan


def dot64(a: torch.Tensor, b: torch.Tensor) -> float:
    """<a, b> accumulated in FP64."""
    return float(torch.dot(a.reshape(-1).double(), b.reshape(-1).double()).item())


class KirchhoffMigration(torch.autograd.Function):
    """image = L^T traces (optionally followed by the post-stack chain);
    backward applies L (and the post-stack transpose)."""

    @staticmethod
    def forward(ctx, traces, plan: Plan, coords, poststack: bool = False, fast: bool = False):
        km, fold = plan.migrate_lin(traces.contiguous(), coords, kahan=True, fast=fast)
        geom = plan.geometry(coords)
        ctx.plan, ctx.geom, ctx.fold, ctx.poststack = plan, geom, fold, poststack
        ctx.fast = fast
        if poststack:
            return plan.poststack(km, fold)
        return km

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.contiguous()
        if ctx.poststack:
            g = ctx.plan.poststack_adjoint(g, ctx.fold)
        return ctx.plan.demigrate(g, ctx.geom, fast=ctx.fast), None, None, None, None


def migrate_op(traces, plan: Plan, coords, poststack: bool = False, fast: bool = False):
    return KirchhoffMigration.apply(traces, plan, coords, poststack, fast)


class Operator:
    """L / L^T on one survey with the geometry cached.  L^T accumulates with
    the Kahan-compensated FP32 kernel by default: at ~2e4 traces per image
    sample a plain FP32 sum carries ~1e-5 relative error, the size of the
    adjointness tolerance.  ``fast`` selects the pair with the relaxed
    traveltime index (INDEX_FAST) on both sides."""

    def __init__(self, plan: Plan, coords: torch.Tensor, kahan: bool = True, fast: bool = False):
        self.plan = plan
        self.coords = coords
        self.kahan = kahan
        self.fast = fast
        self.geom = plan.geometry(coords)

    def adjoint(self, d):      # L^T: traces -> image
        return self.plan.migrate_lin(d, self.coords, kahan=self.kahan, fast=self.fast)[0]

    def forward(self, m):      # L: image -> traces
        return self.plan.demigrate(m, self.geom, fast=self.fast)


def dot_test(op: Operator, d: torch.Tensor, m: torch.Tensor) -> tuple[float, float, float]:
    """(<L^T d, m>, <d, L m>, relative difference)."""
    lhs = dot64(op.adjoint(d), m)
    rhs = dot64(d, op.forward(m))
    return lhs, rhs, abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)


def time_gain(op: Operator, image_shape, nprobe: int = 4, seed: int = 0) -> torch.Tensor:
    """Diagonal preconditioner w[t] = 1 / sqrt(g[t]) of least-squares
    migration, g[t] the mean over image columns and offset classes of
    diag(L^T L) at output sample t, estimated from nprobe Rademacher probes z
    as mean(z * L^T L z) (Hutchinson).  The 1 / (4 t2) weights of the
    trace integration (Kirchhoff.c:1315-1316) make diag(L^T L) fall by six
    orders of magnitude down the trace; w undoes that spread.  Returns a
    tensor broadcastable to the image, [ns]."""
    dev = op.coords.device
    g = torch.Generator(device=dev).manual_seed(seed)
    est = torch.zeros(image_shape, dtype=torch.float32, device=dev)
    for _ in range(nprobe):
        z = torch.randint(0, 2, image_shape, generator=g, device=dev).float() * 2.0 - 1.0
        est += z * op.adjoint(op.forward(z))
    prof = (est / nprobe).double().mean(dim=tuple(range(len(image_shape) - 1)))
    prof = prof.clamp_min(float(prof.max()) * 1e-12)
    return (1.0 / prof.sqrt()).float()


class _Preconditioned:
    """A = L diag(w) for right-preconditioned CGLS (m = w * y)."""

    def __init__(self, op: Operator, w: torch.Tensor):
        self.op, self.w = op, w

    def forward(self, y):
        return self.op.forward(self.w * y)

    def adjoint(self, r):
        return self.w * self.op.adjoint(r)


def cgls(op: Operator, d_obs: torch.Tensor, niter: int = 20, m0: torch.Tensor | None = None,
         callback=None, precond: torch.Tensor | None = None):
    """Least-squares migration: min_m ||L m - d_obs||^2 by conjugate gradients
    on the normal equations (CGLS), FP32 vectors, FP64 scalars.  Returns
    (m, history) with history[k] = (||r_k||, ||A^T r_k||), A = L, or
    A = L diag(precond) when a diagonal preconditioner is given (CGLS then
    runs on y with m = precond * y; r = d_obs - L m either way)."""
    if precond is not None:
        y0 = None if m0 is None else m0 / precond
        y, hist = cgls(_Preconditioned(op, precond), d_obs, niter, y0, callback)
        return precond * y, hist
    m = torch.zeros_like(op.adjoint(d_obs)) if m0 is None else m0.clone()
    r = d_obs - op.forward(m) if m0 is not None else d_obs.clone()
    s = op.adjoint(r)
    p = s.clone()
    gamma = dot64(s, s)
    hist = [(dot64(r, r) ** 0.5, gamma ** 0.5)]
    for k in range(niter):
        q = op.forward(p)
        qq = dot64(q, q)
        if qq <= 0.0:
            break
        alpha = gamma / qq
        m += alpha * p
        r -= alpha * q
        s = op.adjoint(r)
        gamma_new = dot64(s, s)
        hist.append((dot64(r, r) ** 0.5, gamma_new ** 0.5))
        if callback:
            callback(k, m, hist[-1])
        if gamma_new == 0.0:
            break
        p = s + (gamma_new / gamma) * p
        gamma = gamma_new
    return m, hist
