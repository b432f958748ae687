"""The adjoint pair (migration L^T / demigration L), its autograd wrapper and
least-squares migration (the adjoint + LSM configuration: 200 shots x 240
receivers, 1501 samples).  Dot products accumulate in FP64."""
import numpy as np
import pytest
import torch

from random_amd import surveys
from random_amd.ktm import Plan
from random_amd.lsm import Operator, cgls, dot64, dot_test, migrate_op, time_gain

pytestmark = pytest.mark.gpu


def survey_cases():
    yield "line2d", surveys.line2d(nshots=6, ncols=240, ns=501, nscat=4, jobs_x=3)
    yield "area3d", surveys.area3d(ntr=600, ns=700, ncols=24, extent=800.0, max_offset=2000.0,
                                   nscat=4, jobs_x=2, jobs_y=3)
    yield "offsets", surveys.line2d(nshots=6, ncols=160, ns=401, nscat=3, nh=8, oh=0.0, dh=200.0)


@pytest.mark.parametrize("name,sv", list(survey_cases()), ids=lambda x: x if isinstance(x, str) else "")
def test_dot_product(cuda, name, sv):
    print(f"case: dot test {name}", flush=True)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    op = Operator(plan, torch.from_numpy(sv.coords).to(cuda))
    g = torch.Generator(device=cuda).manual_seed(11)
    d = torch.randn((sv.ntr, sv.ns), generator=g, device=cuda)
    m = torch.randn(plan.image_shape, generator=g, device=cuda)
    lhs, rhs, rel = dot_test(op, d, m)
    print(f"  <L^T d, m> = {lhs:.9e}  <d, L m> = {rhs:.9e}  rel {rel:.2e}")
    assert rel <= 1e-5


def test_poststack_adjoint_dot(cuda):
    sv = surveys.line2d(nshots=3, ncols=64, ns=1001, nscat=2)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    g = torch.Generator(device=cuda).manual_seed(3)
    km = torch.randn(plan.image_shape, generator=g, device=cuda)
    img = torch.randn(plan.image_shape, generator=g, device=cuda)
    fold = torch.randint(1, 40, plan.image_shape, generator=g, device=cuda).float()
    lhs = dot64(plan.poststack(km, fold), img)
    rhs = dot64(km, plan.poststack_adjoint(img, fold))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"case: post-stack transpose <S km, i> {lhs:.9e} <km, S^T i> {rhs:.9e} rel {rel:.2e}",
          flush=True)
    assert rel <= 1e-6


@pytest.mark.parametrize("poststack", [False, True])
def test_autograd_finite_differences(cuda, poststack):
    """J(d) = 0.5 ||F(d) - m_obs||^2 with F the migration (autograd forward);
    the gradient from the backward (demigration) against central differences
    along 16 random directions."""
    sv = surveys.line2d(nshots=4, ncols=160, ns=501, nscat=3)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    coords = torch.from_numpy(sv.coords).to(cuda)
    g = torch.Generator(device=cuda).manual_seed(5)
    d = torch.randn((sv.ntr, sv.ns), generator=g, device=cuda, requires_grad=True)
    with torch.no_grad():
        out0 = migrate_op(d, plan, coords, poststack)
        m_obs = (0.5 * out0 + out0.std() * torch.randn(plan.image_shape, generator=g,
                                                        device=cuda)).double()

    def J(x):  # residual in FP64 so that only the operator's own rounding remains
        r = migrate_op(x, plan, coords, poststack).double() - m_obs
        return 0.5 * float((r ** 2).sum().item())

    out = migrate_op(d, plan, coords, poststack)
    loss = 0.5 * ((out.double() - m_obs) ** 2).sum()
    loss.backward()
    grad = d.grad.detach()
    worst = 0.0
    with torch.no_grad():
        for k in range(16):
            u = torch.randn(d.shape, generator=g, device=cuda)
            u /= u.norm()
            eps = 0.05 * float(d.norm())  # J is quadratic: no truncation error
            fd = (J(d + eps * u) - J(d - eps * u)) / (2 * eps)
            an = dot64(grad, u)
            rel = abs(fd - an) / max(abs(an), 1e-30)
            worst = max(worst, rel)
    print(f"case: autograd FD (poststack={poststack}): worst relative difference {worst:.2e}",
          flush=True)
    assert worst <= 1e-3


def test_lsm_configuration(cuda):
    """Adjoint + LSM configuration: dot test at full size, then 20 CG
    iterations from data modelled by L on 50 point scatterers."""
    sv = surveys.line2d_lsm()
    print(f"case: LSM {sv.ntr} traces x {sv.ns}, image {sv.cfg.end_iline} columns", flush=True)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    op = Operator(plan, torch.from_numpy(sv.coords).to(cuda))
    g = torch.Generator(device=cuda).manual_seed(1)
    d = torch.randn((sv.ntr, sv.ns), generator=g, device=cuda)
    m = torch.randn(plan.image_shape, generator=g, device=cuda)
    lhs, rhs, rel = dot_test(op, d, m)
    print(f"  dot test rel {rel:.2e}")
    assert rel <= 1e-5
    m_true = torch.from_numpy(surveys.point_reflectivity(sv, plan.image_shape)).to(cuda)
    d_obs = op.forward(m_true)
    m_est, hist = cgls(op, d_obs, niter=20)
    res = [h[0] / hist[0][0] for h in hist]
    print("  relative residual per iteration: " + " ".join(f"{r:.3f}" for r in res))
    corr = dot64(m_est, m_true) / (dot64(m_est, m_est) * dot64(m_true, m_true)) ** 0.5
    mig = op.adjoint(d_obs)
    corr0 = dot64(mig, m_true) / (dot64(mig, mig) * dot64(m_true, m_true)) ** 0.5
    print(f"  correlation with the true reflectivity: migration {corr0:.3f}, LSM {corr:.3f}")
    # CGLS decreases ||L m - d|| monotonically; the 1/t^2 weights of the
    # filter make L^T L badly conditioned, so 20 unpreconditioned iterations
    # remove ~18% of the residual while doubling the correlation with the truth
    assert all(b <= a * (1 + 1e-6) for a, b in zip(res, res[1:]))
    assert res[-1] < 0.9
    assert corr > 1.5 * corr0
    # with the time-gain preconditioner (4 Hutchinson probes of diag(L^T L))
    w = time_gain(op, plan.image_shape, nprobe=4, seed=5)
    m_pc, hist_pc = cgls(op, d_obs, niter=20, precond=w)
    res_pc = [h[0] / hist_pc[0][0] for h in hist_pc]
    corr_pc = dot64(m_pc, m_true) / (dot64(m_pc, m_pc) * dot64(m_true, m_true)) ** 0.5
    print("  preconditioned: relative residual " + " ".join(f"{r:.3f}" for r in res_pc[::4]) +
          f" ... {res_pc[-1]:.3f}; correlation {corr_pc:.3f}")
    assert all(b <= a * (1 + 1e-6) for a, b in zip(res_pc, res_pc[1:]))
    assert res_pc[-1] < 0.65 * res[-1]
    assert corr_pc > 2.0 * corr
