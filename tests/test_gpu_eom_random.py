"""Seeded random configurations of the equivalent-offset operator (EOM.c
csp()) against the oracle: a split-spread line or a rotated 3-D CSP grid
with its own origin, CSP spacing, offset bins (dh, nh, the integer oh),
aperture, job slabs, block size (ntblock), velocity and sample count drawn
per case.  The evaluated-sample count and fold match bit for bit, the
gathers to FP32 rounding (the taper's double atan may differ from the C
library's in its last bit), and the post filter to 1e-6."""
import numpy as np
import pytest
import torch

import oracle
from random_amd import surveys
from random_amd.eom import EomConfig, EomPlan

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def random_eom_case(seed):
    rng = np.random.default_rng(seed)
    ns = int(rng.integers(120, 600))
    t1 = 0.8 * ns * 0.004
    v = float(rng.uniform(1500.0, 3500.0))
    common = dict(dh=float(rng.uniform(8.0, 50.0)), nh=int(rng.integers(4, 48)),
                  oh=int(rng.integers(0, 30)), aperture=float(rng.uniform(250.0, 1500.0)),
                  ntblock=int(rng.integers(16, 300)), velocity_start=v, velocity_end=v)
    if rng.integers(0, 2):
        ncols = int(rng.integers(20, 90))
        dx = float(rng.choice([6.25, 12.5, 25.0]))
        sv = surveys.line2d(nshots=int(rng.integers(3, 9)), nrecv_side=int(rng.integers(4, 20)),
                            ns=ns, ncols=ncols, col_dx=dx, nscat=3, x0=float(rng.uniform(150, 700)),
                            t0_range=(0.1, t1), seed=seed)
        cfg = EomConfig(ox=0.0, oy=0.0, angle_e_to_n=0.0, oxcsp=0.0, oycsp=0.0,
                        excsp=ncols * dx, eycsp=0.0, scale_xy=0.01, dcsp_x=dx, dcsp_y=0.0,
                        jobs_x=int(rng.integers(1, 5)), jobs_y=1, **common)
    else:
        sv = surveys.area3d(ntr=int(rng.integers(150, 500)), ns=ns, ncols=16, extent=700.0,
                            max_offset=float(rng.uniform(300.0, 1200.0)), nscat=6, seed=seed)
        sv.scatterers[:, 2] = rng.uniform(0.1, t1, sv.scatterers.shape[0])   # inside the traces
        # header words are metres + 5000: a rotated survey frame with its own origin
        cfg = EomConfig(ox=float(rng.uniform(4800.0, 5000.0)), oy=float(rng.uniform(4800.0, 5200.0)),
                        angle_e_to_n=float(rng.uniform(-30.0, 30.0)),
                        oxcsp=float(rng.uniform(4950.0, 5050.0)), oycsp=float(rng.uniform(4950.0, 5050.0)),
                        excsp=float(rng.uniform(5300.0, 5600.0)), eycsp=float(rng.uniform(5300.0, 5600.0)),
                        scale_xy=1.0, dcsp_x=float(rng.uniform(15.0, 45.0)),
                        dcsp_y=float(rng.uniform(15.0, 45.0)), jobs_x=int(rng.integers(1, 4)),
                        jobs_y=int(rng.integers(1, 4)), **common)
    return sv, cfg


@pytest.mark.parametrize("seed", range(16))
def test_random_eom_configuration(cuda, seed):
    sv, cfg = random_eom_case(7000 + seed)
    print(f"case: EOM seed {7000 + seed}: {sv.name}, ns {sv.ns}, {sv.ntr} traces, angle "
          f"{cfg.angle_e_to_n:.2f}, dcsp {cfg.dcsp_x:.2f} x {cfg.dcsp_y:.2f}, dh {cfg.dh:.2f}, "
          f"nh {cfg.nh}, oh {cfg.oh}, aperture {cfg.aperture:.0f}, jobs {cfg.jobs_x}x{cfg.jobs_y}, "
          f"ntblock {cfg.ntblock}", flush=True)
    tr = surveys.traces_numpy(sv)
    vel = cfg.velocity(sv.ns)
    c0, f0, n0 = oracle.eom_all(cfg, sv.ns, np.float32(sv.dt), tr, sv.coords, vel)
    plan = EomPlan(cfg, sv.ns, sv.dt)
    csp, fold, n = plan.migrate(torch.from_numpy(tr).to(cuda), torch.from_numpy(sv.coords).to(cuda))
    c, f = csp.cpu().numpy(), fold.cpu().numpy()
    r = rel_l2(c, c0)
    print(f"  evaluated {n} (oracle {n0}); gathers relL2 {r:.2e}", flush=True)
    assert n == n0
    assert np.array_equal(f, f0)
    assert r <= 1e-6
    post = plan.post(csp, fold).cpu().numpy()
    lp, rr = oracle.eom_filters(cfg, sv.ns, np.float32(sv.dt))
    ref = oracle.eom_post(c0, f0, np.float32(sv.dt), lp, rr)
    assert rel_l2(post, ref) <= 1e-6
