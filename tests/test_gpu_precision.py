"""Accuracy of the accurate-filter and reduced-precision paths against the
FP64 oracle, which uses the same (FP32, reference) traveltime index and an
FP64 data path -- so the errors below are those of trace storage, filter
arithmetic and accumulation only."""
import numpy as np
import pytest
import torch

import oracle
from random_amd import (ACCUM_KAHAN, ACCUM_PLAIN, PREP_CLOSED, PREP_REFERENCE, TRACE_BF16,
                        TRACE_F16, TRACE_F32, surveys)
from random_amd.ktm import Plan

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def area():
    sv = surveys.area3d(ntr=3000, ncols=32, extent=1500.0, max_offset=4000.0, nscat=20,
                        jobs_x=4, jobs_y=2)
    tr = surveys.traces_numpy(sv)
    km64, fold64, n = oracle.migrate(sv.cfg, sv.ns, sv.dt, tr, sv.coords, sv.vel,
                                     mode=oracle.FP64_DATA)
    return sv, tr, km64, fold64, n


def test_closed_prep_matches_exact_filter(cuda, area):
    sv, tr, *_ = area
    print("case: CLOSED prep vs FP64 prefix-sum filter, 200 traces", flush=True)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    sel = np.arange(200)
    prepared, _ = plan.prep(torch.from_numpy(tr[sel]).to(cuda),
                            torch.from_numpy(sv.coords[sel]).to(cuda), PREP_CLOSED, TRACE_F32)
    F = prepared.cpu().numpy()[:, :sv.ns]
    tm = plan.setup["trange_max"]
    ref = np.stack([oracle.filtered_trace(oracle.trace_prep(tr[i], sv.dt, tm, fp64=True), tm,
                                          sv.ns) for i in sel])
    r = rel_l2(F, ref)
    print(f"  relL2 {r:.3e}")
    assert r < 1e-6


@pytest.mark.parametrize("name,mode,dtype,accum,tol", [
    ("fp32-closed-kahan", PREP_CLOSED, TRACE_F32, ACCUM_KAHAN, 1e-6),
    ("fp32-closed-plain", PREP_CLOSED, TRACE_F32, ACCUM_PLAIN, 1e-5),
    ("fp16-kahan", PREP_CLOSED, TRACE_F16, ACCUM_KAHAN, 1e-3),
    ("bf16-kahan", PREP_CLOSED, TRACE_BF16, ACCUM_KAHAN, 3e-3),
])
def test_migration_error_vs_fp64(cuda, area, name, mode, dtype, accum, tol):
    sv, tr, km64, fold64, n64 = area
    print(f"case: {name} vs FP64 oracle, 3000 traces, 32x32x1500", flush=True)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    km, fold, n = plan.migrate(torch.from_numpy(tr).to(cuda), torch.from_numpy(sv.coords).to(cuda),
                               mode=mode, dtype=dtype, accum=accum, block=1000)
    km = km.cpu().numpy()
    r = rel_l2(km, km64)
    mx = float(np.abs(km - km64).max() / np.abs(km64).max())
    print(f"  relL2 {r:.3e}  max-abs/max {mx:.3e}")
    assert n == n64
    assert np.array_equal(fold.cpu().numpy(), fold64.astype(np.float32))
    assert r <= tol


def test_reference_fp32_error_is_the_references_own(cuda, area):
    """The FP32 reference path reproduces Kirchhoff.c bit for bit; its error
    against FP64 is the cancellation of the FP32 prefix sums (large at late
    times), identical for the oracle's FP32 restatement."""
    sv, tr, km64, _, _ = area
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    km, _, _ = plan.migrate(torch.from_numpy(tr).to(cuda), torch.from_numpy(sv.coords).to(cuda),
                            mode=PREP_REFERENCE)
    km = km.cpu().numpy()
    early = rel_l2(km[..., :300], km64[..., :300])
    late = rel_l2(km[..., 1000:], km64[..., 1000:])
    print(f"case: FP32 reference vs FP64: relL2 t<300 {early:.3e}, t>=1000 {late:.3e}", flush=True)
    assert early < late
