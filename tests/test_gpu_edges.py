"""Edge cases of the migration path on the GPU against the CPU oracle: every
kernel variant (the sample-count boundaries where the per-column lane layout
changes, from the smallest plan, ns = 4, to the largest, ns = 3072), empty
blocks, blocks with no trace in aperture, offset classes with traces outside
[0, nh), a one-column image, the plan's size limits, and the adjoint of every
variant."""
import numpy as np
import pytest
import torch

import oracle
from random_amd import KtmError, _lib, surveys
from random_amd.ktm import Plan
from random_amd.lsm import Operator, dot_test

pytestmark = pytest.mark.gpu

# sample counts at and around the lane-layout boundaries of ktm_migrate_block
# (K samples per lane x W waves per column: 8x1 <= 512 < 8x2 <= 1024 < 12x2
# <= 1536 < 8x4 <= 2048 < 12x4 <= 3072)
NS_CASES = [4, 63, 512, 513, 1024, 1025, 1536, 1537, 2048, 2049, 3072]


def small_line(ns, **kw):
    dt = 0.004
    t1 = max(0.02, 0.8 * ns * dt)
    # shots from x = 100 m, so the 40 columns of 6.25 m from 0 straddle the
    # CMPs and even a 4-sample trace reaches the image
    args = dict(nshots=3, nrecv_side=10, ns=ns, ncols=40, nscat=3, jobs_x=2, x0=100.0,
                t0_range=(min(0.01, t1 / 2), t1))
    args.update(kw)
    return surveys.line2d(**args)


def bitexact_case(sv, tr, cuda, block=None):
    km_o, fold_o, n_o = oracle.migrate(sv.cfg, sv.ns, sv.dt, tr, sv.coords, sv.vel)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    km, fold, n = plan.migrate(torch.from_numpy(tr).to(cuda), torch.from_numpy(sv.coords).to(cuda),
                               block=block)
    km, fold = km.cpu().numpy(), fold.cpu().numpy()
    same = float(np.mean(km.view(np.uint32) == km_o.view(np.uint32)))
    print(f"  contributions {n} (oracle {n_o}); bit-identical fraction {same:.8f}", flush=True)
    assert n == n_o
    assert np.array_equal(fold, fold_o)
    assert same == 1.0
    return plan, km, fold, n


@pytest.mark.parametrize("ns", NS_CASES)
def test_every_kernel_variant_bitexact(cuda, ns):
    print(f"case: variant at ns={ns}, 60 traces, 40 columns, 2 jobs", flush=True)
    sv = small_line(ns)
    tr = surveys.traces_numpy(sv)
    assert np.abs(tr).max() > 0
    _, km, _, n = bitexact_case(sv, tr, cuda, block=25)
    assert n > 0 and np.abs(km).max() > 0


@pytest.mark.parametrize("ns", [4, 513, 1537, 2049, 3072])
def test_adjoint_of_every_variant(cuda, ns):
    print(f"case: dot test at ns={ns}", flush=True)
    sv = small_line(ns)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    op = Operator(plan, torch.from_numpy(sv.coords).to(cuda))
    g = torch.Generator(device=cuda).manual_seed(ns)
    d = torch.randn((sv.ntr, sv.ns), generator=g, device=cuda)
    m = torch.randn(plan.image_shape, generator=g, device=cuda)
    lhs, rhs, rel = dot_test(op, d, m)
    print(f"  rel {rel:.2e}", flush=True)
    assert rel <= 1e-5


def test_empty_and_out_of_aperture_blocks(cuda):
    print("case: empty block, block with no trace in aperture", flush=True)
    sv = small_line(301)
    tr = surveys.traces_numpy(sv)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    km, fold, _ = plan.new_image(device=cuda)
    counter = torch.zeros(1, dtype=torch.int64, device=cuda)
    d_tr = torch.from_numpy(tr).to(cuda)
    # an empty block is a no-op
    prepared, geom = plan.prep(d_tr[:0], torch.zeros((0, 4), dtype=torch.int32, device=cuda))
    plan.migrate_block(prepared, geom, km, fold, counter=counter)
    # traces 1000 km down the line: no column within the 2000 m aperture
    far = sv.coords.copy()
    far[:, 0] += 100_000_000
    far[:, 2] += 100_000_000
    prepared, geom = plan.prep(d_tr, torch.from_numpy(far).to(cuda))
    plan.migrate_block(prepared, geom, km, fold, counter=counter)
    torch.cuda.synchronize()
    assert int(counter.item()) == 0
    assert not km.any() and bool((fold == 1.0).all())
    km_o, fold_o, n_o = oracle.migrate(sv.cfg, sv.ns, sv.dt, tr, far, sv.vel)
    assert n_o == 0 and not km_o.any()


def test_offset_classes_with_traces_outside(cuda):
    """nh = 3 classes of 100 m on a split spread up to 375 m: classes 0-2 are
    filled, the traces of class 3 are skipped (Kirchhoff.c:1387)."""
    print("case: offset classes, nh=3, traces beyond the last class", flush=True)
    sv = small_line(401, nshots=4, nrecv_side=30, nh=3, oh=0.0, dh=100.0)
    tr = surveys.traces_numpy(sv)
    _, geo_off = oracle.trace_geometry(sv.cfg, sv.coords)
    assert set(np.unique(geo_off)) >= {0, 1, 2, 3}
    plan, km, fold, _ = bitexact_case(sv, tr, cuda, block=100)
    assert plan.image_shape[2] == 3
    for h in range(3):
        assert np.abs(km[:, :, h]).max() > 0


def test_one_column_image(cuda):
    print("case: one image column, two traces", flush=True)
    sv = small_line(257, nshots=1, nrecv_side=1, ncols=1, jobs_x=1)
    tr = surveys.traces_numpy(sv)
    assert tr.shape[0] == 2
    plan, km, _, n = bitexact_case(sv, tr, cuda)
    assert plan.image_shape[:2] == (1, 1)


def test_plan_size_limits(cuda):
    sv = small_line(64)
    for ns in (3, 3073):
        with pytest.raises(KtmError) as e:
            Plan(sv.cfg, ns, sv.dt, np.ones((sv.cfg.nvel_ilines, 1, ns), np.float32) * 2000.0)
        assert e.value.status == _lib.KTM_ERR_UNSUPPORTED
