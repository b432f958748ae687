"""Edge cases of the migration path on the GPU against the CPU oracle: every
kernel variant (the sample-count boundaries where the per-column lane layout
changes, from the smallest plan, ns = 4, to the largest, ns = 6144), empty
blocks, blocks with no trace in aperture, offset classes with traces outside
[0, nh), a one-column image, the plan's size limits, and the adjoint of every
variant, and the post-stack filter and its transpose for long traces (the
traces-per-workgroup and kernel-in-LDS choices of ktm_poststack)."""
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
# <= 1536 < 8x4 <= 2048 < 12x4 <= 3072 < 8x8 <= 4096 < 12x8 <= 6144; images
# too small for two workgroups per CU take 4x2 and 4x4 below 1025 samples,
# as the 40-column line here does -- test_wide_image_layouts_bitexact keeps
# the default ones covered)
NS_CASES = [4, 63, 512, 513, 1024, 1025, 1536, 1537, 2048, 2049, 3072, 3073, 4096, 4097, 6144]


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


@pytest.mark.parametrize("ns", [63, 512, 513, 1024])
def test_wide_image_layouts_bitexact(cuda, ns):
    """A 4104-column line (more than two workgroups per CU at 8 samples per
    lane) keeps the default layouts below 1025 samples."""
    print(f"case: wide-image variant at ns={ns}, 60 traces, 4104 columns, 2 jobs", flush=True)
    sv = small_line(ns, ncols=4104)
    tr = surveys.traces_numpy(sv)
    _, km, _, n = bitexact_case(sv, tr, cuda, block=25)
    assert n > 0 and np.abs(km).max() > 0


@pytest.mark.parametrize("ns", [4, 513, 1537, 2049, 3072, 4097, 6144])
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
    for ns in (3, 6145):
        with pytest.raises(KtmError) as e:
            Plan(sv.cfg, ns, sv.dt, np.ones((sv.cfg.nvel_ilines, 1, ns), np.float32) * 2000.0)
        assert e.value.status == _lib.KTM_ERR_UNSUPPORTED


@pytest.mark.parametrize("ns", [4, 128, 129, 512, 1024, 1025, 1536, 1537, 2048, 2049, 4096, 4097,
                                6144])
def test_poststack_long_traces(cuda, ns):
    """Forward against the oracle's DFT path and the dot test of the
    transpose on both sides of every transform length of the FFT form
    (N = 256 up to 128 samples, 512 from 129, 1024, 2048, 3072 from 1025 to
    1536 (radix 16 x 16 x 12), 4096 from 1537 to 2048, 8192 from 2049 to
    4096) and of the switch to the FP64 matrix product beyond 4096 samples;
    8 image traces at even ns, 7 at odd ns (the last transform then carries
    one trace)."""
    print(f"case: post-stack and transpose at ns={ns}", flush=True)
    sv = small_line(ns, ncols=7 if ns % 2 else 8, jobs_x=1)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    g = torch.Generator(device=cuda).manual_seed(ns + 1)
    km = torch.randn(plan.image_shape, generator=g, device=cuda)
    img = torch.randn(plan.image_shape, generator=g, device=cuda)
    fold = torch.randint(1, 40, plan.image_shape, generator=g, device=cuda).float()
    out = plan.poststack(km, fold).cpu().numpy()
    tm, aa = oracle.alias_filter(sv.cfg, sv.ns, sv.dt)
    ref = oracle.poststack(km.cpu().numpy(), fold.cpu().numpy(), sv.dt, aa)
    r = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    skm = plan.poststack(km, fold).double().ravel()
    lhs = float(torch.dot(skm, img.double().ravel()))
    rhs = float(torch.dot(km.double().ravel(), plan.poststack_adjoint(img, fold).double().ravel()))
    rel = abs(lhs - rhs) / abs(lhs)
    # the outputs are rounded to float, so the difference is ~1e-7 of the
    # Cauchy-Schwarz scale over sqrt(n); relative to <S km, img> unless that
    # product is below 1 % of the scale by chance
    scale = float(torch.linalg.vector_norm(skm) * torch.linalg.vector_norm(img.double()))
    print(f"  forward relL2 vs oracle {r:.2e}; transpose dot rel {rel:.2e}, "
          f"|<S km, i>| / scale {abs(lhs) / scale:.2e}", flush=True)
    assert r <= 1e-6
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), 1e-2 * scale)


@pytest.mark.parametrize("ns", [4097, 6144])
def test_compensated_and_16bit_long_traces(cuda, ns):
    """The Kahan and BF16 variants of the long-trace layouts against the FP64
    oracle (same traveltime index, FP64 data path)."""
    from random_amd import ACCUM_KAHAN, PREP_CLOSED, TRACE_BF16, TRACE_F32
    print(f"case: Kahan FP32 and BF16 at ns={ns}", flush=True)
    sv = small_line(ns)
    tr = surveys.traces_numpy(sv)
    km64, fold64, n64 = oracle.migrate(sv.cfg, sv.ns, sv.dt, tr, sv.coords, sv.vel,
                                       mode=oracle.FP64_DATA)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    d_tr, d_co = torch.from_numpy(tr).to(cuda), torch.from_numpy(sv.coords).to(cuda)
    for dtype, tol in ((TRACE_F32, 1e-6), (TRACE_BF16, 3e-3)):
        km, fold, n = plan.migrate(d_tr, d_co, mode=PREP_CLOSED, dtype=dtype, accum=ACCUM_KAHAN,
                                   block=25)
        r = float(np.linalg.norm(km.cpu().numpy() - km64) / np.linalg.norm(km64))
        print(f"  dtype {dtype}: relL2 vs FP64 {r:.2e}", flush=True)
        assert n == n64 and np.array_equal(fold.cpu().numpy(), fold64.astype(np.float32))
        assert r <= tol
