"""ktm_migrate_block_static (the allocation-free entry point) against
ktm_migrate_block through the C ABI: BF16 and FP16 storage (decoded into
the library's static device array) and FP32, two calls in a row with the
second on fewer traces, bit for bit; a block beyond the array's capacity is
refused before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

from random_amd import (ACCUM_KAHAN, ACCUM_PLAIN, INDEX_FAST, PREP_CLOSED, PREP_REFERENCE,
                        TRACE_BF16, TRACE_F16, TRACE_F32, _lib, surveys)
from random_amd.ktm import Plan, _stream

pytestmark = pytest.mark.gpu

KTM_ERR_UNSUPPORTED = -7
STAGE_FLOATS = 16 << 20


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _call(fn, plan, prepared, dtype, mode, geom, accum, km, fold, comp, counter):
    return fn(plan._h, _p(prepared), dtype, mode, _p(geom), geom.shape[0], accum, -1, _p(km),
              _p(fold), _p(comp), _p(counter), _stream())


@pytest.mark.parametrize("dtype,mode,accum", [
    (TRACE_BF16, PREP_CLOSED, ACCUM_KAHAN),
    (TRACE_F16, PREP_CLOSED, ACCUM_PLAIN | INDEX_FAST),
    (TRACE_F32, PREP_REFERENCE, ACCUM_PLAIN),
], ids=["bf16-kahan", "fp16-fast", "fp32-reference"])
def test_static_entry_point_matches_migrate_block(cuda, dtype, mode, accum):
    L = _lib.lib()
    sv = surveys.area3d(ntr=6000, ns=700, ncols=40, extent=1500.0, max_offset=2000.0,
                        aperture=1200.0, nscat=10, jobs_x=2, jobs_y=2)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    tr = torch.from_numpy(surveys.traces_numpy(sv)).to(cuda)
    co = torch.from_numpy(sv.coords).to(cuda)
    kahan = (accum & ~INDEX_FAST) == ACCUM_KAHAN
    for n in (6000, 2500):                    # the second call on fewer traces
        print(f"case: dtype {dtype}, mode {mode}, accum {accum:#x}, {n} traces", flush=True)
        prepared, geom = plan.prep(tr[:n], co[:n], mode, dtype)
        out = []
        for fn in (L.ktm_migrate_block, L.ktm_migrate_block_static):
            km, fold, comp = plan.new_image(kahan=kahan, device=cuda)
            counter = torch.zeros(1, dtype=torch.int64, device=cuda)
            _lib.check(_call(fn, plan, prepared, dtype, mode, geom, accum, km, fold, comp, counter),
                       "migrate")
            torch.cuda.synchronize()
            out.append((km.cpu().numpy(), fold.cpu().numpy(), int(counter.item()),
                        None if comp is None else comp.cpu().numpy()))
        (k0, f0, n0, c0), (k1, f1, n1, c1) = out
        assert n0 == n1 > 0
        assert np.array_equal(k0.view(np.uint32), k1.view(np.uint32))
        assert np.array_equal(f0, f1)
        if c0 is not None:
            assert np.array_equal(c0.view(np.uint32), c1.view(np.uint32))


def test_static_entry_point_refuses_blocks_beyond_its_array(cuda):
    L = _lib.lib()
    sv = surveys.line2d(nshots=2, ncols=32, ns=1001, nscat=2)
    plan = Plan(sv.cfg, sv.ns, sv.dt, sv.vel)
    cap = STAGE_FLOATS // plan.row
    km, fold, comp = plan.new_image(device=cuda)
    geom = torch.zeros((cap + 1, 8), dtype=torch.float32, device=cuda)
    prepared = torch.zeros(8, dtype=torch.uint16, device=cuda)
    st = L.ktm_migrate_block_static(plan._h, _p(prepared), TRACE_BF16, PREP_CLOSED, _p(geom),
                                    cap + 1, ACCUM_PLAIN, -1, _p(km), _p(fold), None, None,
                                    _stream())
    assert st == KTM_ERR_UNSUPPORTED
