"""The numerical assumptions behind the fast traveltime path of k_migrate.

The kernel computes q = dsr/dt with FMA and the hardware square root on
radicands pre-scaled by 1/dt^2, and falls back to the reference arithmetic of
Kirchhoff.c:1381-1384 whenever an integer lies within 4.8e-7 (q + 1) of q.
That is exact only if |q_fast - q_ref| <= 4.8e-7 (q_ref + 1), which rests on
v_sqrt_f32 being within one ulp.  Both facts are checked here.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from random_amd import _lib

pytestmark = pytest.mark.gpu

EPS = 4.8e-7   # the kernel's tolerance of its fast form (8 x 2^-24; bound 7 x 2^-24)


def _p(t):
    return C.c_void_p(t.data_ptr())


def ulp_diff(a, b):
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_hw_sqrt_within_one_ulp_exhaustive(cuda):
    # every float in [1, 4): sqrt(4^k x) = 2^k sqrt(x), so this covers every
    # mantissa of both exponent parities
    print("case: v_sqrt_f32 on all 2^24 floats of [1, 4)", flush=True)
    x = np.arange(1 << 24, dtype=np.int64) + np.float32(1.0).view(np.int32)
    x = x.astype(np.int32).view(np.float32)
    assert x[0] == 1.0 and x[-1] < 4.0 and x[-1] > 3.99
    d_x = torch.from_numpy(x).to(cuda)
    out = torch.empty_like(d_x)
    _lib.check(_lib.lib().ktm_selftest_sqrt(_p(d_x), x.size, _p(out), None), "selftest_sqrt")
    hw = out.cpu().numpy()
    ref = np.sqrt(x)  # correctly rounded
    u = ulp_diff(hw, ref)
    print(f"  max ulp error {u.max()}, fraction not correctly rounded {np.mean(u > 0):.3e}")
    assert u.max() <= 1


def _run_traveltime(cuda, inp, dt, ns):
    d_in = torch.from_numpy(np.ascontiguousarray(inp, np.float32)).to(cuda)
    n = inp.shape[0]
    qf = torch.empty(n, dtype=torch.float32, device=cuda)
    qr = torch.empty_like(qf)
    tk = torch.empty(n, dtype=torch.int32, device=cuda)
    tr = torch.empty_like(tk)
    _lib.check(_lib.lib().ktm_selftest_traveltime(_p(d_in), n, dt, ns, _p(qf), _p(qr), _p(tk),
                                                  _p(tr), None), "selftest_traveltime")
    return qf.cpu().numpy(), qr.cpu().numpy(), tk.cpu().numpy(), tr.cpu().numpy()


def test_fast_traveltime_bound_and_index(cuda):
    print("case: fast vs reference traveltime, 2^24 random samples", flush=True)
    rng = np.random.default_rng(7)
    n = 1 << 24
    dt = np.float32(0.004)
    ns = 1500
    t = rng.integers(0, ns, n)
    d = rng.uniform(0.0, 8000.0, (n, 2)) ** 2
    v = rng.uniform(1400.0, 5000.0, n).astype(np.float32)
    inp = np.stack([d[:, 0], d[:, 1], t, v * v], axis=1).astype(np.float32)
    qf, qr, tk, tr = _run_traveltime(cuda, inp, dt, ns)
    err = np.abs(qf.astype(np.float64) - qr) / (qr.astype(np.float64) + 1.0)
    near = np.abs(qr - np.rint(qr)) <= EPS * (qr + 1.0)
    print(f"  max |q_fast - q_ref| / (q_ref + 1) = {err.max():.3e} (bound {EPS}); "
          f"{near.sum()} samples within the bound of an integer; T mismatches {np.sum(tk != tr)}")
    assert err.max() <= EPS
    assert np.array_equal(tk, tr)
    # the T the kernel keeps is floor of the reference quotient
    ref_T = np.where(qr < ns, np.floor(qr), ns).astype(np.int32)
    assert np.array_equal(tr, ref_T)


def test_fast_traveltime_near_integers(cuda):
    """Samples constructed to land within ~1e-5 samples of an integer (the
    cases the fallback exists for): the kernel's index equals the
    reference's on all of them."""
    print("case: fast vs reference traveltime, 2^22 near-integer samples", flush=True)
    rng = np.random.default_rng(11)
    n = 1 << 22
    dt = np.float32(0.004)
    ns = 1500
    t = rng.integers(0, ns, n)
    t2 = 0.25 * t * t * np.float64(dt) ** 2
    v = rng.uniform(1400.0, 5000.0, n).astype(np.float32)
    v2 = (v * v).astype(np.float64)
    d1 = rng.uniform(0.0, 4000.0, n) ** 2
    s1 = np.sqrt(t2 + d1 / v2)
    target = np.floor(2.0 * s1 / np.float64(dt)) + rng.integers(1, 400, n) + rng.uniform(-1e-5, 1e-5, n)
    rest = target * np.float64(dt) - s1
    d2 = (rest * rest - t2) * v2
    ok = (rest > 0) & (d2 >= 0)
    inp = np.stack([d1[ok], d2[ok], t[ok], v2[ok]], axis=1).astype(np.float32)
    qf, qr, tk, tr = _run_traveltime(cuda, inp, dt, ns)
    dist = np.abs(qr - np.rint(qr))
    err = np.abs(qf.astype(np.float64) - qr) / (qr.astype(np.float64) + 1.0)
    print(f"  {ok.sum()} samples, {np.sum(dist < 1e-4)} within 1e-4 of an integer, "
          f"max rel err {err.max():.3e}, T mismatches {np.sum(tk != tr)}")
    assert np.sum(dist < 1e-4) > ok.sum() // 4
    assert err.max() <= EPS
    assert np.array_equal(tk, tr)
