"""Pins of the oracle's differentiable offset-class migration and velocity
gradient (orc_vgrad): the analytic gradient against FP64 central differences
of J at every node, the spline's partition of unity and linear reproduction,
the Catmull-Rom interpolant's reproduction of quadratics, and the zero-offset
identity I = F."""
import numpy as np
import pytest

import oracle
from random_amd import surveys
from random_amd.vgrad import linear_nodes


def tiny_survey():
    sv = surveys.line2d(nshots=6, nrecv_side=12, ns=201, ncols=60, col_dx=12.5, aperture=600.0,
                        nscat=4, nh=6, oh=6.25, dh=25.0, x0=400.0, t0_range=(0.1, 0.7))
    tr = surveys.traces_numpy(sv)
    return sv, oracle.filtered_traces_fp64(tr, sv.dt)


def test_gradient_matches_fp64_central_differences_at_every_node():
    sv, F = tiny_survey()
    V = linear_nodes(sv.ns, sv.dt, 8, 7, 1500.0, 400.0, scale=0.95).astype(np.float64)
    J, g, _ = oracle.vgrad(sv.cfg, sv.ns, sv.dt, F, sv.coords, V)
    assert J > 0
    h = 1e-3
    fd = np.zeros_like(g)
    for i in range(V.shape[0]):
        for j in range(V.shape[1]):
            Vp, Vm = V.copy(), V.copy()
            Vp[i, j] += h
            Vm[i, j] -= h
            fd[i, j] = (oracle.vgrad(sv.cfg, sv.ns, sv.dt, F, sv.coords, Vp, grad=False)[0] -
                        oracle.vgrad(sv.cfg, sv.ns, sv.dt, F, sv.coords, Vm, grad=False)[0]) / (2 * h)
    scale = np.abs(g).max()
    assert scale > 0
    # FP64 rounding of J (~1e-16 J) over 2h bounds the FD noise
    noise = 4e-16 * J / (2 * h)
    assert np.abs(fd - g).max() <= 1e-6 * scale + 10 * noise


def test_spline_partition_of_unity_and_linear_reproduction():
    ns, nxp = 301, 57
    V = np.full((9, 11), 1834.5)
    v = oracle.vgrad_velocity(ns, nxp, V)
    assert np.abs(v - 1834.5).max() < 1e-9
    dt = 0.004
    V = linear_nodes(ns, dt, 9, 11, 1500.0, 400.0).astype(np.float64)
    v = oracle.vgrad_velocity(ns, nxp, V)
    t = np.arange(ns) * dt
    # the nodes are float32; compare with the line they were rounded from
    assert np.abs(v - (1500.0 + 400.0 * t)[None, :]).max() < 1e-3


def _single_trace_cfg(nh=1):
    sv = surveys.line2d(nshots=1, nrecv_side=1, ns=120, ncols=40, col_dx=10.0, aperture=1000.0,
                        nscat=1, nh=nh, oh=0.0, dh=5000.0, x0=100.0)
    return sv


def test_zero_offset_trace_images_its_own_samples():
    sv = _single_trace_cfg()
    X = 15
    xs = int(round(sv.cfg.iline_distance * X * 100))         # centimetres, scale_xy = 0.01
    coords = np.array([[xs, 0, xs, 0]], np.int32)
    rng = np.random.default_rng(3)
    F = rng.standard_normal((1, sv.ns))
    V = np.full((6, 6), 2000.0)
    _, _, cig = oracle.vgrad(sv.cfg, sv.ns, sv.dt, F, coords, V, cig=True, grad=False)
    # q = 2 sqrt(t2) / dt = t: every sample is read back at its own index
    np.testing.assert_allclose(cig[X, 0], F[0], rtol=0, atol=1e-12)


def test_catmull_rom_reproduces_quadratics():
    sv = _single_trace_cfg()
    ns, dt = sv.ns, float(np.float32(sv.dt))
    xs, xg = 20000, 26000                                     # 200 m and 260 m
    coords = np.array([[xs, 0, xg, 0]], np.int32)
    k = np.arange(ns, dtype=np.float64)
    a, b, c = 0.3, -0.01, 2.5e-4
    F = (a + b * k + c * k * k)[None, :]
    v0 = 1800.0
    V = np.full((6, 6), v0)
    _, _, cig = oracle.vgrad(sv.cfg, ns, dt, F, coords, V, cig=True, grad=False)
    geo, _ = oracle.trace_geometry(sv.cfg, coords)
    colx = oracle.column_coords(sv.cfg, cig.shape[0])
    checked = 0
    for X in range(cig.shape[0]):
        IX = float(colx[X])
        if not abs(IX - float(geo[0, 4])) < sv.cfg.aperture:
            continue
        A = (IX - float(geo[0, 0])) ** 2
        B = (IX - float(geo[0, 2])) ** 2
        t2 = (0.5 * k * dt) ** 2
        q = (np.sqrt(t2 + A / v0 ** 2) + np.sqrt(t2 + B / v0 ** 2)) / dt
        ok = (q >= 1.0) & (q < ns - 2)       # all four taps inside the trace
        np.testing.assert_allclose(cig[X, 0][ok], (a + b * q + c * q * q)[ok], rtol=1e-10,
                                   atol=1e-12)
        checked += int(ok.sum())
    assert checked > 1000


def test_objective_is_zero_for_identical_classes_and_grows_with_a_difference():
    sv = _single_trace_cfg(nh=2)
    sv.cfg.dh = 50.0
    sv.cfg.oh = 0.0
    # two zero-offset-like traces in classes 0 and 1 at the same CMP
    coords = np.array([[40000, 0, 40000 + 1000, 0], [40000 - 2000, 0, 40000 + 3000, 0]], np.int32)
    geo, off = oracle.trace_geometry(sv.cfg, coords)
    assert list(off) == [0, 1]
    rng = np.random.default_rng(5)
    f = rng.standard_normal(sv.ns)
    V = np.full((6, 6), 1700.0)
    J0, _, cig = oracle.vgrad(sv.cfg, sv.ns, sv.dt, np.stack([f, f]), coords, V, cig=True,
                              grad=False)
    d = cig[:, 1] - cig[:, 0]
    assert J0 == pytest.approx(float((d * d).sum()), rel=1e-12)
    J1, _, _ = oracle.vgrad(sv.cfg, sv.ns, sv.dt, np.stack([f, 0 * f]), coords, V, grad=False)
    _, _, c1 = oracle.vgrad(sv.cfg, sv.ns, sv.dt, np.stack([f, 0 * f]), coords, V, cig=True,
                            grad=False)
    assert J1 == pytest.approx(float((c1[:, 0] ** 2).sum()), rel=1e-12)
