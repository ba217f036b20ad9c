"""Pins of the oracle's Hancock predictor-corrector (time_approximation = 1,
src/PatchMHDAM.cpp:858-1189), on the CPU.

What fixes it, independently of its code:
  - a uniform state is left unchanged (no spurious predictor terms);
  - on a periodic box without sources every conserved component is
    conserved to round-off (the corrector is in conservation form);
  - an entropy wave (density varying, velocity / pressure / field uniform,
    field along the flow) advected across a periodic box converges at
    second order in space and time: halving dx divides the L1 error by
    about 4, where the first-order time scheme with the same reconstruction
    only reaches about 2 and is clearly less accurate;
  - Brio-Wu: the plateaus of the shock tube at t = 0.1 as for RK2.
"""
import numpy as np
import pytest

import mfmhd
from helpers import periodic_fill, smooth_state, valid_view


def periodic_params(n, dims=3, eight=False, time=mfmhd.TIME_HANCOCK):
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.spacedim = dims
    p.time_approx = time
    p.eight_wave = 1 if eight else 0
    for d in range(3):
        p.domain_lo[d] = 0
        p.domain_hi[d] = n[d] - 1
        p.periodic[d] = 1
    return p


def test_uniform_state_unchanged(oracle_mod):
    n, ng = (8, 8, 8), 3
    p = periodic_params(n, eight=True)
    w = np.array([1.3, 0.2, -0.1, 0.05, 0.9, 0.4, 0.3, -0.2])
    u = np.zeros(8)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    u[5:8] = w[5:8]
    u[4] = w[4] / (p.gamma - 1) + 0.5 * w[0] * (w[1:4] ** 2).sum() + (w[5:8] ** 2).sum() / (8 * np.pi)
    U = np.broadcast_to(u[:, None, None, None], (8, 14, 14, 14)).copy()
    lo = (-ng,) * 3
    r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (7, 7, 7), 1e-3, 1.0 / 8)
    assert r["rc"] == 0
    v = valid_view(r["U"], lo, (0, 0, 0), (7, 7, 7))
    assert np.allclose(v, u[:, None, None, None], rtol=1e-14, atol=1e-15)


def test_periodic_conservation(oracle_mod):
    n, ng = (12, 12, 12), 3
    p = periodic_params(n, eight=False)
    lo = (-ng,) * 3
    shape = (8, 18, 18, 18)
    U = smooth_state(p, shape, lo, seed=4)
    periodic_fill(U, ng)
    dx = 1.0 / 12
    dt0, _ = oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (11, 11, 11), dx)
    before = valid_view(U, lo, (0, 0, 0), (11, 11, 11)).sum(axis=(1, 2, 3))
    for _ in range(3):
        r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (11, 11, 11), 0.3 * dt0, dx)
        assert r["rc"] == 0
        U = periodic_fill(r["U"], ng)
    after = valid_view(U, lo, (0, 0, 0), (11, 11, 11)).sum(axis=(1, 2, 3))
    scale = np.abs(valid_view(U, lo, (0, 0, 0), (11, 11, 11))).sum(axis=(1, 2, 3))
    assert np.all(np.abs(after - before) <= 1e-13 * scale), (after - before) / scale


def _entropy_wave_error(oracle_mod, nx, time):
    """L1 error of rho after advecting rho = 1 + 0.2 sin(2 pi x) at u = 1
    (p = 1, B = (0.5, 0, 0)) to t = 0.25 on a periodic nx x 4 x 4 box."""
    ng = 3
    n = (nx, 4, 4)
    p = periodic_params(n, time=time)
    dx = 1.0 / nx
    lo = (-ng,) * 3
    shape = (8, 4 + 2 * ng, 4 + 2 * ng, nx + 2 * ng)
    x = (np.arange(nx + 2 * ng) - ng + 0.5) * dx
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x)
    U = np.zeros(shape)
    U[0] = rho
    U[1] = rho * 1.0
    U[5] = 0.5
    U[4] = 1.0 / (p.gamma - 1) + 0.5 * rho + 0.25 / (8 * np.pi)
    periodic_fill(U, ng)
    t, T = 0.0, 0.25
    cfl = 0.4
    while t < T - 1e-14:
        dtr, _ = oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (nx - 1, 3, 3), dx)
        dt = min(cfl * dtr, T - t)
        r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (nx - 1, 3, 3), dt, dx, t)
        assert r["rc"] == 0
        U = periodic_fill(r["U"], ng)
        t += dt
    xc = (np.arange(nx) + 0.5) * dx
    exact = 1.0 + 0.2 * np.sin(2 * np.pi * (xc - T))
    got = U[0, ng, ng, ng:ng + nx]
    return np.mean(np.abs(got - exact))


def test_entropy_wave_second_order(oracle_mod):
    e32 = _entropy_wave_error(oracle_mod, 32, mfmhd.TIME_HANCOCK)
    e64 = _entropy_wave_error(oracle_mod, 64, mfmhd.TIME_HANCOCK)
    rate = np.log2(e32 / e64)
    assert rate > 1.6, (e32, e64, rate)
    f64 = _entropy_wave_error(oracle_mod, 64, mfmhd.TIME_FIRST_ORDER)
    assert e64 < 0.5 * f64, (e64, f64)


def test_brio_wu_hancock(oracle_mod):
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.time_approx = mfmhd.TIME_HANCOCK
    n, ng = 400, 3
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 1]
    p.spacedim = 2
    dx = 1.0 / n
    lo = (-ng, -ng, 0)
    shape = (8, 1, 1 + 2 * ng, n + 2 * ng)
    U = oracle_mod.init_problem(p, shape, lo, lo, (n - 1 + ng, ng, 0), dx)
    t = 0.0
    dt, _ = oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dx)
    dt *= 0.8
    while t < 0.1 - 1e-14:
        dt = min(dt, 0.1 - t)
        U[:, :, :, :] = U[:, :, ng:ng + 1, :]
        r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dt, dx, t)
        assert r["rc"] == 0
        U = r["U"]
        t += dt
        dt = 0.8 * r["dt"]
    rho = U[0, 0, ng, ng:ng + n]
    assert rho[int(0.05 * n)] == pytest.approx(1.0, abs=1e-10)
    assert rho[int(0.97 * n)] == pytest.approx(0.125, abs=1e-6)
    # compound wave plateau and the region right of the slow shock
    assert rho[int(0.45 * n)] == pytest.approx(0.676, rel=0.02)
    assert rho[int(0.75 * n)] == pytest.approx(0.117, rel=0.02)
    assert np.all(rho > 0.1) and np.all(rho <= 1.0 + 1e-12)
