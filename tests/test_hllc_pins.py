"""Pins of the oracle's HLLC flux of ideal MHD (flux = 21; S. Li, J. Comput.
Phys. 203 (2005) 344-357).  The reference has no HLLC solver, so nothing
here is reference parity: each check is a property the method fixes in
closed form, chosen so that a dropped or mis-signed term fails it --
consistency, upwinding beyond the outer waves, exact stationary and moving
contacts (which the HLL and Rusanov fluxes smear), mirror symmetry, the
GLM block with cleaning, rotation between directions, and the Brio-Wu
plateaus of the whole step.  CPU only."""
import numpy as np
import pytest

import mfmhd
from test_oracle_pins import D4, mhd_flux, params, rand_prim


def hllc(p, wl, wr, d, ch=1.0):
    p.flux = mfmhd.FLUX_HLLC
    return _oracle().flux_face(p, wl, wr, d, ch)


_ORACLE = []


def _oracle():
    return _ORACLE[0]


@pytest.fixture(autouse=True)
def _bind(oracle_mod):
    _ORACLE[:] = [oracle_mod]


def fast_speed(p, w, d):
    a2 = p.gamma * w[4] / w[0]
    va2 = D4 * np.dot(w[5:8], w[5:8]) / w[0]
    vn2 = D4 * w[5 + d] ** 2 / w[0]
    return np.sqrt(0.5 * (a2 + va2 + np.sqrt((a2 + va2) ** 2 - 4 * a2 * vn2)))


def test_consistency():
    """F(W, W) is the physical flux of W."""
    rng = np.random.default_rng(31)
    p = params()
    for t in range(60):
        d = t % 3
        w = rand_prim(rng)
        np.testing.assert_allclose(hllc(p, w, w, d), mhd_flux(p, w, d), rtol=1e-12, atol=1e-12)


def test_upwind_beyond_the_outer_waves():
    """Both states moving faster than their fast speeds: the flux is the
    upwind state's, to the left and to the right."""
    rng = np.random.default_rng(32)
    p = params()
    for t in range(30):
        d = t % 3
        wl, wr = rand_prim(rng), rand_prim(rng)
        cmax = max(fast_speed(p, wl, d), fast_speed(p, wr, d))
        wl[1 + d] = wr[1 + d] = cmax + 0.5
        np.testing.assert_allclose(hllc(p, wl, wr, d), mhd_flux(p, wl, d), rtol=1e-12, atol=1e-12)
        wl[1 + d] = wr[1 + d] = -(cmax + 0.5)
        np.testing.assert_allclose(hllc(p, wl, wr, d), mhd_flux(p, wr, d), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("un", [0.0, 0.3, -0.4])
def test_contact_is_exact(un):
    """A contact discontinuity (density jump only; velocity, pressure and
    field equal, B_n != 0) moving at u_n: the flux is the upwind state's
    physical flux, the mass flux rho_upwind * u_n exactly up to round-off;
    the HLL-type Rusanov flux smears it."""
    rng = np.random.default_rng(33)
    p = params()
    for t in range(24):
        d = t % 3
        w = rand_prim(rng)
        w[1 + d] = un
        wl, wr = w.copy(), w.copy()
        wl[0], wr[0] = 1.7, 0.4
        f = hllc(p, wl, wr, d)
        up = wl if un >= 0 else wr
        scale = np.max(np.abs(mhd_flux(p, up, d))) + 1.0
        np.testing.assert_allclose(f, mhd_flux(p, up, d), rtol=0, atol=1e-13 * scale)
        p.flux = mfmhd.FLUX_RUSANOV
        fr = _oracle().flux_face(p, wl, wr, d)
        assert abs(fr[0] - up[0] * un) > 1e-3


def test_mirror_symmetry():
    """Reflecting the Riemann problem in the face plane -- the states
    swapped, their normal velocity and normal field negated -- reflects
    the flux: the mass and energy fluxes change sign, the normal momentum
    flux keeps it (a wrong sign in one side's star state breaks this)."""
    rng = np.random.default_rng(34)
    p = params()
    for t in range(30):
        d = t % 3
        wl, wr = rand_prim(rng), rand_prim(rng)
        f = hllc(p, wl, wr, d)

        def mirror(w):
            m = w.copy()
            m[1 + d] = -m[1 + d]
            m[5 + d] = -m[5 + d]
            return m
        g = hllc(p, mirror(wr), mirror(wl), d)
        scale = np.max(np.abs(f)) + 1.0
        assert abs(f[1 + d] - g[1 + d]) <= 1e-12 * scale
        assert abs(f[0] + g[0]) <= 1e-12 * scale
        assert abs(f[4] + g[4]) <= 1e-12 * scale


def test_rotation_between_directions():
    """The flux of direction d is the x flux of the rotated states."""
    rng = np.random.default_rng(35)
    p = params()
    for _ in range(20):
        wl, wr = rand_prim(rng), rand_prim(rng)
        fx = hllc(p, wl, wr, 0)
        for d in (1, 2):
            perm = [(0 + d) % 3, (1 + d) % 3, (2 + d) % 3]
            rl, rr = wl.copy(), wr.copy()
            for k in range(3):
                rl[1 + perm[k]], rl[5 + perm[k]] = wl[1 + k], wl[5 + k]
                rr[1 + perm[k]], rr[5 + perm[k]] = wr[1 + k], wr[5 + k]
            fd = hllc(p, rl, rr, d)
            for k in range(3):
                assert fd[1 + perm[k]] == pytest.approx(fx[1 + k], rel=1e-13, abs=1e-14)
                assert fd[5 + perm[k]] == pytest.approx(fx[5 + k], rel=1e-13, abs=1e-14)
            assert fd[0] == pytest.approx(fx[0], rel=1e-13, abs=1e-14)
            assert fd[4] == pytest.approx(fx[4], rel=1e-13, abs=1e-14)


def test_glm_block_replaces_normal_field_flux():
    """flux 21 with GLM cleaning: the (B_n, psi) fluxes are DEDNER2X2F's
    closed forms, the other seven HLLC's."""
    rng = np.random.default_rng(36)
    p = params(dedner=1)
    q = params()
    ch = 1.3
    for d in range(3):
        wl = np.append(rand_prim(rng), 0.2)
        wr = np.append(rand_prim(rng), -0.3)
        f = hllc(p, wl, wr, d, ch)
        assert f[5 + d] == pytest.approx(0.5 * (wl[8] + wr[8]) + 0.5 * ch * (wl[5 + d] - wr[5 + d]),
                                         rel=1e-13)
        assert f[8] == pytest.approx(0.5 * ch * ch * (wl[5 + d] + wr[5 + d]) +
                                     0.5 * ch * (wl[8] - wr[8]), rel=1e-13)
        g = hllc(q, wl[:8], wr[:8], d)
        mask = [c for c in range(8) if c != 5 + d]
        np.testing.assert_allclose(f[mask], g[mask], rtol=1e-14, atol=1e-15)


def test_brio_wu_plateaus_with_hllc(oracle_mod):
    """The Brio-Wu tube to t = 0.1 on 400 cells with the HLLC flux reaches
    the plateaus of the exact solution (compound-wave region rho ~ 0.676,
    right of the slow shock rho ~ 0.117) and keeps the end states."""
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.time_approx = mfmhd.TIME_RK2
    p.flux = mfmhd.FLUX_HLLC
    n, ng = 400, 4
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 1]
    p.spacedim = 2
    dx = 1.0 / n
    lo = (-ng, -ng, 0)
    U = oracle_mod.init_problem(p, (8, 1, 1 + 2 * ng, n + 2 * ng), lo, lo, (n - 1 + ng, ng, 0), dx)
    t = 0.0
    dt = 0.8 * oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dx)[0]
    for _ in range(1000):   # about 270 steps
        if t >= 0.1 - 1e-14:
            break
        dt = min(dt, 0.1 - t)
        U[:, :, :, :] = U[:, :, ng:ng + 1, :]
        r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dt, dx, t)
        assert r["rc"] == 0 and r["dt"] > 1e-6, (t, r["dt"])
        U = r["U"]
        t += dt
        dt = 0.8 * r["dt"]
    assert t >= 0.1 - 1e-14
    rho = U[0, 0, ng, ng:ng + n]
    assert rho[int(0.05 * n)] == pytest.approx(1.0, abs=1e-10)
    assert rho[int(0.97 * n)] == pytest.approx(0.125, abs=1e-6)
    assert rho[int(0.45 * n)] == pytest.approx(0.676, rel=0.02)
    assert rho[int(0.75 * n)] == pytest.approx(0.117, rel=0.02)
    assert np.all(rho > 0.1) and np.all(rho <= 1.0 + 1e-12)
