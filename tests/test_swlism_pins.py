"""Pins of the oracle's 2-D axisymmetric SW/LISM problem (the ensemble
workload: swlism/SWLISMProblem.cpp, swlism/SWLISMF.ChF, the coord_sys = 1
sources of srcIdeal/PatchIdealMHDF.ChF and srcIdeal/PatchMHDMFF.ChF), on the
CPU, from closed forms and symmetries rather than from the code:
  - SETSWLISM's non-dimensionalisation (physical constants of
    srcIdeal/EosCommon.fh);
  - the geometric source of every conserved variable is the extra term of
    the divergence in cylindrical coordinates, -(1/y) F_y, with the pressure
    left out of the radial momentum (it enters as a gradient) -- written here
    from the ideal-MHD flux, for swirl-free states (w = 0, B_z = 0, as the
    reference's source assumes);
  - a uniform axial flow with an axial field is a steady solution of the
    axisymmetric equations;
  - the LISM state of the ghost cells and of the initial data outside initR
    is the potential flow past a sphere: no radial velocity on the sphere and
    no divergence in cylindrical coordinates;
  - the solar wind inside R0 is radial with rho r^2 and p r^(2 gamma) fixed;
  - the mirror image at the symmetry axis carries no mass, axial momentum or
    energy through it;
  - the supersonic branch of the neutral outflow flux is the upwind Euler
    flux; the charge exchange of the problem conserves neutral mass, plasma +
    neutral momentum and energy, and the born population gets the others'
    losses."""
import numpy as np
import pytest

import mfmhd
from helpers import grid_shape, valid_view

MP, KB = 1.67262171e-24, 1.3806505e-16
PI4 = 4 * np.pi


def small_swlism(nx=24, ny=16, nf=2, time=mfmhd.TIME_RK2, dx=10.0):
    p = mfmhd.swlism_params(sunXY=(0.4 * nx * dx, 0.0), R0=2.6 * dx, initR=6.0 * dx,
                            dx=dx, num_cells=(nx, ny), nfluids=nf, lismB=3e-6)
    p.time_approx = time
    return p


def test_swlism_setter_nondimensionalisation():
    dx = 1500.0 / 96
    p = mfmhd.swlism_params(lismN=0.06, lismV=2.64e6, lismT=6527.0, lismB=1.5e-6,
                            sunXY=(600.0, 0.0), R0=10.0, sunN=7.4, sunV=4.5e7,
                            sunT=51100.0, initR=80.0, netN=0.1, dx=dx, num_cells=(96, 64))
    pref = 0.06 * MP * 2.64e6 ** 2
    lismP = 2 * KB * 0.06 * 6527.0 / pref
    assert p.problem == mfmhd.PROBLEM_SWLISM and p.coord_sys == mfmhd.COORD_AXISYMMETRIC
    assert p.spacedim == 2 and tuple(p.domain_hi[:2]) == (95, 63)
    assert p.lism_p == pytest.approx(lismP, rel=1e-14)
    assert p.lism_m == pytest.approx(1 / np.sqrt(5 / 3 * lismP), rel=1e-14)
    assert p.lism_bx == pytest.approx(1.5e-6 / np.sqrt(pref), rel=1e-14)
    assert (p.lism_ux, p.lism_uy, p.lism_uz, p.lism_by, p.lism_bz) == (1, 0, 0, 0, 0)
    assert p.sun_rho == pytest.approx(7.4 / 0.06, rel=1e-15)
    assert p.sun_p == pytest.approx(2 * KB * 7.4 * 51100.0 / pref, rel=1e-14)
    assert p.sun_v == pytest.approx(4.5e7 / 2.64e6, rel=1e-15)
    assert p.net_rho == pytest.approx(0.1 / 0.06, rel=1e-15)
    assert p.net_p == pytest.approx(KB * 0.1 * 6527.0 / pref, rel=1e-14)
    # the sun on the nearest node: 600 / 15.625 = 38.4 -> node 38
    assert p.sun_xc == pytest.approx(38 * dx) and p.sun_yc == 0.0
    assert p.r0_dt_ignore == pytest.approx(10.0 - dx)
    assert p.smallr == 1e-8
    assert p.smallp == pytest.approx(1e-8 * 0.5 / p.lism_m ** 2, rel=1e-14)
    assert p.ce_sigma_lindsay == pytest.approx(1e-16 * 1.5e13 * 0.06, rel=1e-15)


def mhd_flux_y(w, gamma):
    """Ideal-MHD flux in y (Gaussian units) of one fluid (plasma: 8 comps)."""
    rho, u, v, ww, p, bx, by, bz = w[:8]
    b2 = bx * bx + by * by + bz * bz
    E = p / (gamma - 1) + 0.5 * rho * (u * u + v * v + ww * ww) + b2 / (2 * PI4)
    pt = p + b2 / (2 * PI4)
    ub = u * bx + v * by + ww * bz
    return np.array([rho * v, rho * u * v - bx * by / PI4, rho * v * v + pt - by * by / PI4,
                     rho * v * ww - bz * by / PI4, v * (E + pt) - by * ub / PI4,
                     v * bx - u * by, 0.0, v * bz - ww * by]), pt


@pytest.mark.parametrize("nf", [1, 2, 3])
def test_axisymmetric_source_is_the_cylindrical_divergence(oracle_mod, nf):
    p = small_swlism(nf=nf)
    g = p.gamma
    rng = np.random.default_rng(nf)
    dt, dx = 0.013, 0.7
    for _ in range(30):
        w = [rng.uniform(0.3, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0,
             rng.uniform(0.1, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0]
        for f in range(nf - 1):
            w += [rng.uniform(0.1, 1), rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0,
                  rng.uniform(0.05, 1)]
        w = np.array(w)
        j = int(rng.integers(0, 40))
        y = (j + 0.5) * dx
        S = oracle_mod.axisym_source(p, w, j, dt, dx)
        Fy, pt = mhd_flux_y(w, g)
        Fy[2] -= pt  # the pressure enters the radial momentum as a gradient
        want = -dt / y * Fy
        assert np.allclose(S[:8], want, rtol=1e-13, atol=1e-15 * np.max(np.abs(want)))
        for f in range(nf - 1):
            s = 8 + 5 * f
            r, u, v, ww, pn = w[s:s + 5]
            E = pn / (g - 1) + 0.5 * r * (u * u + v * v)
            wantn = -dt / y * np.array([r * v, r * u * v, r * v * v, 0.0, v * (E + pn)])
            assert np.allclose(S[s:s + 5], wantn, rtol=1e-13,
                               atol=1e-15 * np.max(np.abs(wantn)))


def test_uniform_axial_flow_is_steady(oracle_mod):
    """Caller-filled ghosts (problem NONE), a uniform flow along the axis
    with an axial field: no flux difference and no geometric source."""
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.spacedim = 2
    p.coord_sys = mfmhd.COORD_AXISYMMETRIC
    p.nfluids = 2
    p.time_approx = mfmhd.TIME_RK2
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 9, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    w = np.array([1.3, 0.7, 0.0, 0.0, 0.9, 1.1, 0.0, 0.0, 0.4, 0.2, 0.0, 0.0, 0.3])
    u = np.zeros(nc)
    u[0], u[1:4], u[5:8] = w[0], w[0] * w[1:4], w[5:8]
    u[4] = w[4] / (2 / 3) + 0.5 * w[0] * w[1] ** 2 + (w[5] ** 2) / (2 * PI4)
    u[8], u[9] = w[8], w[8] * w[9]
    u[12] = w[12] / (2 / 3) + 0.5 * w[8] * w[9] ** 2
    U = np.broadcast_to(u[:, None, None, None], shape).copy()
    r = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.01, 0.1, 0.0)
    assert r["rc"] == 0
    V = valid_view(r["U"], lo, vlo, vhi)
    assert np.allclose(V, u[:, None, None, None], rtol=1e-14, atol=1e-15)


def test_lism_state_is_potential_flow_past_a_sphere(oracle_mod):
    """The initial state outside initR (SWLISMINIT) and the LISM ghost state
    (LISMINIT) are the incompressible potential flow past a sphere of radius
    initR in the axisymmetric (x, y) plane: u_r = 0 on the sphere and
    du/dx + (1/y) d(y v)/dy = 0."""
    p = small_swlism(nx=40, ny=30, nf=1, dx=1.0)
    p.init_r = 6.0
    n0, n1 = 40, 30
    lo, hi = (0, 0, 0), (n0 - 1, n1 - 1, 0)
    U = oracle_mod.init_problem(p, (8, 1, n1, n0), lo, lo, hi, 1.0)
    x = (np.arange(n0) + 0.5) - p.sun_xc
    y = (np.arange(n1) + 0.5) - p.sun_yc
    X, Y = np.meshgrid(x, y)
    R = np.sqrt(X * X + Y * Y)
    ux, uy = U[1, 0] / U[0, 0], U[2, 0] / U[0, 0]
    # analytic potential flow with the same radius, checked where r > initR
    m = R > 6.0 + 1.5
    h = (6.0 / R) ** 3
    st, ct = Y / R, X / R
    ur = ux * ct + uy * st
    assert np.allclose(ur[m], (ct * (1 - h))[m], rtol=0, atol=1e-12)
    # cylindrical divergence by central differences falls with the grid
    d = (ux[1:-1, 2:] - ux[1:-1, :-2]) / 2 + \
        (Y[2:, 1:-1] * uy[2:, 1:-1] - Y[:-2, 1:-1] * uy[:-2, 1:-1]) / (2 * Y[1:-1, 1:-1])
    mm = R[1:-1, 1:-1] > 6.0 + 3
    assert np.max(np.abs(d[mm])) < 0.02


def test_solar_wind_inside_r0(oracle_mod):
    p = small_swlism(nx=32, ny=24, nf=3, time=mfmhd.TIME_FIRST_ORDER, dx=1.0)
    p.sun_r0 = 5.5
    p.init_r = 8.0
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (31, 23, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 2, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, 1.0)
    r = oracle_mod.patch_update(p, U, lo, vlo, vhi, 1e-6, 1.0, 0.0)
    assert r["rc"] == 0
    V = valid_view(r["U"], lo, vlo, vhi)[:, 0]
    x = (np.arange(32) + 0.5) - p.sun_xc
    y = (np.arange(24) + 0.5) - p.sun_yc
    X, Y = np.meshgrid(x, y)
    R = np.sqrt(X * X + Y * Y)
    m = R <= 5.5
    assert m.sum() > 10
    g = p.gamma
    rho = V[0][m]
    assert np.allclose(rho * R[m] ** 2, p.sun_rho, rtol=1e-13)
    vx, vy = V[1][m] / rho, V[2][m] / rho
    assert np.allclose(vx, p.sun_v * X[m] / R[m], rtol=1e-12)
    assert np.allclose(vy, p.sun_v * Y[m] / R[m], rtol=1e-12, atol=1e-14)
    pr = (g - 1) * (V[4][m] - 0.5 * rho * p.sun_v ** 2)
    assert np.allclose(pr * R[m] ** (2 * g), p.sun_p, rtol=1e-9)
    assert np.all(V[5:8][:, m] == 0)
    # population 2 back to its floor state, population 1 untouched by the re-init
    assert np.allclose(V[13][m], 1e-6) and np.allclose(V[14:16][:, m], 0)


def test_axis_carries_no_mass_or_axial_momentum(oracle_mod):
    """At the symmetry axis the ghost cells mirror the interior with v and
    B_y reversed: the Riemann problem there is symmetric, so no mass, axial
    momentum, axial field or energy crosses it, and no normal field."""
    p = small_swlism(nx=20, ny=12, nf=2, dx=8.0)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (19, 11, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, 8.0)
    rng = np.random.default_rng(5)
    U[:, :, 4:, 4:-4] *= 1 + 0.05 * rng.uniform(-1, 1, U[:, :, 4:, 4:-4].shape)
    r = oracle_mod.patch_update(p, U, lo, vlo, vhi, 1e-3, 8.0, 0.0, want_F=True)
    assert r["rc"] == 0
    Fy = r["F"][1][:, 0, 0, :]  # faces j = 0: the axis
    scale = np.max(np.abs(r["F"][1]), axis=(1, 2, 3))
    for c in (0, 1, 4, 5, 8, 9, 12):
        assert np.max(np.abs(Fy[c])) <= 1e-13 * scale[c], c


def test_bfluxgas_supersonic_outflow_is_upwind(oracle_mod):
    """On the right boundary a neutral population leaving faster than its
    sound speed gets the Euler flux of its inside state (BFLUXGAS)."""
    p = small_swlism(nx=12, ny=8, nf=2, time=mfmhd.TIME_FIRST_ORDER, dx=5.0)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 7, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 2, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, 5.0)
    # a fast, cold population 1 everywhere (Mach ~ 5 along +x)
    rn, un, pn = 0.3, 3.0, 0.01
    U[8], U[9], U[10], U[11] = rn, rn * un, 0.0, 0.0
    U[12] = pn / (2 / 3) + 0.5 * rn * un ** 2
    r = oracle_mod.patch_update(p, U, lo, vlo, vhi, 1e-4, 5.0, 0.0, want_F=True)
    F = r["F"][0][:, 0, :, -1]  # the right boundary faces
    want = np.array([rn * un, rn * un * un + pn, 0.0, 0.0,
                     un * (pn / (2 / 3) + 0.5 * rn * un ** 2 + pn)])
    assert np.allclose(F[8:13].T, want, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("nf", [2, 3, 4])
@pytest.mark.parametrize("reg", [1, 2, 3, 4])
def test_charge_exchange_mf_balance(oracle_mod, nf, reg):
    """CHARGE_EXCHANGE_MF (2 and 3 fluids) and CHARGE_EXCHANGE_4F (4 fluids,
    swlism/SWLISMProblem.cpp:945-1021): every population not born in the
    cell's region loses mass, and the one born there (MF: REG == k; 4F:
    population 1 also in region 4) gains exactly the others' losses, so
    neutral mass, plasma + neutral momentum and energy are conserved where a
    population is born; where none is (one neutral outside region 1), the
    neutral only loses mass."""
    p = small_swlism(nf=nf)
    rng = np.random.default_rng(10 * nf + reg)
    born = reg - 1 if not (nf == 4 and reg == 4) else 0
    if nf != 4 and reg == 4:
        born = 3
    for _ in range(20):
        w = [rng.uniform(0.1, 2), *rng.uniform(-2, 2, 3), rng.uniform(0.05, 1), 0.2, -0.1, 0.0]
        for f in range(nf - 1):
            w += [rng.uniform(0.05, 0.5), *rng.uniform(-1.5, 1.5, 3), rng.uniform(0.01, 0.3)]
        w = np.array(w)
        S = oracle_mod.charge_exchange(p, w, reg, 0.01)
        rho = S[8::5][: nf - 1]
        if born < nf - 1:
            for c in range(1, 5):
                tot = S[c] + S[8 + c::5][: nf - 1].sum()
                assert abs(tot) <= 1e-12 * np.max(np.abs(S))
        assert S[0] == 0 and np.all(S[5:8] == 0)
        lost = [k for k in range(nf - 1) if k != born]
        assert all(rho[k] < 0 for k in lost)
        if born < nf - 1:
            assert rho[born] == pytest.approx(-sum(rho[k] for k in lost), rel=1e-13, abs=0)
            assert abs(rho.sum()) <= 1e-13 * np.max(np.abs(rho))
        else:
            assert rho.sum() < 0


def test_swlism_dt_excludes_the_inner_sphere(oracle_mod):
    p = small_swlism(nx=20, ny=14, nf=2, dx=4.0)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (19, 13, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, 4.0)
    dt, cell = oracle_mod.compute_dt(p, U, lo, vlo, vhi, 4.0)
    best, bc = np.inf, None
    for j in range(14):
        for i in range(20):
            x, y = (i + 0.5) * 4.0 - p.sun_xc, (j + 0.5) * 4.0 - p.sun_yc
            if x * x + y * y <= p.r0_dt_ignore ** 2:
                continue
            u = U[:, 0, j + 4, i + 4]
            d = 4.0 / oracle_mod.max_wave_speed(p, u)
            if d < best:
                best, bc = d, (i, j, 0)
    assert dt == best and tuple(cell) == bc


@pytest.mark.parametrize("time", [mfmhd.TIME_FIRST_ORDER, mfmhd.TIME_RK2])
def test_swlism_steps_stay_positive(oracle_mod, time):
    """Twelve steps from SWLISMINIT at the deck's initial_cfl and cfl with
    Roe-averaged fluxes: the density stays at the wind's level and dt does
    not collapse.  (The deck's 8-wave Roe on arithmetic means loses
    positivity at the wind / LISM interface of this initial state within a
    few steps; the deck itself restarts from a developed checkpoint.)"""
    dx = 1500.0 / 48
    p = mfmhd.swlism_params(sunXY=(19.2 * dx, 0.0), R0=3.1 * dx, initR=8.0 * dx,
                            dx=dx, num_cells=(48, 32), nfluids=2, lismB=3e-6)
    p.time_approx = time
    p.flux = mfmhd.FLUX_ROE8_ROEAVG
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (47, 31, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dt0, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    rho0 = valid_view(U, lo, vlo, vhi)[0].min()
    dt = 0.1 * dt0
    for _ in range(12):
        r = oracle_mod.patch_update(p, U, lo, vlo, vhi, dt, dx, 0.0)
        assert r["rc"] == 0
        U = r["U"]
        assert valid_view(U, lo, vlo, vhi)[0].min() > 0.5 * rho0
        assert r["dt"] > 0.5 * dt0
        dt = min(0.25 * r["dt"], 1.02 * dt)
