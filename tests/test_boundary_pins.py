"""Pins of the oracle's characteristic boundary fluxes of the heliospheric and
SW/LISM problems -- BFLUX / BFLUXR (swlism/HeliosphericF.ChF:1363-1774) for
the plasma and BFLUXGAS (swlism/SWLISMF.ChF:1434-1553) for a neutral
population -- against what gas dynamics and MHD fix for them:

- supersonic (super-fast) out- or inflow: the physical flux of the upwind
  state;
- BFLUXGAS's sub-sonic branches: the sonic state reached along the
  characteristic that leaves the domain (out) or enters it (in), i.e. the
  Riemann invariant u +- 2a/(gamma - 1) and the entropy p / rho^gamma of
  the state it starts from, with u = +-a: solved here by a root finder
  from those three conditions, not from the reference's closed form;
- BFLUX's sub-fast outflow branch: a state moved from the interior along
  the right eigenvector of the u - c_f wave (from the ideal-MHD Jacobian
  written from the equations, Gaussian units) until it is fast-sonic,
  linearised in rho: the state agrees with W + h r to first order in
  c_f - u (h from the linearised sonic condition), its u - c_f vanishes
  to second order, and the flux is that state's physical flux;
- the sub-fast inflow branch: the 8-wave Roe flux of (interior, exterior).

CPU only."""
import numpy as np
import pytest
from scipy.optimize import fsolve

import mfmhd
from test_oracle_closed_forms import PI4, mhd_flux, prim_jacobian
from test_oracle_pins import euler_flux

G = 5.0 / 3.0


def plasma_params():
    p = mfmhd.orszag_params(G)
    p.flux = mfmhd.FLUX_ROE8
    return p


def fast_speed(w):
    rho, _, _, _, pr, bx, by, bz = w
    a2 = G * pr / rho
    b2 = (bx * bx + by * by + bz * bz) / (PI4 * rho)
    return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bx * bx / (PI4 * rho))))


STATES = [  # (rho, vn, vt1, vt2, p, Bn, Bt1, Bt2); vn is set per case
    np.array([1.3, 0.0, 0.2, -0.1, 0.9, 1.1, 0.8, -0.5]),
    np.array([0.7, 0.0, 0.3, 0.0, 1.6, -2.2, 0.4, 1.3]),
    np.array([2.1, 0.0, 0.1, -0.3, 0.4, 0.6, -1.7, 0.2]),
]


@pytest.mark.parametrize("radial", [False, True], ids=["BFLUX", "BFLUXR"])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_bflux_super_fast(oracle_mod, radial, k):
    """u >= c_f: the interior state's flux and normal field; u < -c_f: the
    exterior state's."""
    p = plasma_params()
    w = STATES[k].copy()
    out = STATES[(k + 1) % 3].copy()
    w[1] = 1.2 * fast_speed(w)
    f, bn, ws = oracle_mod.helio_bflux(p, w, out, radial, (1.0, 0.2, 0.1))
    np.testing.assert_allclose(f, mhd_flux(w, G), rtol=1e-14, atol=1e-14)
    assert bn == w[5] and ws is None
    w[1] = -1.2 * fast_speed(w)
    f, bn, ws = oracle_mod.helio_bflux(p, w, out, radial, (1.0, 0.2, 0.1))
    np.testing.assert_allclose(f, mhd_flux(out, G), rtol=1e-14, atol=1e-14)
    assert bn == out[5] and ws is None


@pytest.mark.parametrize("radial", [False, True], ids=["BFLUX", "BFLUXR"])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_bflux_sub_fast_outflow_is_fast_sonic(oracle_mod, radial, k):
    """0 <= u = c_f (1 - delta): the boundary state is W + h r + O(delta^2)
    with r the u - c_f eigenvector (normalised to d rho = rho) and h the root
    of the linearised condition u(W + h r) = c_f(W + h r); the residual
    u - c_f of the state is O(delta^2); the flux is its physical flux and
    the face's normal field its Bn.  Checked at delta = 1e-2, 1e-3, 1e-4:
    the state error falls tenfold and the residual a hundredfold per step
    (a wrong term in the linearisation leaves an O(delta) residual)."""
    p = plasma_params()
    errs, resid = [], []
    for delta in (1e-2, 1e-3, 1e-4):
        w = STATES[k].copy()
        w[1] = fast_speed(w) * (1 - delta)
        out = w.copy()
        out[1] = 5.0          # exterior leaving faster than its c_f (BFLUX's condition)
        lam, R = np.linalg.eig(prim_jacobian(w, G))
        lam, R = lam.real, R.real
        j = np.argmin(lam)
        assert lam[j] == pytest.approx(w[1] - fast_speed(w), rel=1e-10)
        r = R[:, j] / R[0, j] * w[0]
        e = 1e-7
        dcf = (fast_speed(w + e * r) - fast_speed(w - e * r)) / (2 * e)
        h = (fast_speed(w) - w[1]) / (r[1] - dcf)
        f, bn, ws = oracle_mod.helio_bflux(p, w, out, radial, (1.0, 0.2, 0.1))
        assert ws is not None
        np.testing.assert_allclose(f, mhd_flux(ws, G), rtol=1e-13, atol=1e-14)
        assert bn == ws[5] == w[5]
        errs.append(np.max(np.abs((ws - w) - h * r)) / np.max(np.abs(h * r)))
        resid.append(abs(ws[1] - fast_speed(ws)) / fast_speed(w))
    for a, b in zip(errs, errs[1:]):
        assert b < 0.15 * a
    for a, b in zip(resid, resid[1:]):
        assert b < 0.015 * a
    assert errs[-1] < 1e-3 and resid[-1] < 1e-8


@pytest.mark.parametrize("radial", [False, True], ids=["BFLUX", "BFLUXR"])
def test_bflux_sub_fast_inflow_is_roe(oracle_mod, radial):
    """-c_f < u < 0: the 8-wave Roe flux of (interior, exterior) and the mean
    normal field; BFLUX also takes it for a sub-fast outflow whose exterior
    is not super-fast."""
    p = plasma_params()
    rng = np.random.default_rng(61)
    for t in range(20):
        w = STATES[t % 3].copy()
        out = STATES[(t + 1) % 3].copy() * rng.uniform(0.8, 1.2, 8)
        w[1] = -rng.uniform(0.05, 0.95) * fast_speed(w)
        out[1] = rng.uniform(-0.5, 0.5)
        f, bn, ws = oracle_mod.helio_bflux(p, w, out, radial, (1.0, 0.2, 0.1))
        assert ws is None
        np.testing.assert_allclose(f, oracle_mod.flux_roe8(p, w, out, 0), rtol=1e-14, atol=1e-14)
        assert bn == pytest.approx(0.5 * (w[5] + out[5]), rel=1e-15)
        if not radial:
            w[1] = -w[1]                       # sub-fast outflow
            out[1] = 0.5 * fast_speed(out)     # exterior not super-fast
            f, bn, ws = oracle_mod.helio_bflux(p, w, out, radial)
            assert ws is None
            np.testing.assert_allclose(f, oracle_mod.flux_roe8(p, w, out, 0),
                                       rtol=1e-14, atol=1e-14)


# ---------------------------------------------------------------- BFLUXGAS --
def gas_params():
    p = mfmhd.orszag_params(G)
    p.smallr = 1e-12
    p.smallp = 1e-12
    return p


def sonic_state(w, sign):
    """The state with u = sign * a on the characteristic of w carrying
    u + sign 2a/(gamma-1), at w's entropy; tangential velocity kept."""
    rho, u, v, ww, pr = w
    J = u + sign * 2 * np.sqrt(G * pr / rho) / (G - 1)
    s = pr / rho ** G

    def eqs(x):
        r, un, pg = np.exp(x[0]), x[1], np.exp(x[2])
        a = np.sqrt(G * pg / r)
        return [un + sign * 2 * a / (G - 1) - J, np.log(pg / r ** G) - np.log(s), un - sign * a]
    x = fsolve(eqs, [np.log(rho), sign * np.sqrt(G * pr / rho), np.log(pr)], xtol=1e-12)
    assert np.max(np.abs(eqs(x))) < 1e-11
    return np.array([np.exp(x[0]), x[1], v, ww, np.exp(x[2])])


GAS = [np.array([1.0, 0.0, 0.3, -0.2, 0.6]), np.array([0.3, 0.0, -0.1, 0.4, 0.05]),
       np.array([2.5, 0.0, 0.0, 0.0, 3.0])]


@pytest.mark.parametrize("k", range(len(GAS)))
def test_bfluxgas_branches(oracle_mod, k):
    p = gas_params()
    w = GAS[k].copy()
    a = np.sqrt(G * w[4] / w[0])
    ext = GAS[(k + 1) % 3].copy()
    a1 = np.sqrt(G * ext[4] / ext[0])
    riemann = np.array([0.11, 0.22, 0.33, 0.44, 0.55])
    # supersonic outflow: the interior's flux; supersonic inflow: the exterior's
    w[1], ext[1] = 1.3 * a, 0.0
    np.testing.assert_allclose(oracle_mod.bfluxgas(p, w, ext, riemann), euler_flux(G, w, 0),
                               rtol=1e-14, atol=1e-15)
    w[1], ext[1] = -1.3 * a, -0.2
    np.testing.assert_allclose(oracle_mod.bfluxgas(p, w, ext, riemann), euler_flux(G, ext, 0),
                               rtol=1e-14, atol=1e-15)
    # sub-sonic outflow: the sonic state on the outgoing characteristic
    # (J+ and the entropy of the interior) when the exterior leaves
    # supersonically, else the Riemann flux is kept
    for frac in (0.0, 0.4, 0.9):
        w[1], ext[1] = frac * a, 1.5 * a1
        s = sonic_state(w, +1)
        np.testing.assert_allclose(oracle_mod.bfluxgas(p, w, ext, riemann), euler_flux(G, s, 0),
                                   rtol=1e-12, atol=1e-14)
        ext[1] = 0.5 * a1
        assert np.array_equal(oracle_mod.bfluxgas(p, w, ext, riemann), riemann)
    # sub-sonic inflow: the sonic state on the entering characteristic (J-
    # and the entropy of the exterior) when the exterior arrives
    # supersonically, else the Riemann flux is kept
    for frac in (0.1, 0.5, 0.95):
        w[1], ext[1] = -frac * a, -1.5 * a1
        s = sonic_state(ext, -1)
        np.testing.assert_allclose(oracle_mod.bfluxgas(p, w, ext, riemann), euler_flux(G, s, 0),
                                   rtol=1e-12, atol=1e-14)
        ext[1] = -0.5 * a1
        assert np.array_equal(oracle_mod.bfluxgas(p, w, ext, riemann), riemann)


# ------------------------------------------------- HELIOBC, neutral fluids --
def helio4():
    p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, sunN=7.0, sunV=4.5e7, netN=0.1,
                           R0=0.1, dx=1.0 / 64, num_cells=(64, 64, 64), sunXYZ=(0.5, 0.5, 0.5))
    p.nfluids = 4
    return p


# the floors HELIOBC applies to neutral populations 2 and 3 (rho, p)
FLOORS = {2: (1.0e-6, 1.0e-8), 3: (1.0e-7, 1.0e-9)}


def neutral_slice(f):
    b = 8 + 5 * (f - 1)
    return slice(b, b + 5)


def lab_euler(w5, d):
    return euler_flux(G, w5, d)


def helio_face(oracle_mod, p, side, d, interior, exterior):
    """HELIOBC's flux at the face of direction d on side +-1 with the
    interior cell's face state and the exterior's."""
    wl, wr = (interior, exterior) if side == 1 else (exterior, interior)
    ijk = [32, 32, 32]
    ijk[d] = 63 if side == 1 else 0
    return oracle_mod.helio_bc_face(p, side, d, 1.0 / 64, ijk, wl, wr)[0]


def two_states(rng):
    from test_oracle_pins import rand_prim
    a, b = rand_prim(rng, nf=4), rand_prim(rng, nf=4)
    return a, b


@pytest.mark.parametrize("f", [2, 3])
def test_helio_bc_populations_2_3_outflow_only(oracle_mod, f):
    """Populations 2 and 3 leave through every boundary face with the
    physical flux of the interior's floored state, the normal velocity
    clamped to leave (none enters); the exterior state plays no part."""
    p = helio4()
    rng = np.random.default_rng(70 + f)
    sl = neutral_slice(f)
    dmin, pmin = FLOORS[f]
    for t in range(36):
        d, side = t % 3, (1 if (t // 3) % 2 else -1)
        w_in, w_out = two_states(rng)
        if t % 4 == 0:   # below the floors
            w_in[sl.start], w_in[sl.start + 4] = 0.1 * dmin, 0.1 * pmin
        got = helio_face(oracle_mod, p, side, d, w_in, w_out)[sl]
        s = w_in[sl].copy()
        s[0], s[4] = max(dmin, s[0]), max(pmin, s[4])
        s[1 + d] = max(0.0, s[1 + d]) if side == 1 else min(0.0, s[1 + d])
        np.testing.assert_allclose(got, lab_euler(s, d), rtol=1e-14, atol=1e-300)
        w_out2 = w_out.copy()
        w_out2[sl] *= 1.7
        assert np.array_equal(helio_face(oracle_mod, p, side, d, w_in, w_out2)[sl], got)


def test_helio_bc_population_1_low_x_outflow(oracle_mod):
    """At the low-x face (the downwind boundary of the LISM flow along -x)
    population 1 leaves with the interior state's physical flux, the
    normal velocity clamped to leave (at least 1e-6)."""
    p = helio4()
    rng = np.random.default_rng(81)
    sl = neutral_slice(1)
    for _ in range(20):
        w_in, w_out = two_states(rng)
        got = helio_face(oracle_mod, p, -1, 0, w_in, w_out)[sl]
        s = w_in[sl].copy()
        s[1] = min(-1e-6, s[1])
        np.testing.assert_allclose(got, lab_euler(s, 0), rtol=1e-14, atol=1e-300)


def test_helio_bc_population_1_van_leer(oracle_mod):
    """At the other boundary faces population 1 takes van Leer's flux of
    (below, above) in the lab frame; when the interior state leaves, the
    exterior takes its velocity and density, so that an
    interior leaving at the exterior's pressure has its own physical flux
    (consistency) -- on the low side as the point image of the high side."""
    p = helio4()
    rng = np.random.default_rng(91)
    sl = neutral_slice(1)
    n = 0
    for t in range(60):
        d, side = t % 3, (1 if (t // 3) % 2 else -1)
        if side == -1 and d == 0:
            continue
        w_in, w_out = two_states(rng)
        got = helio_face(oracle_mod, p, side, d, w_in, w_out)[sl]
        a, b = w_in[sl].copy(), w_out[sl].copy()
        if side * a[1 + d] > 0:          # the interior leaves
            b[0:4] = a[0:4]
        lo, hi = (a, b) if side == 1 else (b, a)
        want = oracle_mod.flux_vanleer(p, lo, hi, d)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14 * np.max(np.abs(want)))
        # consistency: leaving at the exterior's pressure (and density)
        w_out3 = w_out.copy()
        w_out3[sl.start + 4] = w_in[sl.start + 4]
        w_out3[sl.start] = w_in[sl.start]
        w_in3 = w_in.copy()
        w_in3[sl.start + 1 + d] = side * abs(w_in[sl.start + 1 + d])
        got3 = helio_face(oracle_mod, p, side, d, w_in3, w_out3)[sl]
        np.testing.assert_allclose(got3, lab_euler(w_in3[sl], d), rtol=1e-12, atol=1e-14)
        n += 1
    assert n >= 40


# ------------------------------------------------------------ ghost copies --
def _box_fab(rng, nc, lo, hi):
    shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    return rng.uniform(0.5, 2.0, shape)


@pytest.mark.parametrize("side", [1, -1])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_heliogs_ghosts(oracle_mod, d, side):
    """HELIOGS (swlism/HeliosphericF.ChF:978-1149) on every face of the
    heliospheric box: the plasma ghosts hold the LISM state, neutral
    population 1 the LISM velocity at its own density and pressure, and
    populations 2 and 3 copy the interior cell next to the boundary in their
    row (zero-order extrapolation); the interior is untouched."""
    p = helio4()
    nc = mfmhd.num_components(p)
    rng = np.random.default_rng(100 + 3 * d + side)
    lo, hi = [28, 28, 28], [35, 35, 35]
    if side == 1:
        lo[d], hi[d] = 58, 63 + 4
    else:
        lo[d], hi[d] = -4, 5
    W0 = _box_fab(rng, nc, lo, hi)
    W = oracle_mod.fill_ghost(p, W0.copy(), lo, d, 1.0 / 64)
    ax = 3 - d                       # numpy axis of direction d
    idx = np.arange(lo[d], hi[d] + 1)
    ghost = idx > 63 if side == 1 else idx < 0
    edge = (63 if side == 1 else 0) - lo[d]

    def take(A, sel):
        return np.take(A, np.nonzero(sel)[0] if sel.dtype == bool else sel, axis=ax)
    assert np.array_equal(take(W, ~ghost), take(W0, ~ghost))
    G = take(W, ghost)
    lism = [1.0, p.lism_ux, p.lism_uy, p.lism_uz, p.lism_p, p.lism_bx, p.lism_by, p.lism_bz]
    for c, v in enumerate(lism):
        assert np.all(G[c] == v), c
    net = [p.net_rho, p.lism_ux, p.lism_uy, p.lism_uz, p.net_p]
    for c, v in enumerate(net):
        assert np.all(G[8 + c] == v), c
    E = take(W0, np.array([edge]))
    for c in range(13, 23):
        assert np.array_equal(G[c], np.broadcast_to(E[c], G[c].shape)), c


@pytest.mark.parametrize("side,d", [(1, 0), (-1, 0), (1, 1)])
def test_swlism_ghost_copies_of_populations_2_3(oracle_mod, side, d):
    """LISMINIT's ghosts (swlism/SWLISMF.ChF:512-652) on the x boundaries
    and the top: populations 2 and 3 copy the interior cell next to the
    boundary in their row; the interior is untouched."""
    nx, ny, dx = 48, 32, 1.0
    p = mfmhd.swlism_params(sunXY=(0.4 * nx * dx, 0.0), R0=2.6 * dx, initR=6.0 * dx, dx=dx,
                            num_cells=(nx, ny), nfluids=4)
    nc = mfmhd.num_components(p)
    rng = np.random.default_rng(120 + 3 * d + side)
    lo, hi = [20, 10, 0], [27, 17, 0]
    top = (nx - 1, ny - 1)[d]
    if side == 1:
        lo[d], hi[d] = top - 5, top + 4
    else:
        lo[d], hi[d] = -4, 5
    W0 = _box_fab(rng, nc, lo, hi)
    W = oracle_mod.fill_ghost(p, W0.copy(), lo, d, dx)
    ax = 3 - d
    idx = np.arange(lo[d], hi[d] + 1)
    ghost = idx > top if side == 1 else idx < 0
    edge = (top if side == 1 else 0) - lo[d]
    Wi = np.take(W, np.nonzero(~ghost)[0], axis=ax)
    assert np.array_equal(Wi, np.take(W0, np.nonzero(~ghost)[0], axis=ax))
    G = np.take(W, np.nonzero(ghost)[0], axis=ax)
    E = np.take(W0, [edge], axis=ax)
    for c in range(13, 23):
        assert np.array_equal(G[c], np.broadcast_to(E[c], G[c].shape)), c


# ---------------------------------------------------------- SWLISMREINIT --
def test_swlism_reinit_states(oracle_mod):
    """SWLISMREINIT (swlism/SWLISMF.ChF:305-403) inside R0: the plasma is the
    radial supersonic wind (rho_sun / r^2, speed sun_v along r, pressure
    p_sun / r^(2 gamma), no field), populations 2 and 3 are at rest at their
    floor density and pressure (energy p / (gamma - 1)) and population 1 is
    untouched; outside R0 nothing changes."""
    nx, ny, dx = 48, 32, 1.0
    p = mfmhd.swlism_params(sunXY=(0.4 * nx * dx, 0.0), R0=5.2 * dx, initR=8.0 * dx, dx=dx,
                            num_cells=(nx, ny), nfluids=4)
    nc = mfmhd.num_components(p)
    rng = np.random.default_rng(130)
    lo, hi = (10, 0, 0), (29, 9, 0)
    U0 = rng.uniform(0.5, 2.0, (nc, 1, hi[1] + 1, hi[0] - lo[0] + 1))
    U = oracle_mod.swlism_reinit(p, U0.copy(), lo, lo, hi, dx)
    g = p.gamma
    n_in = 0
    for j in range(hi[1] + 1):
        for i in range(lo[0], hi[0] + 1):
            x, y = (i + 0.5) * dx - p.sun_xc, (j + 0.5) * dx - p.sun_yc
            r = np.hypot(x, y)
            u, u0 = U[:, 0, j, i - lo[0]], U0[:, 0, j, i - lo[0]]
            if r > p.sun_r0:
                assert np.array_equal(u, u0)
                continue
            n_in += 1
            rho, pr = p.sun_rho / r ** 2, p.sun_p / r ** (2 * g)
            want = [rho, rho * p.sun_v * x / r, rho * p.sun_v * y / r, 0.0,
                    pr / (g - 1) + 0.5 * rho * p.sun_v ** 2, 0.0, 0.0, 0.0]
            np.testing.assert_allclose(u[:8], want, rtol=1e-14, atol=0)
            assert np.array_equal(u[8:13], u0[8:13])
            for f, (dmin, pmin) in FLOORS.items():
                b = 8 + 5 * (f - 1)
                assert list(u[b:b + 5]) == [dmin, 0.0, 0.0, 0.0, pmin / (g - 1)], f
    assert n_in > 20


@pytest.mark.parametrize("side", [1, -1])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_helio_bc_plasma(oracle_mod, d, side):
    """HELIOBC's plasma flux (swlism/HeliosphericF.ChF:221-529): the high-x
    face (the LISM inflow) carries the LISM state's physical flux and field;
    the other faces take BFLUX of (interior, exterior) in the face frame, the
    low faces as the point image of the high ones (velocity and field
    reversed, mass and energy fluxes and the normal field negated back) and
    the low-x face with BFLUXR's radial branch."""
    p = helio4()
    rng = np.random.default_rng(140 + 3 * d + side)
    from test_oracle_pins import rand_prim
    n, t1, t2 = 1 + d, 1 + (d + 1) % 3, 1 + (d + 2) % 3
    for _ in range(8):
        w_in, w_out = rand_prim(rng, nf=4), rand_prim(rng, nf=4)
        wl, wr = (w_in, w_out) if side == 1 else (w_out, w_in)
        ijk = [30, 33, 35]
        ijk[d] = 63 if side == 1 else 0
        f, bn = oracle_mod.helio_bc_face(p, side, d, 1.0 / 64, ijk, wl, wr)
        if side == 1 and d == 0:
            lism = np.array([1.0, p.lism_ux, p.lism_uy, p.lism_uz, p.lism_p,
                             p.lism_bx, p.lism_by, p.lism_bz])
            np.testing.assert_allclose(f[:8], mhd_flux(lism, G), rtol=1e-14, atol=1e-300)
            assert bn == p.lism_bx
            continue
        sgn = 1.0 if side == 1 else -1.0

        def frame(w):
            return np.array([w[0], sgn * w[n], sgn * w[t1], sgn * w[t2], w[4],
                             sgn * w[4 + n], sgn * w[4 + t1], sgn * w[4 + t2]])
        radial = side == -1 and d == 0
        dxyz = (0.0, 0.0, 0.0)
        if radial:
            dxyz = tuple(-((ijk[q] + 0.5) / 64 - (p.sun_xc, p.sun_yc, p.sun_zc)[q]) for q in range(3))
        fl, bnv, _ = oracle_mod.helio_bflux(p, frame(w_in), frame(w_out), radial, dxyz)
        want = np.zeros(8)
        want[0], want[4] = sgn * fl[0], sgn * fl[4]
        want[n], want[t1], want[t2] = fl[1], fl[2], fl[3]
        want[4 + n], want[4 + t1], want[4 + t2] = fl[5], fl[6], fl[7]
        np.testing.assert_allclose(f[:8], want, rtol=1e-14, atol=1e-300)
        assert bn == sgn * bnv
