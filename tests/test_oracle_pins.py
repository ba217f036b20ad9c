"""Pins of the CPU oracle (no GPU).  The reference ships no tests or golden
vectors (its ChomboFortran kernels cannot be built here: Chombo and its
Fortran preprocessor are absent), so each oracle function is checked against
what the mathematics fixes: closed forms from a different reference routine
or a textbook form, conservation invariants, eigen-system identities,
symmetries and brute force.  A dropped term, a wrong sign or index, or a
transposed operand fails at least one of these."""
import numpy as np
import pytest

import mfmhd

D4 = 1.0 / (4 * np.pi)
D8 = 1.0 / (8 * np.pi)


def params(nf=1, gamma=5.0 / 3.0, dedner=0):
    p = mfmhd.default_params()
    p.gamma = gamma
    p.nfluids = nf
    p.dedner = dedner
    p.smallr = 1e-12
    p.smallp = 1e-12
    return p


def rand_prim(rng, nf=1, bscale=1.0):
    w = [rng.uniform(0.5, 2.0), *rng.uniform(-1, 1, 3), rng.uniform(0.3, 2.0),
         *(bscale * rng.uniform(-2, 2, 3))]
    for _ in range(nf - 1):
        w += [rng.uniform(0.2, 1.5), *rng.uniform(-1, 1, 3), rng.uniform(0.1, 1.0)]
    return np.array(w)


def prim_to_cons(p, w):
    """PRIMTOCONS / PRIMTOCONS_MF (srcIdeal/PatchIdealMHDF.ChF:191-233,
    srcIdeal/PatchMHDMFF.ChF:82-150)."""
    g = p.gamma
    u = np.zeros_like(w)
    r = w[0]
    u[0] = r
    u[1:4] = r * w[1:4]
    u[4] = w[4] / (g - 1) + 0.5 * (r * np.dot(w[1:4], w[1:4]) + D4 * np.dot(w[5:8], w[5:8]))
    u[5:8] = w[5:8]
    c = 8
    for _ in range(p.nfluids - 1):
        rn = w[c]
        u[c] = rn
        u[c + 1:c + 4] = rn * w[c + 1:c + 4]
        u[c + 4] = w[c + 4] / (g - 1) + 0.5 * rn * np.dot(w[c + 1:c + 4], w[c + 1:c + 4])
        c += 5
    return u


def mhd_flux(p, w, d):
    """Ideal-MHD flux in Gaussian units (textbook form)."""
    g = p.gamma
    r, v, pg, B = w[0], w[1:4], w[4], w[5:8]
    vn, Bn = v[d], B[d]
    ptot = pg + D8 * np.dot(B, B)
    E = pg / (g - 1) + 0.5 * r * np.dot(v, v) + D8 * np.dot(B, B)
    f = np.zeros(8)
    f[0] = r * vn
    f[1:4] = r * vn * v - D4 * Bn * B
    f[1 + d] += ptot
    f[4] = vn * (E + ptot) - D4 * Bn * np.dot(v, B)
    f[5:8] = vn * B - Bn * v
    f[5 + d] = 0.0
    return f


def cons_vec(p, w):
    return prim_to_cons(p, w)[:8]


# ------------------------------------------------------------ cons/prim ---
@pytest.mark.parametrize("nf", [1, 2, 4])
def test_cons_prim_round_trip(oracle_mod, nf):
    rng = np.random.default_rng(1)
    p = params(nf)
    for _ in range(50):
        w = rand_prim(rng, nf)
        w2 = oracle_mod.cons_to_prim(p, prim_to_cons(p, w))
        np.testing.assert_allclose(w2, w, rtol=1e-13, atol=1e-14)


def test_cons_prim_floors(oracle_mod):
    p = params(2)
    p.smallr, p.smallp = 1e-6, 1e-7
    u = prim_to_cons(p, np.array([1.0, 0.1, 0.2, 0.3, 1e-3, 0.5, 0.5, 0.5,
                                  1.0, 0.1, 0.1, 0.1, 1.0]))
    u[4] -= 1.0  # negative plasma pressure
    u[8] = 1e-9  # neutral density below smallr (multi-fluid floors it)
    w = oracle_mod.cons_to_prim(p, u)
    assert w[4] == p.smallp
    assert w[8] == p.smallr


@pytest.mark.parametrize("nf", [1, 4])
def test_positivity_floors(oracle_mod, nf):
    """POSTPROCESSING / POSTPROCESSING_MF (srcIdeal/PatchIdealMHDF.ChF:1019-1070,
    srcIdeal/PatchMHDMFF.ChF:301-511): a fluid whose density or pressure is
    below its floor gets the floor and a rebuilt energy (the multi-fluid
    version keeps its velocity); the floors are smallr/smallp for the plasma
    and neutral 1, 1e-6/1e-8 for neutral 2, 1e-7/1e-9 for neutral 3; a fluid
    above its floors is left bit-for-bit unchanged."""
    p = params(nf)
    p.smallr, p.smallp = 1e-5, 1e-6
    g = p.gamma
    ok = np.array([1.0, 0.1, 0.2, 0.3, 0.5, 0.4, 0.3, 0.2])
    fl = [ok]
    for f in range(1, nf):
        fl.append(np.array([0.7, 0.1, -0.2, 0.05, 0.3]))
    u = prim_to_cons(p, np.concatenate(fl))
    assert np.array_equal(oracle_mod.floors(p, u), u)
    floors = {0: (p.smallr, p.smallp), 1: (p.smallr, p.smallp), 2: (1e-6, 1e-8),
              3: (1e-7, 1e-9)}
    for f in range(nf):
        for which in ("rho", "p"):
            v = u.copy()
            b = 0 if f == 0 else 8 + 5 * (f - 1)
            vel = v[b + 1:b + 4] / v[b]
            if which == "rho":
                v[b] = 0.1 * floors[f][0]
                v[b + 1:b + 4] = v[b] * vel
            else:
                ke = 0.5 * v[b] * np.dot(vel, vel)
                mag = np.dot(v[5:8], v[5:8]) / (8 * np.pi) if f == 0 else 0.0
                v[b + 4] = ke + mag - 1.0  # negative pressure
            out = oracle_mod.floors(p, v)
            rho_f, p_f = floors[f]
            assert out[b] == max(v[b], rho_f)
            r = out[b]
            pres = (g - 1) * (out[b + 4] - 0.5 * np.dot(out[b + 1:b + 4], out[b + 1:b + 4]) / r
                              - (np.dot(out[5:8], out[5:8]) / (8 * np.pi) if f == 0 else 0.0))
            assert pres == pytest.approx(max(p_f, pres), rel=1e-9, abs=1e-12)
            if which == "p":
                assert pres == pytest.approx(p_f, rel=1e-6)
            if nf > 1 or f > 0:
                assert np.allclose(out[b + 1:b + 4] / r, vel, rtol=1e-12)
            # the other fluids are untouched
            mask = np.ones(len(u), bool)
            mask[b:b + 5] = False
            if f == 0:
                mask[:8] = False
            assert np.array_equal(out[mask], v[mask])


def test_helio_initial_state_invariants(oracle_mod):
    """HELIOINIT (swlism/HeliosphericF.ChF:1150-1361): inside initR the
    solar wind is radial with |v| = sunV, rho r^2 and p r^(2 gamma) constant;
    outside, rho = 1 and velocity and field are the potential flow past a
    sphere of radius initR: radial part (1 - (R/r)^3) (V.r), tangential part
    (1 + (R/r)^3 / 2) V_t, for V the LISM vector.  Checked as properties of
    the generated field."""
    from helpers import grid_shape
    n = 24
    dx = 1.0 / n
    p = mfmhd.helio_params(lismB=3e-6, lismBdir=(0.6, 0.8, 0.0), sunXYZ=(0.5, 0.5, 0.5),
                           R0=0.1, initR=0.3, netN=0.1, sunV=4.5e7, dx=dx,
                           num_cells=(n, n, n))
    p.nfluids = 2
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(13, vlo, vhi, 4)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    rho_r2, p_r2g, n_in, n_out = [], [], 0, 0
    for k in range(n):
        for j in range(n):
            for i in range(n):
                x, y, z = ((i + 0.5) * dx - 0.5, (j + 0.5) * dx - 0.5, (k + 0.5) * dx - 0.5)
                r = np.sqrt(x * x + y * y + z * z)
                u = U[:, k - lo[2], j - lo[1], i - lo[0]]
                w = oracle_mod.cons_to_prim(p, u)
                if 0.12 < r < 0.28:
                    n_in += 1
                    v = w[1:4]
                    assert np.linalg.norm(v) == pytest.approx(p.sun_v, rel=1e-12)
                    assert np.dot(v, (x, y, z)) / (np.linalg.norm(v) * r) == pytest.approx(1.0, abs=1e-12)
                    rho_r2.append(w[0] * r * r)
                    p_r2g.append(w[4] * r ** (2 * p.gamma))
                elif r > 0.32:
                    n_out += 1
                    assert w[0] == pytest.approx(1.0, rel=1e-12)
                    e = np.array((x, y, z)) / r
                    q = (0.3 / r) ** 3
                    for got, far in ((w[1:4], (p.lism_ux, p.lism_uy, p.lism_uz)),
                                     (w[5:8], (p.lism_bx, p.lism_by, p.lism_bz))):
                        far = np.array(far)
                        assert np.dot(got, e) == pytest.approx((1 - q) * np.dot(far, e), abs=1e-12)
                        tg, tf = got - np.dot(got, e) * e, far - np.dot(far, e) * e
                        assert np.allclose(tg, (1 + 0.5 * q) * tf, rtol=1e-10, atol=1e-12)
    assert n_in > 50 and n_out > 1000
    assert np.ptp(rho_r2) <= 1e-10 * np.mean(rho_r2)
    assert np.ptp(p_r2g) <= 1e-9 * np.mean(p_r2g)


# ------------------------------------------------------------- MHD fluxes --
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("avg", [0, 1])
def test_roe8_consistency(oracle_mod, d, avg):
    """F(W, W) equals the physical flux: FLUXESHANCOCK and a textbook form."""
    rng = np.random.default_rng(2 + d)
    p = params()
    for _ in range(20):
        w = rand_prim(rng)
        f = oracle_mod.flux_roe8(p, w, w, d, avg)
        np.testing.assert_allclose(f, mhd_flux(p, w, d), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(f, oracle_mod.physical_flux(p, w, d),
                                   rtol=1e-12, atol=1e-12)


def lf_expected(p, wl, wr, d):
    """Rusanov form 1/2(FL+FR) + 1/2 lam (UL-UR), lam = |u| + c_fast of the
    arithmetically averaged state with the sound speed squared clamped to
    [min, max](aL^2, aR^2), as RIEMANNF's averaging defines it."""
    g = p.gamma
    ra = 0.5 * (wl[0] + wr[0])
    un = 0.5 * (wl[1 + d] + wr[1 + d])
    B = 0.5 * (wl[5:8] + wr[5:8])
    PL = wl[4] + D8 * np.dot(wl[5:8], wl[5:8])
    PR = wr[4] + D8 * np.dot(wr[5:8], wr[5:8])
    pg = 0.5 * (PL + PR) - D8 * np.dot(B, B)
    cc = g * pg / ra
    a2 = sorted([g * wl[4] / wl[0], g * wr[4] / wr[0]])
    cc = min(max(cc, a2[0]), a2[1])
    bb = D4 * B[d] ** 2 / ra
    b2 = D4 * np.dot(B, B) / ra
    cf = np.sqrt(0.5 * (cc + b2 + np.sqrt((cc + b2) ** 2 - 4 * cc * bb)))
    lam = abs(un) + cf
    return 0.5 * (mhd_flux(p, wl, d) + mhd_flux(p, wr, d) +
                  lam * (cons_vec(p, wl) - cons_vec(p, wr)))


@pytest.mark.parametrize("d", [0, 1, 2])
def test_roe8_eigenbasis(oracle_mod, d):
    """With all |lambda| equal (iLaxFriedrix = 1) R|L|L reduces to
    lambda * I iff the left and right eigenvector matrices are inverse: the
    flux must equal the Rusanov form exactly.  Includes the degenerate
    branch B_t = 0."""
    rng = np.random.default_rng(10 + d)
    p = params()
    for k in range(30):
        wl = rand_prim(rng)
        wr = wl + rng.uniform(-0.2, 0.2, 8)
        wr[0] = abs(wr[0]) + 0.3
        wr[4] = abs(wr[4]) + 0.3
        if k % 5 == 0:  # B transverse to the face = 0 on both sides
            for q in range(3):
                if q != d:
                    wl[5 + q] = wr[5 + q] = 0.0
        f = oracle_mod.flux_roe8(p, wl, wr, d, 0, 1)
        e = lf_expected(p, wl, wr, d)
        np.testing.assert_allclose(f, e, rtol=1e-11, atol=1e-11 * np.abs(e).max())


def test_roe8_rotation_invariance(oracle_mod):
    """The cyclic (normal, t1, t2) mapping makes a rotated state give the
    rotated flux, bit for bit."""
    rng = np.random.default_rng(4)
    p = params()
    for _ in range(10):
        wl, wr = rand_prim(rng), rand_prim(rng)
        f0 = oracle_mod.flux_roe8(p, wl, wr, 0)
        for d in (1, 2):
            perm = [(q - d) % 3 for q in range(3)]  # new comp q takes old perm[q]

            def rot(w):
                r = w.copy()
                r[1:4] = w[1:4][perm]
                r[5:8] = w[5:8][perm]
                return r
            fd = oracle_mod.flux_roe8(p, rot(wl), rot(wr), d)
            assert np.array_equal(fd, rot(f0))


@pytest.mark.parametrize("d", [0, 2])
def test_rusanov_closed_form(oracle_mod, d):
    rng = np.random.default_rng(5)
    p = params()
    g = p.gamma
    for _ in range(20):
        wl, wr = rand_prim(rng), rand_prim(rng)
        f = oracle_mod.flux_rusanov(p, wl, wr, d, 0)
        ra = 0.5 * (wl[0] + wr[0])
        B = 0.5 * (wl[5:8] + wr[5:8])
        a2 = g * 0.5 * (wl[4] + wr[4]) / ra
        b2 = D4 * np.dot(B, B) / ra
        bb = D4 * B[d] ** 2 / ra
        cf = np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bb)))
        lam = abs(0.5 * (wl[1 + d] + wr[1 + d])) + cf
        e = 0.5 * (mhd_flux(p, wl, d) + mhd_flux(p, wr, d) +
                   lam * (cons_vec(p, wl) - cons_vec(p, wr)))
        np.testing.assert_allclose(f, e, rtol=1e-12, atol=1e-12)


def test_roe7_selection(oracle_mod):
    """flux 11 / 13 select the 7-wave Roe solver (src/RiemannSolver.cpp:
    76-89): zero normal-field flux at every face, and where the normal
    field is continuous the 8-wave solver's divergence wave has zero
    strength, so the 7-wave flux equals the 8-wave flux on the same means
    (11 ~ arithmetic, 13 ~ Roe); flux 0 is the factory's default case, the
    8-wave flux on arithmetic means."""
    rng = np.random.default_rng(12)
    p = params()
    for t in range(60):
        d = t % 3
        wl, wr = rand_prim(rng), rand_prim(rng)
        p.flux = 0
        f0 = oracle_mod.flux_face(p, wl, wr, d)
        p.flux = 1
        assert np.array_equal(f0, oracle_mod.flux_face(p, wl, wr, d))
        for flux, avg in ((11, 0), (13, 1)):
            p.flux = flux
            f7 = oracle_mod.flux_face(p, wl, wr, d)
            assert f7[5 + d] == 0.0
            f8 = oracle_mod.flux_roe8(p, wl, wr, d, averaging=avg)
            assert abs(f8[5 + d]) > 1e-6    # the divergence wave of a jump in B_n
            wc = wr.copy()
            wc[5 + d] = wl[5 + d]
            f7 = oracle_mod.flux_face(p, wl, wc, d)
            f8 = oracle_mod.flux_roe8(p, wl, wc, d, averaging=avg)
            np.testing.assert_allclose(f7, f8, rtol=0, atol=1e-13 * np.max(np.abs(f8)))


@pytest.mark.parametrize("flux7,flux8", [(11, 1), (13, 3)])
def test_roe7_patch_with_uniform_normal_field(oracle_mod, flux7, flux8):
    """A Brio-Wu tube along x (B_x uniform, one periodic cell in y): every
    face has a continuous normal field, so two RK2 steps with the 7-wave
    and the 8-wave solver give the same state."""
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    n, ng, dx = 64, 4, 1.0 / 64
    out = []
    for flux in (flux7, flux8):
        p = mfmhd.riemann_params(2.0, L, R, 0.5)
        p.time_approx = mfmhd.TIME_RK2
        p.flux = flux
        p.domain_lo[:] = [0, 0, 0]
        p.domain_hi[:] = [n - 1, 0, 0]
        p.periodic[:] = [0, 1, 1]
        p.spacedim = 2
        lo = (-ng, -ng, 0)
        U = oracle_mod.init_problem(p, (8, 1, 1 + 2 * ng, n + 2 * ng), lo, lo,
                                    (n - 1 + ng, ng, 0), dx)
        dt = 0.4 * oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dx)[0]
        for _ in range(2):
            U[:, :, :, :] = U[:, :, ng:ng + 1, :]
            r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (n - 1, 0, 0), dt, dx, 0.0,
                                        want_F=True)
            assert r["rc"] == 0
            U = r["U"]
        out.append((U, r["F"][0]))
    (U7, F7), (U8, F8) = out
    assert np.max(np.abs(U7[0] - 1.0)) > 1e-3     # the tube has evolved
    np.testing.assert_allclose(U7, U8, rtol=0, atol=1e-12 * np.max(np.abs(U8)))
    assert np.all(F7[5] == 0.0)


def test_dedner_glm_block(oracle_mod):
    """DEDNER2X2F: upwinding of (B_n, psi) with speed Ch; the other
    components come from the 7-wave Roe flux, consistent at W_L = W_R."""
    rng = np.random.default_rng(6)
    p = params(dedner=1)
    ch = 1.7
    for d in range(3):
        wl = np.append(rand_prim(rng), 0.3)
        wr = np.append(rand_prim(rng), -0.1)
        f = oracle_mod.flux_dedner(p, ch, wl, wr, d)
        assert f[5 + d] == pytest.approx(0.5 * (wl[8] + wr[8]) + 0.5 * ch * (wl[5 + d] - wr[5 + d]), rel=1e-13)
        assert f[8] == pytest.approx(0.5 * ch * ch * (wl[5 + d] + wr[5 + d]) + 0.5 * ch * (wl[8] - wr[8]), rel=1e-13)
        fw = oracle_mod.flux_dedner(p, ch, wl, wl, d)
        e = mhd_flux(p, wl[:8], d)
        mask = [q for q in range(8) if q != 5 + d]
        np.testing.assert_allclose(fw[mask], e[mask], rtol=1e-12, atol=1e-12)


def euler_flux(g, w, d):
    r, v, pg = w[0], w[1:4], w[4]
    E = pg / (g - 1) + 0.5 * r * np.dot(v, v)
    f = np.zeros(5)
    f[0] = r * v[d]
    f[1:4] = r * v[d] * v
    f[1 + d] += pg
    f[4] = v[d] * (E + pg)
    return f


@pytest.mark.parametrize("d", [0, 1, 2])
def test_vanleer_splitting(oracle_mod, d):
    """F+(W) + F-(W) = F(W) (consistency of the Haenel form) for sub- and
    supersonic states; fully upwind for |M| >= 1."""
    rng = np.random.default_rng(7 + d)
    p = params(2)
    g = p.gamma
    for k in range(20):
        w = np.concatenate([[rng.uniform(0.3, 2)], rng.uniform(-1, 1, 3), [rng.uniform(0.2, 1)]])
        if k % 2:
            w[1 + d] = np.sign(w[1 + d] + 0.1) * 3.0 * np.sqrt(g * w[4] / w[0])
        f = oracle_mod.flux_vanleer(p, w, w, d)
        np.testing.assert_allclose(f, euler_flux(g, w, d), rtol=1e-13, atol=1e-13)
    wl = np.array([1.0, 0, 0, 0, 1.0])
    wl[1 + d] = 2.0 * np.sqrt(g)  # M = 2 on the left, supersonic to the right
    wr = np.array([0.5, 0.1, 0.2, 0.3, 0.4])
    wr[1 + d] = 5.0
    np.testing.assert_allclose(oracle_mod.flux_vanleer(p, wl, wr, d),
                               euler_flux(g, wl, d), rtol=1e-14)


# --------------------------------------------------------------- limiters --
SECOND_ORDER = [mfmhd.LIM_MINMOD, mfmhd.LIM_SUPERBEE, mfmhd.LIM_VANALBADA,
                mfmhd.LIM_MC, mfmhd.LIM_HARMONIC, mfmhd.LIM_HYPERBEE,
                mfmhd.LIM_KOREN, mfmhd.LIM_VENKAT, mfmhd.LIM_VANLEER,
                mfmhd.LIM_WENO3, mfmhd.LIM_WENO3YC]
TVD = [mfmhd.LIM_MINMOD, mfmhd.LIM_SUPERBEE, mfmhd.LIM_VANALBADA, mfmhd.LIM_MC,
       mfmhd.LIM_HARMONIC, mfmhd.LIM_HYPERBEE, mfmhd.LIM_KOREN,
       mfmhd.LIM_VENKAT, mfmhd.LIM_VANLEER]


@pytest.mark.parametrize("lim", SECOND_ORDER)
def test_limiter_linear_exact(oracle_mod, lim):
    """Every second-order limiter reproduces linear data: W -/+ b/2."""
    for w0, b in ((1.0, 0.1), (-2.0, 3.0), (5.0, -0.7)):
        l, r = oracle_mod.limiter(lim, w0 - b, w0, w0 + b)
        assert l == pytest.approx(w0 - 0.5 * b, rel=1e-9, abs=1e-9)
        assert r == pytest.approx(w0 + 0.5 * b, rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("lim", TVD)
def test_limiter_extremum_and_bounds(oracle_mod, lim):
    rng = np.random.default_rng(8)
    for _ in range(200):
        a, b, c = rng.uniform(-1, 1, 3)
        l, r = oracle_mod.limiter(lim, a, b, c)
        if (b - a) * (c - b) <= 0:
            assert l == b and r == b   # first order at extrema
        else:
            lo, hi = min(a, b, c), max(a, b, c)
            if lim != mfmhd.LIM_HYPERBEE:  # slope -> 2 dWL when dWR >> dWL
                assert lo - 1e-6 <= l <= hi + 1e-6 and lo - 1e-6 <= r <= hi + 1e-6


@pytest.mark.parametrize("lim", SECOND_ORDER + [mfmhd.LIM_CD])
def test_limiter_reflection_symmetry(oracle_mod, lim):
    """Mirroring the data swaps the face values."""
    rng = np.random.default_rng(9)
    for _ in range(100):
        a, b, c = rng.uniform(-1, 1, 3)
        l, r = oracle_mod.limiter(lim, a, b, c)
        l2, r2 = oracle_mod.limiter(lim, c, b, a)
        tol = 1e-8 if lim == mfmhd.LIM_VENKAT else 1e-13
        assert l2 == pytest.approx(r, abs=tol) and r2 == pytest.approx(l, abs=tol)


# -------------------------------------------------------- charge exchange --
def helio_p(nf):
    p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, sunN=7.0, sunV=4.5e7,
                           netN=0.1, R0=0.1, dx=1.0 / 64, num_cells=(64, 64, 64),
                           sunXYZ=(0.5, 0.5, 0.5))
    p.nfluids = nf
    return p


@pytest.mark.parametrize("nf", [2, 3, 4])
@pytest.mark.parametrize("reg", [1, 2, 3, 4])
def test_charge_exchange_conservation(oracle_mod, nf, reg):
    """Pauls-Zank sources: momentum and energy leave the plasma and enter
    the neutrals exactly (h + PPLU + PMI = 0 per fluid), total neutral mass
    is conserved and the plasma mass is untouched -- whenever the population
    born in the cell's region exists (neutral 1 in region 1, and 4 except in
    the 3-fluid kernel; neutral k in region k)."""
    born = {1: True, 4: nf != 3, 2: nf >= 3, 3: nf >= 4}[reg]
    rng = np.random.default_rng(20 + nf * 5 + reg)
    p = helio_p(nf)
    for _ in range(10):
        w = rand_prim(rng, nf, bscale=0.05)
        w[0] *= 3
        s = oracle_mod.charge_exchange(p, w, reg, 0.01)
        assert s[0] == 0.0 and np.all(s[5:8] == 0.0)
        neut = [8 + 5 * f for f in range(nf - 1)]
        scale = np.abs(s).max()
        assert scale > 0
        if not born:  # every neutral only loses mass
            assert all(s[b] < 0 for b in neut)
            continue
        assert abs(sum(s[b] for b in neut)) <= 1e-12 * scale
        for q in range(1, 5):
            tot = s[q] + sum(s[b + q] for b in neut)
            assert abs(tot) <= 1e-12 * scale, (q, tot, scale)


def test_charge_exchange_equal_velocities(oracle_mod):
    """Equal bulk velocities: no momentum transfer."""
    p = helio_p(4)
    w = rand_prim(np.random.default_rng(3), 4)
    for f in range(3):
        w[9 + 5 * f:12 + 5 * f] = w[1:4]
    s = oracle_mod.charge_exchange(p, w, 2, 0.01)
    # plasma momentum: h_k ~ (u - u_k) = 0
    assert np.all(np.abs(s[1:4]) == 0.0)


def test_cross_section_monotone_literals(oracle_mod):
    """Single- vs double-precision Fortran literals change the cross
    section by ~1e-8 relative, not more."""
    p = helio_p(2)
    w = rand_prim(np.random.default_rng(4), 2)
    s1 = oracle_mod.charge_exchange(p, w, 1, 0.01)
    p.ce_single_literals = 1
    s2 = oracle_mod.charge_exchange(p, w, 1, 0.01)
    rel = np.abs(s2 - s1).max() / np.abs(s1).max()
    assert 0 < rel < 1e-6


# ---------------------------------------------------------------- regions --
def test_region_truth_table(oracle_mod):
    p = helio_p(4)
    dx = 1.0 / 64
    i, j, k = 45, 32, 32  # on the +x axis from the sun (32 dx away... + 13)
    rhat = np.array([(i + 0.5) * dx - p.sun_xc, (j + 0.5) * dx - p.sun_yc,
                     (k + 0.5) * dx - p.sun_zc])
    rhat /= np.linalg.norm(rhat)
    coeff = p.gamma * p.lism_m ** 2 * p.lism_t

    def w_of(ur, T):
        w = np.zeros(23)
        w[0] = 1.0
        w[1:4] = ur * rhat
        w[4] = T / coeff
        w[5:8] = 0.01
        return w
    fast = 0.7 * p.sun_v
    slow = 0.5 * p.sun_v
    assert oracle_mod.region(p, w_of(fast, 1e5), i, j, k, dx) == 3
    assert oracle_mod.region(p, w_of(fast, 2e6), i, j, k, dx) == 2
    assert oracle_mod.region(p, w_of(-fast, 1e5), i, j, k, dx) == 3
    assert oracle_mod.region(p, w_of(slow, 6e4), i, j, k, dx) == 2
    assert oracle_mod.region(p, w_of(slow, 4e4), i, j, k, dx) == 1


# ----------------------------------------------------------- wave speed ---
def test_max_wave_speed_fast_magnetosonic(oracle_mod):
    rng = np.random.default_rng(11)
    p = params(2)
    g = p.gamma
    for _ in range(20):
        w = rand_prim(rng, 2)
        u = prim_to_cons(p, w)
        a2 = g * w[4] / w[0]
        b2 = D4 * np.dot(w[5:8], w[5:8]) / w[0]
        cmax = 0.0
        for d in range(3):
            bn2 = D4 * w[5 + d] ** 2 / w[0]
            cf = np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bn2)))
            cmax = max(cmax, abs(w[1 + d]) + cf)
        an = np.sqrt(g * w[12] / w[8])
        cmax = max(cmax, np.abs(w[9:12]).max() + an)
        assert oracle_mod.max_wave_speed(p, u) == pytest.approx(cmax, rel=1e-13)


# ------------------------------------------------------------- full steps --
def test_periodic_step_conserves(oracle_mod):
    """With no source terms the update is conservative: every component's
    sum over a periodic box is unchanged to round-off (flux telescoping)."""
    from helpers import grid_shape, periodic_fill, smooth_state, valid_view
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.nfluids = 2
    p.eight_wave = 0
    p.time_approx = mfmhd.TIME_RK2
    ng = 4
    vlo, vhi = (0, 0, 0), (15, 15, 15)
    lo, hi, shape = grid_shape(13, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=2)
    periodic_fill(U, ng)
    r = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.002, 1 / 16)
    a = valid_view(U, lo, vlo, vhi).reshape(13, -1).sum(axis=1)
    b = valid_view(r["U"], lo, vlo, vhi).reshape(13, -1).sum(axis=1)
    scale = np.abs(valid_view(U, lo, vlo, vhi)).reshape(13, -1).max(axis=1) * 4096
    assert np.all(np.abs(a - b) <= 1e-13 * scale), (a - b) / scale


def test_compute_dt_brute_force(oracle_mod):
    from helpers import grid_shape, smooth_state
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.nfluids = 3
    lo, hi, shape = grid_shape(18, (0, 0, 0), (5, 4, 3), 2)
    U = smooth_state(p, shape, lo, seed=9)
    dt, cell = oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (5, 4, 3), 0.1)
    best, bc = None, None
    for k in range(4):
        for j in range(5):
            for i in range(6):
                u = U[:, k - lo[2], j - lo[1], i - lo[0]]
                d = 0.1 / oracle_mod.max_wave_speed(p, u)
                if best is None or d < best:
                    best, bc = d, (i, j, k)
    assert dt == best and cell == bc


def test_compute_dt_helio_excludes_inner_sphere(oracle_mod):
    """HeliosphericProblem::computeDt skips cells within R0 - 2 dx of the
    sun (swlism/HeliosphericProblem.cpp:3809-3860, m_R0dtIgnore at :497):
    a fast spike at the sun does not set dt, one outside the radius does."""
    from helpers import grid_shape
    n = 16
    dx = 1.0 / n
    p = mfmhd.helio_params(sunXYZ=(0.5, 0.5, 0.5), R0=0.3, initR=0.4, netN=0.1,
                           dx=dx, num_cells=(n, n, n))
    p.nfluids = 2
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(13, vlo, vhi, 4)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    r_ign = 0.3 - 2 * dx

    def brute(V):
        best, bc = None, None
        for k in range(n):
            for j in range(n):
                for i in range(n):
                    x, y, z = ((i + 0.5) * dx - 0.5, (j + 0.5) * dx - 0.5, (k + 0.5) * dx - 0.5)
                    if x * x + y * y + z * z <= r_ign * r_ign:
                        continue
                    u = V[:, k - lo[2], j - lo[1], i - lo[0]]
                    d = dx / oracle_mod.max_wave_speed(p, u)
                    if best is None or d < best:
                        best, bc = d, (i, j, k)
        return best, bc

    for (i, j, k) in ((8, 8, 8), (1, 2, 3)):
        V = U.copy()
        V[1, k - lo[2], j - lo[1], i - lo[0]] *= 1e4  # very fast cell
        dt, cell = oracle_mod.compute_dt(p, V, lo, vlo, vhi, dx)
        best, bc = brute(V)
        assert dt == best and cell == bc
        assert (cell == (i, j, k)) == ((i, j, k) == (1, 2, 3))


def test_brio_wu_known_answer(oracle_mod):
    """Brio & Wu (1988) shock tube at t = 0.1 on 400 cells: undisturbed
    states outside the fans and the well-known plateaus (compound-wave
    region rho ~ 0.676, right of the slow shock rho ~ 0.117)."""
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.time_approx = mfmhd.TIME_RK2
    n, ng = 400, 4
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
    assert rho[int(0.45 * n)] == pytest.approx(0.676, rel=0.02)
    assert rho[int(0.75 * n)] == pytest.approx(0.117, rel=0.02)
    assert np.all(rho > 0.1) and np.all(rho <= 1.0 + 1e-12)
