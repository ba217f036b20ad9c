"""Pins of the oracle's characteristic reconstruction (mhdam.char_reconstruction
= 1; Reconstruction::faceValues, src/Reconstruction.cpp:178-204, with
CHARANALYSISF / CHARSYNTHESISF, srcIdeal/PatchIdealMHDF.ChF:1167-1505, and
the *FACESLOPES limiters of src/reconstructionF.ChF), on the CPU.

What fixes it, independently of its code:
  - every column of the synthesis is a right eigenvector of the 1-D ideal
    MHD system in primitive variables (normal field a parameter, Gaussian
    units) with the eigenvalue of its family: entropy and div(B) u, Alfven
    u -+ |Bn|/sqrt(4 pi rho), fast u -+ c_f, slow u -+ c_s, in the ordering
    the analysis gives;
  - the analysis is the inverse of the synthesis;
  - each *FACESLOPES limiter gives the slopes of the matching
    *FACEVALUES limiter (already pinned), centred differences halve;
  - on data linear in the primitive variables every limiter reproduces the
    linear face values W -+ dW/2, whatever the eigenvectors;
  - when the two one-sided differences of a cell are multiples a, b of one
    right eigenvector r_k of that cell, the face values are
    W -+ slope(a, b) r_k with slope the limiter of variable k;
  - check_tvd (CHECKTVD): the clipped face values lie between the cell
    value and its neighbour, within the smaller one-sided difference, on
    exactly the variables the option names; linear data pass unchanged;
  - whole steps: a uniform state stays, a periodic box conserves, the
    Brio-Wu shock tube keeps its plateaus.
"""
import numpy as np
import pytest

import mfmhd
from helpers import periodic_fill, smooth_state, valid_view

FOURPI = 4 * np.pi
LIMS = [mfmhd.LIM_MINMOD, mfmhd.LIM_SUPERBEE, mfmhd.LIM_VANALBADA, mfmhd.LIM_MC,
        mfmhd.LIM_HARMONIC, mfmhd.LIM_HYPERBEE, mfmhd.LIM_KOREN, mfmhd.LIM_VENKAT,
        mfmhd.LIM_VANLEER, mfmhd.LIM_WENO3, mfmhd.LIM_WENO3YC]


def char_params(lim=mfmhd.LIM_VANLEER, time=mfmhd.TIME_RK2, n=(8, 8, 8)):
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.char_reconstruction = 1
    p.time_approx = time
    for c in range(8):
        p.limiter[c] = lim
    for d in range(3):
        p.domain_lo[d] = 0
        p.domain_hi[d] = n[d] - 1
        p.periodic[d] = 1
    return p


def jacobian(w, idir, gamma):
    """1-D ideal MHD in primitive variables (rho, u, v, w, p, B) along idir,
    B in Gaussian units, the normal field a constant parameter."""
    n, t1, t2 = 1 + idir, 1 + (idir + 1) % 3, 1 + (idir + 2) % 3
    bn, b1, b2 = 5 + idir, 5 + (idir + 1) % 3, 5 + (idir + 2) % 3
    rho, un, p = w[0], w[n], w[4]
    Bn, B1, B2 = w[bn], w[b1], w[b2]
    A = np.zeros((8, 8))
    for c in range(8):
        A[c, c] = un
    A[0, n] = rho
    A[n, 4] = 1 / rho
    A[n, b1] = B1 / (FOURPI * rho)
    A[n, b2] = B2 / (FOURPI * rho)
    A[t1, b1] = -Bn / (FOURPI * rho)
    A[t2, b2] = -Bn / (FOURPI * rho)
    A[4, n] = gamma * p
    A[b1, n] = B1
    A[b1, t1] = -Bn
    A[b2, n] = B2
    A[b2, t2] = -Bn
    return A


def speeds(w, idir, gamma):
    n = 1 + idir
    rho, p = w[0], w[4]
    B = w[5:8]
    a2 = gamma * p / rho
    b2 = B @ B / (FOURPI * rho)
    bn2 = w[5 + idir] ** 2 / (FOURPI * rho)
    disc = np.sqrt(max((a2 + b2) ** 2 - 4 * a2 * bn2, 0.0))
    cf = np.sqrt(0.5 * (a2 + b2 + disc))
    cs = np.sqrt(max(0.5 * (a2 + b2 - disc), 0.0))
    return w[n], np.sqrt(bn2), cf, cs


STATES = [
    [1.0, 0.1, -0.2, 0.3, 0.8, 1.1, 0.7, -0.4],
    [0.3, -0.5, 0.2, 0.0, 2.0, -1.7, 0.2, 0.9],      # Bn < 0
    [2.5, 0.0, 0.0, 0.0, 0.1, 0.3, -2.0, 1.5],       # low beta
    [1.0, 0.2, 0.1, 0.0, 1.0, 4.0, 0.0, 0.0],        # Bt = 0, bn > c
    [1.0, 0.2, 0.1, 0.0, 3.0, 0.5, 0.0, 0.0],        # Bt = 0, bn < c
]


@pytest.mark.parametrize("si", range(len(STATES)))
@pytest.mark.parametrize("idir", [0, 1, 2])
def test_synthesis_columns_are_eigenvectors(oracle_mod, si, idir):
    p = char_params()
    w = np.array(STATES[si], dtype=float)
    # the states are given with the normal field in slot 5: rotate them
    B = np.roll(w[5:8], idir)
    V = np.roll(w[1:4], idir)
    w[1:4], w[5:8] = V, B
    A = jacobian(w, idir, p.gamma)
    un, ca, cf, cs = speeds(w, idir, p.gamma)
    fams = [{un}, {un}, {un - ca, un + ca}, {un - ca, un + ca},
            {un - cf, un + cf}, {un - cf, un + cf}, {un - cs, un + cs}, {un - cs, un + cs}]
    R = np.stack([oracle_mod.char_synthesis(p, w, idir, np.eye(8)[k]) for k in range(8)], 1)
    seen = []
    for k in range(8):
        r = R[:, k]
        assert np.linalg.norm(r) > 0, k
        lam = r @ A @ r / (r @ r)
        assert np.linalg.norm(A @ r - lam * r) <= 1e-11 * np.linalg.norm(A) * np.linalg.norm(r), (k, lam)
        assert min(abs(lam - f) for f in fams[k]) <= 1e-10 * (abs(un) + cf), (k, lam, fams[k])
        seen.append(lam)
    # the two members of a family are the two signs
    for a, b in ((2, 3), (4, 5), (6, 7)):
        if min(fams[a]) != max(fams[a]):
            assert abs(seen[a] - seen[b]) > 0.5 * (max(fams[a]) - min(fams[a])), (a, seen)
    assert abs(np.linalg.det(R)) > 1e-8


@pytest.mark.parametrize("si", [0, 1, 2, 4])
@pytest.mark.parametrize("idir", [0, 2])
def test_analysis_inverts_synthesis(oracle_mod, si, idir):
    p = char_params()
    w = np.array(STATES[si], dtype=float)
    R = np.stack([oracle_mod.char_synthesis(p, w, idir, np.eye(8)[k]) for k in range(8)], 1)
    L = np.stack([oracle_mod.char_analysis(p, w, idir, np.eye(8)[k]) for k in range(8)], 1)
    assert np.allclose(L @ R, np.eye(8), atol=1e-12), L @ R


def test_zero_field_sign_does_not_matter(oracle_mod):
    """Where the normal field is exactly zero the transforms stay inverse
    whatever the sign of that zero, and give the same basis for +0 and -0."""
    p = char_params()
    for bx in (0.0, -0.0):
        w = np.array([1.2, 0.1, 0.2, -0.1, 0.9, bx, 0.0, 0.0])
        R = np.stack([oracle_mod.char_synthesis(p, w, 0, np.eye(8)[k]) for k in range(8)], 1)
        L = np.stack([oracle_mod.char_analysis(p, w, 0, np.eye(8)[k]) for k in range(8)], 1)
        assert np.allclose(L @ R, np.eye(8), atol=1e-12), (bx, L @ R)
        if bx == 0.0 and np.copysign(1.0, bx) > 0:
            Lp = L
    assert np.array_equal(L, Lp)


@pytest.mark.parametrize("lim", LIMS)
def test_face_slopes_match_face_values(oracle_mod, lim):
    rng = np.random.default_rng(lim)
    for _ in range(300):
        dsl, dsr = rng.standard_normal(2) * 10.0 ** rng.integers(-3, 2, 2)
        w0 = rng.standard_normal()
        left, right = oracle_mod.limiter(lim, w0 - dsl, w0, w0 + dsr)
        sl, sr = oracle_mod.face_slopes(lim, dsl, dsr)
        scale = abs(w0) + abs(dsl) + abs(dsr)
        assert abs((w0 - sl) - left) <= 1e-13 * scale, (dsl, dsr)
        assert abs((w0 + sr) - right) <= 1e-13 * scale, (dsl, dsr)


def test_centred_face_slopes_halve(oracle_mod):
    for dsl, dsr in ((1.0, -3.0), (0.25, 0.5), (-2.0, 0.0)):
        assert oracle_mod.face_slopes(mfmhd.LIM_CD, dsl, dsr) == (0.5 * dsl, 0.5 * dsr)


def _line(nx, base, grad):
    """Primitive FAB (8, 1, 1, nx) linear in x: base + grad * i."""
    W = np.zeros((8, 1, 1, nx))
    for c in range(8):
        W[c, 0, 0, :] = base[c] + grad[c] * np.arange(nx)
    return W


@pytest.mark.parametrize("lim", LIMS + [mfmhd.LIM_CD])
def test_linear_data_gives_linear_face_values(oracle_mod, lim):
    p = char_params(lim)
    base = np.array([1.0, 0.1, -0.2, 0.3, 0.8, 1.1, 0.7, -0.4])
    grad = np.array([0.01, 0.003, -0.002, 0.004, 0.02, 0.0, 0.005, -0.003])
    W = _line(7, base, grad)
    WM, WP = oracle_mod.face_values(p, W, (-1, 0, 0), 0, (0, 0, 0), (4, 0, 0))
    for c in range(8):
        # limiters regularised by a small dSave differ from the linear values
        # by O(dSave / slope^2) of the slope
        tol = 1e-6 if lim in (mfmhd.LIM_VANALBADA, mfmhd.LIM_WENO3, mfmhd.LIM_WENO3YC,
                              mfmhd.LIM_KOREN, mfmhd.LIM_VENKAT) else 1e-13
        v = W[c, 0, 0, 1:6]
        assert np.allclose(WM[c, 0, 0, 1:6], v - 0.5 * grad[c], atol=tol, rtol=0), (c, lim)
        assert np.allclose(WP[c, 0, 0, 1:6], v + 0.5 * grad[c], atol=tol, rtol=0), (c, lim)


@pytest.mark.parametrize("k", range(8))
@pytest.mark.parametrize("lim", [mfmhd.LIM_MINMOD, mfmhd.LIM_VANLEER, mfmhd.LIM_KOREN,
                                 mfmhd.LIM_SUPERBEE])
def test_single_wave_projection(oracle_mod, k, lim):
    p = char_params(mfmhd.LIM_MC)
    p.limiter[k] = lim          # variable k of the characteristic set takes lim
    w = np.array(STATES[0], dtype=float)
    idir = 0
    r = oracle_mod.char_synthesis(p, w, idir, np.eye(8)[k])
    a, b = 0.01, 0.025          # one-sided differences a r, b r
    W = np.zeros((8, 1, 1, 3))
    W[:, 0, 0, 0] = w - a * r
    W[:, 0, 0, 1] = w
    W[:, 0, 0, 2] = w + b * r
    WM, WP = oracle_mod.face_values(p, W, (0, 0, 0), idir, (1, 0, 0), (1, 0, 0))
    sl, sr = oracle_mod.face_slopes(lim, a, b)
    assert np.allclose(WM[:, 0, 0, 1], w - sl * r, atol=1e-13, rtol=1e-12)
    assert np.allclose(WP[:, 0, 0, 1], w + sr * r, atol=1e-13, rtol=1e-12)


def _rough_line(seed, nx=12):
    rng = np.random.default_rng(seed)
    W = np.zeros((8, 1, 1, nx))
    base = np.array(STATES[0])
    for c in range(8):
        W[c, 0, 0, :] = base[c] + 0.2 * abs(base[c] + 0.5) * rng.standard_normal(nx)
    W[0] = np.abs(W[0]) + 0.2
    W[4] = np.abs(W[4]) + 0.2
    return W


def _tvd_violations(W, WM, WP, comps):
    n = 0
    for c in comps:
        w = W[c, 0, 0]
        for i in range(1, w.size - 1):
            dl, dr = w[i] - w[i - 1], w[i + 1] - w[i]
            for s in (w[i] - WM[c, 0, 0, i], WP[c, 0, 0, i] - w[i]):
                if s * dl < -1e-15 or s * dr < -1e-15 or abs(s) > min(abs(dl), abs(dr)) + 1e-15:
                    n += 1
    return n


@pytest.mark.parametrize("lim", [mfmhd.LIM_VANLEER, mfmhd.LIM_SUPERBEE, mfmhd.LIM_CD])
def test_check_tvd_bounds_face_values(oracle_mod, lim):
    """check_tvd = 1: every plasma face value lies between the cell value and
    its neighbour on that side, within the smaller one-sided difference;
    without the check the characteristic values break that somewhere;
    check_tvd = 2 clips density and pressure and leaves the rest alone."""
    W = _rough_line(11)
    box = ((1, 0, 0), (10, 0, 0))
    p = char_params(lim)
    WM0, WP0 = oracle_mod.face_values(p, W, (0, 0, 0), 0, *box)
    assert _tvd_violations(W, WM0, WP0, range(8)) > 0
    p.check_tvd = 1
    WM1, WP1 = oracle_mod.face_values(p, W, (0, 0, 0), 0, *box)
    assert _tvd_violations(W, WM1, WP1, range(8)) == 0
    p.check_tvd = 2
    WM2, WP2 = oracle_mod.face_values(p, W, (0, 0, 0), 0, *box)
    assert _tvd_violations(W, WM2, WP2, (0, 4)) == 0
    for c in (1, 2, 3, 5, 6, 7):
        assert np.array_equal(WM2[c], WM0[c]) and np.array_equal(WP2[c], WP0[c])
    for c in (0, 4):
        assert np.array_equal(WM2[c], WM1[c]) and np.array_equal(WP2[c], WP1[c])


def test_check_tvd_leaves_linear_data(oracle_mod):
    p = char_params(mfmhd.LIM_MC)
    p.check_tvd = 1
    base = np.array(STATES[0])
    grad = np.array([0.01, 0.003, -0.002, 0.004, 0.02, 0.0, 0.005, -0.003])
    W = _line(7, base, grad)
    WM, WP = oracle_mod.face_values(p, W, (-1, 0, 0), 0, (0, 0, 0), (4, 0, 0))
    for c in range(8):
        v = W[c, 0, 0, 1:6]
        assert np.allclose(WM[c, 0, 0, 1:6], v - 0.5 * grad[c], atol=1e-13, rtol=0)
        assert np.allclose(WP[c, 0, 0, 1:6], v + 0.5 * grad[c], atol=1e-13, rtol=0)


def test_uniform_state_unchanged(oracle_mod):
    ng = 4
    p = char_params(mfmhd.LIM_VANLEER, mfmhd.TIME_RK2)
    p.eight_wave = 1
    w = np.array([1.3, 0.2, -0.1, 0.05, 0.9, 0.4, 0.3, -0.2])
    u = np.zeros(8)
    u[0] = w[0]
    u[1:4] = w[0] * w[1:4]
    u[5:8] = w[5:8]
    u[4] = w[4] / (p.gamma - 1) + 0.5 * w[0] * (w[1:4] ** 2).sum() + (w[5:8] ** 2).sum() / (8 * np.pi)
    U = np.broadcast_to(u[:, None, None, None], (8, 16, 16, 16)).copy()
    lo = (-ng,) * 3
    r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (7, 7, 7), 1e-3, 1.0 / 8)
    assert r["rc"] == 0
    v = valid_view(r["U"], lo, (0, 0, 0), (7, 7, 7))
    assert np.allclose(v, u[:, None, None, None], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("time,ng", [(mfmhd.TIME_RK2, 4), (mfmhd.TIME_HANCOCK, 3)])
def test_periodic_conservation(oracle_mod, time, ng):
    n = (10, 10, 10)
    p = char_params(mfmhd.LIM_VANLEER, time, n)
    p.eight_wave = 0            # the 8-wave source is not in conservation form
    lo = (-ng,) * 3
    shape = (8,) + tuple(x + 2 * ng for x in n)
    U = smooth_state(p, shape, lo, seed=12)
    periodic_fill(U, ng)
    dx = 0.1
    dt0, _ = oracle_mod.compute_dt(p, U, lo, (0, 0, 0), (9, 9, 9), dx)
    vb = lambda A: valid_view(A, lo, (0, 0, 0), (9, 9, 9))
    before = vb(U).sum(axis=(1, 2, 3))
    for _ in range(2):
        r = oracle_mod.patch_update(p, U, lo, (0, 0, 0), (9, 9, 9), 0.3 * dt0, dx)
        assert r["rc"] == 0
        U = periodic_fill(r["U"], ng)
    after = vb(U).sum(axis=(1, 2, 3))
    scale = np.abs(vb(U)).sum(axis=(1, 2, 3))
    assert np.all(np.abs(after - before) <= 1e-13 * scale), (after - before) / scale


def test_char_differs_from_componentwise(oracle_mod):
    """The switch is live: the characteristic and the component-wise
    reconstructions of the same non-linear data differ."""
    p = char_params(mfmhd.LIM_MINMOD)
    W = np.zeros((8, 1, 1, 5))
    rng = np.random.default_rng(3)
    W[:] = np.array(STATES[0])[:, None, None, None] * (1 + 0.1 * rng.standard_normal((8, 1, 1, 5)))
    a = oracle_mod.face_values(p, W, (0, 0, 0), 0, (1, 0, 0), (3, 0, 0))
    p.char_reconstruction = 0
    b = oracle_mod.face_values(p, W, (0, 0, 0), 0, (1, 0, 0), (3, 0, 0))
    assert np.abs(a[0] - b[0]).max() > 1e-4


def test_brio_wu_char(oracle_mod):
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.char_reconstruction = 1
    for c in range(8):
        p.limiter[c] = mfmhd.LIM_MINMOD
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
    assert rho[int(0.45 * n)] == pytest.approx(0.676, rel=0.02)
    assert rho[int(0.75 * n)] == pytest.approx(0.117, rel=0.02)
    assert np.all(rho > 0.1) and np.all(rho <= 1.0 + 1e-12)


def test_rejects_unlimited_plasma_slopes(oracle_mod):
    p = char_params(mfmhd.LIM_MINMOD)
    p.limiter[3] = mfmhd.LIM_NONE
    U = np.ones((8, 10, 10, 10))
    r = oracle_mod.patch_update(p, U, (-1, -1, -1), (0, 0, 0), (7, 7, 7), 1e-3, 0.1)
    assert r["rc"] == mfmhd.ERR_UNSUPPORTED
