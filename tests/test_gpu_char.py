"""GPU parity of characteristic reconstruction (mhdam.char_reconstruction = 1,
the setting of the reference's Orszag-Tang deck main/inputs/orszag.inputs)
against the oracle: FP64 <= 1e-12 relative per component and step (the
FP32-flux build rejects the option).  Covers the interior plasma flux
kernel, the boundary-face kernel (Riemann and heliospheric boundaries), the
Hancock predictor and corrector, GLM, and every limiter."""
import numpy as np
import pytest

import mfmhd
from helpers import S4PI, grid_shape, periodic_fill, smooth_state
from test_gpu_parity import check64, helio_setup, periodic_params, run_pair

pytestmark = pytest.mark.gpu

ALL_LIMS = [mfmhd.LIM_CD, mfmhd.LIM_MINMOD, mfmhd.LIM_SUPERBEE, mfmhd.LIM_VANALBADA,
            mfmhd.LIM_MC, mfmhd.LIM_HARMONIC, mfmhd.LIM_HYPERBEE, mfmhd.LIM_KOREN,
            mfmhd.LIM_VENKAT, mfmhd.LIM_VANLEER, mfmhd.LIM_WENO3, mfmhd.LIM_WENO3YC]


def char_periodic(nf=1, lim=mfmhd.LIM_VANLEER, time=mfmhd.TIME_RK2, dedner=False,
                  prec=mfmhd.PREC_FP64):
    p = periodic_params(nf=nf, flux=mfmhd.FLUX_ROE8, lim=lim, time=time, dedner=dedner,
                        eight=not dedner, prec=prec)
    p.char_reconstruction = 1
    return p


@pytest.mark.parametrize("lim", ALL_LIMS)
def test_periodic_patch_every_limiter(oracle_mod, lim):
    print("char periodic limiter", lim, flush=True)
    p = char_periodic(lim=lim)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (9, 7, 6)
    lo, hi, shape = grid_shape(8, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=40 + lim)
    periodic_fill(U, ng)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.01, 0.05, want_F=True, want_divB=True)
    check64(g, o, lo, vlo, vhi, U_in=U)
    for d in range(3):
        assert np.max(np.abs(g["F"][d] - o["F"][d])) <= 1e-12 * np.max(np.abs(o["F"][d]))


def test_mixed_limiters_four_fluids(oracle_mod):
    """A different limiter per characteristic variable, with neutral fluids
    (which keep their component-wise minmod)."""
    p = char_periodic(nf=4)
    for v, lim in enumerate([mfmhd.LIM_MC, mfmhd.LIM_VANLEER, mfmhd.LIM_MINMOD, mfmhd.LIM_KOREN,
                             mfmhd.LIM_SUPERBEE, mfmhd.LIM_WENO3, mfmhd.LIM_VANALBADA,
                             mfmhd.LIM_HARMONIC]):
        p.limiter[v] = lim
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 9, 7)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=3)
    periodic_fill(U, ng)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.01, 0.05)
    check64(g, o, lo, vlo, vhi, U_in=U)


def test_glm_char(oracle_mod):
    p = char_periodic(dedner=True, lim=mfmhd.LIM_MC)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (9, 9, 9)
    lo, hi, shape = grid_shape(9, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=5)
    periodic_fill(U, ng)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.01, 0.05, ch=2.0, cp=0.5)
    check64(g, o, lo, vlo, vhi)


@pytest.mark.parametrize("nf,lim", [(1, mfmhd.LIM_VANLEER), (2, mfmhd.LIM_MINMOD),
                                    (1, mfmhd.LIM_WENO3YC)])
def test_hancock_char(oracle_mod, nf, lim):
    """The Orszag-Tang deck's scheme: Hancock with characteristic slopes."""
    print("char hancock", nf, lim, flush=True)
    p = char_periodic(nf=nf, lim=lim, time=mfmhd.TIME_HANCOCK)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 9, 7)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=60 + nf)
    periodic_fill(U, ng)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.01, 0.05, want_F=True, want_divB=True)
    check64(g, o, lo, vlo, vhi, U_in=U)


@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_HANCOCK])
def test_helio_char(oracle_mod, time):
    """Heliospheric box (boundary fluxes on every side, charge exchange)."""
    n = 16
    p, dx = helio_setup(oracle_mod, n, 4)
    p.time_approx = time
    p.char_reconstruction = 1
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_VANLEER
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    t = 0.0
    for step in range(2):
        print("char helio", time, step, flush=True)
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.2 * dtn, dx, time=t)
        check64(g, o, lo, vlo, vhi)
        U = o["U"]
        t += 0.2 * dtn
        dtn = o["dt"]


@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_HANCOCK])
def test_brio_wu_char(oracle_mod, time):
    L = [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, 1.0 * S4PI, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -1.0 * S4PI, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.time_approx = time
    p.char_reconstruction = 1
    n = 64
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 1]
    dx = 1.0 / n
    ng = mfmhd.num_ghost_cells(p)
    for vlo, vhi in (((0, 0, 0), (n - 1, 0, 0)), ((30, 0, 0), (n - 1, 0, 0))):
        lo, hi, shape = grid_shape(8, vlo, vhi, ng)
        U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
        dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.8 * dtn, dx)
        check64(g, o, lo, vlo, vhi)


@pytest.mark.parametrize("tvd,time,lim", [(1, mfmhd.TIME_RK2, mfmhd.LIM_VANLEER),
                                          (2, mfmhd.TIME_RK2, mfmhd.LIM_SUPERBEE),
                                          (1, mfmhd.TIME_HANCOCK, mfmhd.LIM_MC)])
def test_check_tvd(oracle_mod, tvd, time, lim):
    """mhdam.check_tvd on top of the characteristic slopes (the reference's
    rotor and blast decks), on rough data so that the clipping acts."""
    print("char check_tvd", tvd, time, lim, flush=True)
    p = char_periodic(lim=lim, time=time)
    p.check_tvd = tvd
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (9, 8, 7)
    lo, hi, shape = grid_shape(8, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=80 + tvd)
    rng = np.random.default_rng(tvd)
    U[1:4] *= 1 + 0.3 * rng.standard_normal(U[1:4].shape)
    U[5:8] *= 1 + 0.3 * rng.standard_normal(U[5:8].shape)
    periodic_fill(U, ng)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.002, 0.05, want_F=True)
    check64(g, o, lo, vlo, vhi, U_in=U)
    p.check_tvd = 0
    o0 = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.002, 0.05)
    assert not np.array_equal(o0["U"], o["U"])  # the clipping acted


def test_helio_check_tvd(oracle_mod):
    n = 16
    p, dx = helio_setup(oracle_mod, n, 4)
    p.char_reconstruction = 1
    p.check_tvd = 1
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.2 * dtn, dx)
    check64(g, o, lo, vlo, vhi)


def test_level_table_char(oracle_mod):
    """mfm_level_update_dev with characteristic slopes over uneven patches ==
    one call per patch, and a patch == the oracle."""
    from helpers import GpuPatch, gpu_update
    p = char_periodic(nf=2, lim=mfmhd.LIM_VANLEER, time=mfmhd.TIME_HANCOCK)
    p.check_tvd = 1
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    rng = np.random.default_rng(19)
    states = []
    for q in range(4):
        e = int(rng.choice([4, 8, 12]))
        lo0 = rng.integers(-20, 20, size=3)
        vlo, vhi = tuple(int(v) for v in lo0), tuple(int(v) + e - 1 for v in lo0)
        lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
        states.append((smooth_state(p, shape, lo, seed=90 + q), lo, vlo, vhi))
    ctx = mfmhd.Context(p)
    singles = [gpu_update(ctx, U, lo, vlo, vhi, 0.01, 0.05) for U, lo, vlo, vhi in states]
    pats = [GpuPatch(U, lo) for U, lo, _, _ in states]
    ctx.update_level([(g.fab, mfmhd.Box.of(s[2], s[3])) for g, s in zip(pats, states)],
                     0.01, 0.05)
    mfmhd.synchronize()
    for q, (g, s) in enumerate(zip(pats, singles)):
        assert np.array_equal(g.numpy(), s["U"]), q
    U, lo, vlo, vhi = states[2]
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.01, 0.05)
    check64(singles[2], o, lo, vlo, vhi)
    ctx.close()


def test_server_char_class():
    """The serving loop keeps characteristic slopes as part of the equation
    class: a request with component-wise slopes is refused."""
    from test_serve import req_params
    n = 12
    p, dx = req_params(n)
    p.char_reconstruction = 1
    srv = mfmhd.Server(p, 1 << 30)
    q = mfmhd.Params.from_buffer_copy(p)
    q.char_reconstruction = 0
    with pytest.raises(mfmhd.MfmError):
        srv.submit(q, dx, 2, 0.5)
    rid = srv.submit(p, dx, 2, 0.5)
    for _ in range(20):
        srv.step()
        if srv.info(rid)["state"] >= mfmhd.REQ_DONE:
            break
    assert srv.info(rid)["state"] == mfmhd.REQ_DONE, srv.error(rid)
    srv.close()


@pytest.mark.parametrize("dedner,time", [(True, mfmhd.TIME_RK2), (False, mfmhd.TIME_HANCOCK),
                                         (False, mfmhd.TIME_FIRST_ORDER)])
def test_orszag_tang_2d_char(oracle_mod, dedner, time):
    """2-D Orszag-Tang with characteristic slopes and check_tvd: with GLM
    cleaning (RK2), with the 8-wave Hancock scheme of the deck, and with
    the first-order time scheme."""
    print("char 2-D orszag", dedner, time, flush=True)
    p = periodic_params(nf=1, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_VANLEER, dims=2,
                        dedner=dedner, eight=not dedner, time=time)
    p.char_reconstruction = 1
    p.check_tvd = 1
    n = 32
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng, dims=2)
    dx = 1.0 / n
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    ch = dx / dtn
    cp = np.sqrt(ch * dx)
    for step in range(2):
        periodic_fill(U, ng, dims=2)
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.5 * dtn, dx, ch=ch, cp=cp)
        check64(g, o, lo, vlo, vhi)
        U = o["U"]
        dtn = o["dt"]
        ch = dx / dtn
        cp = np.sqrt(ch * dx)


def test_char_rejects_unlimited_slopes():
    p = char_periodic()
    p.limiter[2] = mfmhd.LIM_NONE
    with pytest.raises(mfmhd.MfmError) as e:
        mfmhd.Context(p)
    assert e.value.code == mfmhd.ERR_UNSUPPORTED


def test_char_rejects_fp32_build():
    """Characteristic slopes run in the FP64 build only: with FP32 primitives
    the pressure rounding reaches the field slopes through the eigenvectors."""
    p = char_periodic(prec=mfmhd.PREC_FP32_FLUX)
    with pytest.raises(mfmhd.MfmError) as e:
        mfmhd.Context(p)
    assert e.value.code == mfmhd.ERR_UNSUPPORTED
