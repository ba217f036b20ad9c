"""GPU parity of the Hancock predictor-corrector (time_approximation = 1, the
scheme of the reference's 3-D multi-fluid deck main/inputs/helio03.inputs)
against the oracle: FP64 <= 1e-12 relative per component and step, FP32
fluxes <= 1e-5 relative L1 per variable."""
import numpy as np
import pytest

import mfmhd
from helpers import (S4PI, GpuPatch, grid_shape, l1_rel_per_comp, l1_rel_vector_scaled,
                     periodic_fill, smooth_state, valid_view, vector_groups)
from test_gpu_parity import TOL32_L1, check64, helio_setup, periodic_params, run_pair

pytestmark = pytest.mark.gpu

NG = 3  # PatchMHDAM::numGhostCells for Hancock


@pytest.mark.parametrize("nf,lim", [(1, mfmhd.LIM_MINMOD), (1, mfmhd.LIM_VANLEER),
                                    (4, mfmhd.LIM_MINMOD), (2, mfmhd.LIM_MC)])
def test_periodic_patch_hancock(oracle_mod, nf, lim):
    print("hancock periodic", nf, lim, flush=True)
    p = periodic_params(nf=nf, flux=mfmhd.FLUX_ROE8, lim=lim, time=mfmhd.TIME_HANCOCK)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 9, 7)
    lo, hi, shape = grid_shape(nc, vlo, vhi, NG)
    U = smooth_state(p, shape, lo, seed=7 + nf)
    periodic_fill(U, NG)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.01, 0.05, want_F=True, want_divB=True)
    check64(g, o, lo, vlo, vhi, U_in=U)
    for d in range(3):
        assert np.max(np.abs(g["F"][d] - o["F"][d])) <= 1e-12 * np.max(np.abs(o["F"][d]))


def test_helio_patch_hancock(oracle_mod):
    """The 3-D 4-fluid heliospheric box with Hancock: boundary fluxes on every
    side, charge exchange in the predictor and the corrector, re-init and
    floors in the corrector; sub-patches touching the boundary."""
    n = 20
    p, dx = helio_setup(oracle_mod, n, 4)
    p.time_approx = mfmhd.TIME_HANCOCK
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, NG)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    t = 0.0
    for step in range(3):
        print("hancock helio step", step, flush=True)
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.2 * dtn, dx, time=t)
        check64(g, o, lo, vlo, vhi)
        U = o["U"]
        t += 0.2 * dtn
        dtn = o["dt"]
    for (a, b) in (((3, 4, 2), (11, 12, 9)), ((0, 10, 12), (7, 19, 19))):
        g, o = run_pair(oracle_mod, p, U, lo, a, b, 0.2 * dtn, dx, time=t)
        check64(g, o, lo, a, b)


def test_brio_wu_hancock(oracle_mod):
    L = [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, 1.0 * S4PI, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -1.0 * S4PI, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.time_approx = mfmhd.TIME_HANCOCK
    n = 64
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 1]
    dx = 1.0 / n
    for vlo, vhi in (((0, 0, 0), (n - 1, 0, 0)), ((0, 0, 0), (27, 0, 0)),
                     ((30, 0, 0), (n - 1, 0, 0))):
        lo, hi, shape = grid_shape(8, vlo, vhi, NG)
        U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
        dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.8 * dtn, dx)
        check64(g, o, lo, vlo, vhi)


def test_fp32_flux_hancock_error(oracle_mod):
    p, dx = helio_setup(oracle_mod, 20, 4, prec=mfmhd.PREC_FP32_FLUX)
    p.time_approx = mfmhd.TIME_HANCOCK
    vlo, vhi = (0, 0, 0), (19, 19, 19)
    nc = mfmhd.num_components(p)
    lo, hi, shape = grid_shape(nc, vlo, vhi, NG)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.2 * dtn, dx)
    assert g["rc"] == 0
    a = valid_view(g["U"], lo, vlo, vhi)
    b = valid_view(o["U"], lo, vlo, vhi)
    err = l1_rel_vector_scaled(a, b, vector_groups(4))
    print("hancock fp32 per-variable L1:", " ".join("%.2e" % e for e in err))
    assert err.max() <= TOL32_L1, err


def test_hancock_rejects_glm():
    p = periodic_params(nf=1, dedner=True, eight=False, time=mfmhd.TIME_HANCOCK)
    with pytest.raises(mfmhd.MfmError) as e:
        mfmhd.Context(p)
    assert e.value.code == mfmhd.ERR_UNSUPPORTED


def test_level_table_hancock(oracle_mod):
    """mfm_level_update_dev with Hancock over uneven patches == one call
    per patch, and a patch == the oracle."""
    from helpers import gpu_update
    p = periodic_params(nf=4, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_MINMOD,
                        time=mfmhd.TIME_HANCOCK)
    nc = mfmhd.num_components(p)
    rng = np.random.default_rng(9)
    states = []
    for q in range(5):
        e = int(rng.choice([4, 8, 12]))
        lo0 = rng.integers(-20, 20, size=3)
        vlo, vhi = tuple(int(v) for v in lo0), tuple(int(v) + e - 1 for v in lo0)
        lo, hi, shape = grid_shape(nc, vlo, vhi, NG)
        states.append((smooth_state(p, shape, lo, seed=30 + q), lo, vlo, vhi))
    ctx = mfmhd.Context(p)
    singles = [gpu_update(ctx, U, lo, vlo, vhi, 0.01, 0.05) for U, lo, vlo, vhi in states]
    pats = [GpuPatch(U, lo) for U, lo, _, _ in states]
    ctx.update_level([(g.fab, mfmhd.Box.of(s[2], s[3])) for g, s in zip(pats, states)],
                     0.01, 0.05)
    mfmhd.synchronize()
    for q, (g, s) in enumerate(zip(pats, singles)):
        assert np.array_equal(g.numpy(), s["U"]), q
    U, lo, vlo, vhi = states[1]
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.01, 0.05)
    check64(singles[1], o, lo, vlo, vhi)
    ctx.close()


def test_served_hancock_request_matches_oracle(oracle_mod):
    """The serving loop with a Hancock equation class: a request run alone
    equals the oracle stepped with the same CFL rule."""
    from test_serve import req_params
    n = 16
    p, dx = req_params(n)
    p.time_approx = mfmhd.TIME_HANCOCK
    srv = mfmhd.Server(p, 8 << 30)
    rid = srv.submit(p, dx, 3, 0.5)
    for _ in range(50):
        na, npd, _ = srv.step()
        if srv.info(rid)["state"] >= mfmhd.REQ_DONE:
            break
    inf = srv.info(rid)
    assert inf["state"] == mfmhd.REQ_DONE, srv.error(rid)
    Ug = srv.fetch(rid)
    srv.close()
    lo, hi = inf["lo"], inf["hi"]
    assert lo == (-NG, -NG, -NG)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    shape = (mfmhd.num_components(p), hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    t = 0.0
    for _ in range(3):
        dt = 0.5 * dtn
        r = oracle_mod.patch_update(p, U, lo, vlo, vhi, dt, dx, time=t)
        assert r["rc"] == 0
        U, dtn = r["U"], r["dt"]
        t += dt
    from helpers import rel_err_per_comp
    err = rel_err_per_comp(valid_view(Ug, lo, vlo, vhi), valid_view(U, lo, vlo, vhi)).max()
    assert err <= 1e-12, err
