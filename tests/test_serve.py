"""Ensemble serving (mfm_server_*): continuous batching of independent
heliospheric (3-D) and SW/LISM (2-D axisymmetric) solves admitted by an HBM
budget.

A request's result must not depend on what else is in flight: the per-cell
arithmetic of a patch is the same in any batch, so the FAB of a request run
alone and run in a mix must agree bit for bit.  A request run alone is also
checked against the CPU oracle stepped with the same CFL rule.
"""
import numpy as np
import pytest

import mfmhd
from helpers import l1_rel_vector_scaled, rel_err_per_comp, valid_view, vector_groups

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def req_params(n, sunV=4.5e7, lismB=3e-6, nf=2, prec=mfmhd.PREC_FP64):
    dx = 1.0 / n
    p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, lismT=6527.0, lismB=lismB,
                           lismBdir=(2 ** -0.5, 2 ** -0.5, 0.0),
                           lismUdir=(-1.0, 0.0, 0.0), sunN=7.0, sunV=sunV,
                           sunXYZ=(0.5, 0.5, 0.5), R0=0.15, initR=0.3,
                           netN=0.1, dx=dx, num_cells=(n, n, n))
    p.nfluids = nf
    p.time_approx = mfmhd.TIME_RK2
    p.precision = prec
    return p, dx


def run_until_done(srv, ids, max_iter=200, hook=None):
    for it in range(max_iter):
        if hook:
            hook(it)
        na, npd, _ = srv.step()
        if na == 0 and npd == 0 and all(srv.info(r)["state"] >= mfmhd.REQ_DONE for r in ids()):
            return it + 1
    raise AssertionError("server did not drain")


def test_request_alone_equals_request_in_mix():
    pa, dxa = req_params(16)
    big = 8 << 30
    # alone
    s1 = mfmhd.Server(pa, big)
    a = s1.submit(pa, dxa, 4, 0.5)
    run_until_done(s1, lambda: [a])
    assert s1.info(a)["state"] == mfmhd.REQ_DONE
    Ua = s1.fetch(a)
    s1.close()
    # in a mix: two requests already running, A arrives later, one more after
    s2 = mfmhd.Server(pa, big)
    pb, dxb = req_params(20, sunV=6.0e7, lismB=2e-6)
    pc, dxc = req_params(12, sunV=3.5e7, lismB=5e-6)
    pd, dxd = req_params(24, sunV=8.0e7, lismB=4e-6)
    ids = [s2.submit(pb, dxb, 6, 0.5), s2.submit(pc, dxc, 3, 0.5)]
    late = {}

    def arrivals(it):
        if it == 2:
            late["a"] = s2.submit(pa, dxa, 4, 0.5)
            ids.append(late["a"])
        if it == 3:
            ids.append(s2.submit(pd, dxd, 2, 0.5))

    run_until_done(s2, lambda: list(ids), hook=arrivals)
    for r in ids:
        assert s2.info(r)["state"] == mfmhd.REQ_DONE, s2.info(r)
    Um = s2.fetch(late["a"])
    st = s2.stats()
    s2.close()
    assert np.array_equal(Ua, Um)
    assert st["iterations"] >= 6 and st["p99_ms"] >= st["p50_ms"] > 0.0


def test_served_request_matches_oracle(oracle_mod):
    n = 16
    p, dx = req_params(n)
    srv = mfmhd.Server(p, 8 << 30)
    rid = srv.submit(p, dx, 3, 0.5)
    run_until_done(srv, lambda: [rid])
    Ug = srv.fetch(rid)
    inf = srv.info(rid)
    srv.close()
    lo, hi = inf["lo"], inf["hi"]
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
    assert abs(inf["t"] - t) <= 1e-12 * t
    err = rel_err_per_comp(valid_view(Ug, lo, vlo, vhi), valid_view(U, lo, vlo, vhi)).max()
    assert err <= TOL64, err


def _a256(x):
    return (x + 255) // 256 * 256


def request_charge(n, nc, fs=8, ng=4):
    """Device bytes the server charges a cubic RK2 request: state, U*, W and
    the face fluxes / normal B of the first-stage box (valid + 2)."""
    cells = (n + 2 * ng) ** 3
    b = 2 * _a256(cells * nc * 8) + _a256(cells * nc * fs)
    v1 = n + 4
    f = (v1 + 1) * v1 * v1
    b += 3 * (_a256(f * nc * fs) + _a256(f * fs))
    return b + 4096


def test_hbm_budget_limits_admission():
    p, dx = req_params(16)
    nc = mfmhd.num_components(p)
    # a request larger than the whole budget fails instead of waiting
    srv = mfmhd.Server(p, 1.0)
    r = srv.submit(p, dx, 1, 0.5)
    srv.step()
    assert srv.info(r)["state"] == mfmhd.REQ_FAILED
    srv.close()
    # room for two requests at a time: the others wait until one is released
    srv = mfmhd.Server(p, 2.5 * request_charge(16, nc))
    ids = [srv.submit(p, dx, 3, 0.5) for _ in range(4)]
    peak, waited = 0, False
    for _ in range(40):
        na, npd, _ = srv.step()
        peak = max(peak, na)
        waited = waited or npd > 0
        for q in list(ids):
            if srv.info(q)["state"] == mfmhd.REQ_DONE:
                srv.release(q)
                ids.remove(q)
        if na == 0 and npd == 0 and not ids:
            break
    assert not ids
    assert peak == 2 and waited
    srv.close()


def swlism_req(n, sunV=4.5e7, lismB=1.5e-6, nf=2, prec=mfmhd.PREC_FP64):
    """A SW/LISM request as the ensemble's: an n x n grid over the deck's
    1500 AU, the sun at 600 AU on the axis, first order in time, Roe-averaged
    8-wave fluxes."""
    dx = 1500.0 / n
    p = mfmhd.swlism_params(sunV=sunV, lismB=lismB, sunXY=(600.0, 0.0), R0=max(10.0, 2.6 * dx),
                            initR=max(80.0, 8.0 * dx), dx=dx, num_cells=(n, n), nfluids=nf)
    p.time_approx = mfmhd.TIME_FIRST_ORDER
    p.flux = mfmhd.FLUX_ROE8_ROEAVG
    p.precision = prec
    return p, dx


def test_swlism_served_request_matches_oracle(oracle_mod):
    """The deck's step control (initial_cfl 0.1, cfl 0.25, max_dt_growth
    1.02) on a served SW/LISM request against the oracle stepped alike."""
    n = 40
    p, dx = swlism_req(n)
    srv = mfmhd.Server(p, 8 << 30)
    rid = srv.submit(p, dx, 4, 0.25, 0.1, 1.02)
    run_until_done(srv, lambda: [rid])
    Ug = srv.fetch(rid)
    inf = srv.info(rid)
    srv.close()
    lo, hi = inf["lo"], inf["hi"]
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, 0)
    shape = (mfmhd.num_components(p), 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    raw, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    t, dt = 0.0, 0.1 * raw
    for _ in range(4):
        r = oracle_mod.patch_update(p, U, lo, vlo, vhi, dt, dx, time=t)
        assert r["rc"] == 0
        U = r["U"]
        t += dt
        dt = min(0.25 * r["dt"], 1.02 * dt)
    assert abs(inf["t"] - t) <= 1e-12 * t
    err = rel_err_per_comp(valid_view(Ug, lo, vlo, vhi), valid_view(U, lo, vlo, vhi)).max()
    assert err <= TOL64, err


def test_swlism_request_alone_equals_request_in_mix():
    pa, dxa = swlism_req(48, sunV=6.0e7, lismB=2.5e-6)
    s1 = mfmhd.Server(pa, 8 << 30)
    a = s1.submit(pa, dxa, 5, 0.25, 0.1, 1.02)
    run_until_done(s1, lambda: [a])
    Ua = s1.fetch(a)
    s1.close()
    s2 = mfmhd.Server(pa, 8 << 30)
    others = [swlism_req(64, sunV=3.5e7, lismB=5e-6), swlism_req(72, sunV=8.0e7, lismB=2e-6)]
    ids = [s2.submit(p, dx, 7, 0.25, 0.1, 1.02) for p, dx in others]
    late = {}

    def arrivals(it):
        if it == 1:
            late["a"] = s2.submit(pa, dxa, 5, 0.25, 0.1, 1.02)
            ids.append(late["a"])

    run_until_done(s2, lambda: list(ids), hook=arrivals)
    for r in ids:
        assert s2.info(r)["state"] == mfmhd.REQ_DONE, s2.info(r)
    Um = s2.fetch(late["a"])
    s2.close()
    assert np.array_equal(Ua, Um)


def test_server_refuses_another_class():
    """A 2-D axisymmetric request is not of a 3-D heliospheric server's
    class (nor the other way round), and a request of another solver family
    (7-wave Roe, HLLC, Rusanov) is not of an 8-wave Roe server's; the
    averaging of one family may differ (a per-request constant)."""
    ph, dxh = req_params(16)
    ps, dxs = swlism_req(40)
    srv = mfmhd.Server(ph, 8 << 30)
    with pytest.raises(mfmhd.MfmError):
        srv.submit(ps, dxs, 1, 0.25)
    srv.close()
    srv = mfmhd.Server(ps, 8 << 30)
    with pytest.raises(mfmhd.MfmError):
        srv.submit(ph, dxh, 1, 0.5)
    for flux in (mfmhd.FLUX_ROE7, mfmhd.FLUX_ROE7_ROEAVG, mfmhd.FLUX_HLLC, mfmhd.FLUX_RUSANOV):
        q, dq = swlism_req(40)
        q.flux = flux
        with pytest.raises(mfmhd.MfmError):
            srv.submit(q, dq, 1, 0.25)
    q, dq = swlism_req(40)
    q.flux = mfmhd.FLUX_ROE8
    srv.submit(q, dq, 1, 0.25)
    srv.close()


def test_server_stage_structure():
    """The stage structure is fixed per server: AUTO is refused, and so is a
    change after the first request; with the summed stages (FP32-flux
    build) a request is still the same alone and in a mix, bit for bit,
    and within the FP32 bar of the separate kernels' result."""
    srv = mfmhd.Server(swlism_req(40)[0], 8 << 30)
    with pytest.raises(mfmhd.MfmError):
        srv.set_stage_kernel(mfmhd.STAGE_AUTO)
    p, dx = swlism_req(40)
    srv.submit(p, dx, 1, 0.25)
    with pytest.raises(mfmhd.MfmError):
        srv.set_stage_kernel(mfmhd.STAGE_SUMMED)
    srv.close()
    pa, dxa = swlism_req(64, sunV=6.0e7, lismB=2.5e-6, prec=mfmhd.PREC_FP32_FLUX)
    res = {}
    for mode in (mfmhd.STAGE_SUMMED, mfmhd.STAGE_SEPARATE):
        s1 = mfmhd.Server(pa, 8 << 30)
        s1.set_stage_kernel(mode)
        a = s1.submit(pa, dxa, 4, 0.25, 0.1, 1.02)
        run_until_done(s1, lambda: [a])
        res[mode] = s1.fetch(a)
        s1.close()
    s2 = mfmhd.Server(pa, 8 << 30)
    s2.set_stage_kernel(mfmhd.STAGE_SUMMED)
    ids = [s2.submit(*swlism_req(m, sunV=v, lismB=b, prec=mfmhd.PREC_FP32_FLUX), 6, 0.25, 0.1, 1.02)
           for m, v, b in ((72, 3.5e7, 5e-6), (96, 8.0e7, 2e-6))]
    a = s2.submit(pa, dxa, 4, 0.25, 0.1, 1.02)
    run_until_done(s2, lambda: ids + [a])
    Um = s2.fetch(a)
    s2.close()
    assert np.array_equal(res[mfmhd.STAGE_SUMMED], Um)
    # two FP32 results, each within 1e-5 of the FP64 one (vector norm)
    err = l1_rel_vector_scaled(res[mfmhd.STAGE_SUMMED], res[mfmhd.STAGE_SEPARATE],
                               vector_groups(2))
    assert err.max() <= 2e-5, err
