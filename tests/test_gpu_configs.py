"""The stated workloads at their stated sizes, FP64 build, against the
oracle through the C ABI:

- Brio-Wu, 800 cells, Roe 8-wave, minmod, RK2, CFL 0.8, to t = 0.1: every
  step from the oracle's state agrees to 1e-12, and the GPU's own trajectory
  ends within 1e-10 of the oracle's;
- the 3-D shock tube of riemann14.inputs (400 x 4 x 4), 7-wave Roe
  (flux 11), MC limiter, entropy fix 0.3, Hancock, 8-wave terms off, CFL
  0.5 with dt growth <= 1.1, to t = 0.15, with z periodic like y (the
  deck's non-periodic z reads ghost primitives the reference never sets:
  refused): the same two bars;
- Orszag-Tang with GLM cleaning, 512^2, five steps to 1e-12 each;
- 40^3 windows of the 256^3 four-fluid heliospheric run (around the sun, a
  domain corner, the bow): the full-box GPU update restricted to each window
  equals the oracle's update of the window to 1e-12, and the windows hold
  regions 1, 2 and 3 of the charge-exchange partition;
- the nested three-level batched-AMR hierarchy (about 3000 patches of edge
  8/16/32, Gaussian blobs): per level, the hipGraph capture and its replay
  give the eager level update bit for bit, and sampled patches equal the
  oracle to 1e-12."""
import importlib.util
import math
import os
import sys

import numpy as np
import pytest

import mfmhd
from helpers import (gpu_update, grid_shape, periodic_fill, rel_err_per_comp,
                     valid_view)
from test_gpu_parity import TOL64, check64

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(16, os.cpu_count() or 1)


def _bench():
    spec = importlib.util.spec_from_file_location("bench_workloads",
                                                  os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_brio_wu_800_to_t01(oracle_mod):
    n, cfl, t_end = 800, 0.8, 0.1
    S4PI = math.sqrt(4 * math.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, S4PI, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -S4PI, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.spacedim = 2
    p.time_approx = mfmhd.TIME_RK2
    p.flux = mfmhd.FLUX_ROE8
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_MINMOD
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 0]
    ng = mfmhd.num_ghost_cells(p)
    dx = 1.0 / n
    vlo, vhi = (0, 0, 0), (n - 1, 0, 0)
    lo, hi, shape = grid_shape(8, vlo, vhi, ng, dims=2)
    U0 = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    ctx = mfmhd.Context(p)
    # the GPU's own trajectory, on the device throughout
    Ud = mfmhd.DeviceArray(shape, data=U0)
    fd = mfmhd.fab(Ud.ptr, lo, hi, 8)
    valid = mfmhd.Box.of(vlo, vhi)
    Uo, t, dtn, steps = U0.copy(), 0.0, None, 0
    dtn, _ = oracle_mod.compute_dt(p, Uo, lo, vlo, vhi, dx)
    while t < t_end * (1 - 1e-12):
        dt = min(cfl * dtn, t_end - t)
        periodic_fill(Uo, ng, dims=2)
        g = gpu_update(ctx, Uo, lo, vlo, vhi, dt, dx, t)
        o = oracle_mod.patch_update(p, Uo, lo, vlo, vhi, dt, dx, t)
        check64(g, o, lo, vlo, vhi)
        ctx.fill_periodic(fd, valid)
        ctx.update_state(fd, valid, dt, dx, t)
        Uo, dtn = o["U"], o["dt"]
        t += dt
        steps += 1
    assert steps > 300
    a = valid_view(Ud.numpy(), lo, vlo, vhi)
    b = valid_view(Uo, lo, vlo, vhi)
    err = rel_err_per_comp(a, b)
    assert err.max() <= 1e-10, err
    # the fixed states at both ends of the tube are untouched at t = 0.1
    assert abs(b[0, 0, 0, 0] - 1.0) < 1e-12 and abs(b[0, 0, 0, -1] - 0.125) < 1e-12
    ctx.close()
    Ud.free()


def test_riemann14_deck_to_t015(oracle_mod):
    n, t_end = 400, 0.15
    L = [0.18405, 3.89640, 0.53610, 2.48660, 0.36410, 4.0, 2.39400, 1.19700]
    R = [0.1, -5.5, 0.0, 0.0, 0.1, 4.0, 2.0, 1.0]
    p = mfmhd.riemann_params(1.666666666666667, L, R, 0.5)
    p.spacedim = 3
    p.time_approx = mfmhd.TIME_HANCOCK
    p.flux = mfmhd.FLUX_ROE7
    p.entropy_fix = 0.3
    p.eight_wave = 0
    p.divb_method = 1
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_MC
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 3, 3]
    p.periodic[:] = [0, 1, 1]
    ng = mfmhd.num_ghost_cells(p)
    dx = 1.0 / n
    vlo, vhi = (0, 0, 0), (n - 1, 3, 3)
    lo, hi, shape = grid_shape(8, vlo, vhi, ng)
    U0 = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    ctx = mfmhd.Context(p)
    Ud = mfmhd.DeviceArray(shape, data=U0)
    fd = mfmhd.fab(Ud.ptr, lo, hi, 8)
    valid = mfmhd.Box.of(vlo, vhi)
    Uo, t, steps, dt = U0.copy(), 0.0, 0, None
    dtn, _ = oracle_mod.compute_dt(p, Uo, lo, vlo, vhi, dx)
    while t < t_end * (1 - 1e-12):
        dt = 0.5 * dtn if dt is None else min(0.5 * dtn, 1.1 * dt)
        dt = min(dt, t_end - t)
        periodic_fill(Uo, ng)
        g = gpu_update(ctx, Uo, lo, vlo, vhi, dt, dx, t)
        o = oracle_mod.patch_update(p, Uo, lo, vlo, vhi, dt, dx, t, nthreads=THREADS)
        check64(g, o, lo, vlo, vhi)
        ctx.fill_periodic(fd, valid)
        ctx.update_state(fd, valid, dt, dx, t)
        Uo, dtn = o["U"], o["dt"]
        t += dt
        steps += 1
    assert steps > 200
    a = valid_view(Ud.numpy(), lo, vlo, vhi)
    b = valid_view(Uo, lo, vlo, vhi)
    err = rel_err_per_comp(a, b)
    print("riemann14: %d steps, trajectory error %.2e" % (steps, err.max()), flush=True)
    assert err.max() <= 1e-10, err
    # the transverse directions stay uniform; the tube has evolved
    assert np.max(np.abs(b - b[:, :1, :1, :])) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(b[0, 0, 0] - U0[0, ng, ng, ng:-ng])) > 0.1
    ctx.close()
    Ud.free()


def test_orszag_tang_glm_512(oracle_mod):
    n = 512
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.spacedim = 2
    p.nfluids = 1
    p.dedner = 1
    p.eight_wave = 0
    p.flux = mfmhd.FLUX_ROE8
    p.time_approx = mfmhd.TIME_RK2
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_MINMOD
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, n - 1, 0]
    p.periodic[:] = [1, 1, 0]
    ng = mfmhd.num_ghost_cells(p)
    dx = 1.0 / n
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, 0)
    lo, hi, shape = grid_shape(9, vlo, vhi, ng, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    ctx = mfmhd.Context(p)
    t = 0.0
    for step in range(5):
        print("orszag 512 step", step, flush=True)
        ch = dx / dtn
        cp = math.sqrt(ch * dx)
        ctx.set_dedner(ch, cp)
        periodic_fill(U, ng, dims=2)
        dt = 0.5 * dtn
        g = gpu_update(ctx, U, lo, vlo, vhi, dt, dx, t)
        o = oracle_mod.patch_update(p, U, lo, vlo, vhi, dt, dx, t, ch=ch, cp=cp,
                                    nthreads=THREADS)
        check64(g, o, lo, vlo, vhi)
        U, dtn = o["U"], o["dt"]
        t += dt
    ctx.close()


def test_helio_256_windows(oracle_mod):
    """Two steps of the 256^3 box on the GPU, then one more: the windows of
    that last update against the oracle's update of the same windows."""
    b = _bench()
    n, ng, nc = 256, 4, 23
    p, dx = b.c3_params(n, mfmhd.PREC_FP64)
    ctx = mfmhd.Context(p)
    lo, hi = (-ng,) * 3, (n - 1 + ng,) * 3
    U = mfmhd.DeviceArray((nc, n + 2 * ng, n + 2 * ng, n + 2 * ng))
    f = mfmhd.fab(U.ptr, lo, hi, nc)
    ctx.init_problem(f, mfmhd.Box.of(lo, hi), dx)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1,) * 3)
    dtn, _ = ctx.compute_dt(f, valid, dx)
    t, dt = 0.0, None
    for _ in range(2):
        dt = 0.1 * dtn if dt is None else min(0.2 * dtn, 1.01 * dt)
        dtn, _ = ctx.update_state(f, valid, dt, dx, t)
        t += dt
    mfmhd.synchronize()
    S = U.numpy()
    dt = min(0.2 * dtn, 1.01 * dt)
    ctx.update_state(f, valid, dt, dx, t)
    R = U.numpy()
    ctx.close()
    U.free()
    regions = set()
    for name, c0 in (("sun", (108, 108, 108)), ("corner", (0, 0, 0)), ("bow", (20, 108, 108))):
        print("helio window", name, flush=True)
        vlo = c0
        vhi = tuple(v + 39 for v in c0)
        wlo = tuple(v - ng for v in vlo)
        sl = (slice(None),) + tuple(slice(vlo[d] - ng + ng, vhi[d] + ng + ng + 1)
                                    for d in (2, 1, 0))
        W = np.ascontiguousarray(S[sl])
        o = oracle_mod.patch_update(p, W, wlo, vlo, vhi, dt, dx, t, nthreads=THREADS)
        assert o["rc"] == 0
        a = R[(slice(None),) + tuple(slice(vlo[d] + ng, vhi[d] + ng + 1) for d in (2, 1, 0))]
        err = rel_err_per_comp(a, valid_view(o["U"], wlo, vlo, vhi))
        assert err.max() <= TOL64, (name, err)
        for k in range(ng, ng + 40, 3):
            for j in range(ng, ng + 40, 3):
                for i in range(ng, ng + 40, 3):
                    w = oracle_mod.cons_to_prim(p, W[:, k, j, i])
                    regions.add(oracle_mod.region(p, w, i + wlo[0], j + wlo[1], k + wlo[2], dx))
    assert {1, 2, 3} <= regions, regions


def test_nested_amr_3000_patches_graph(oracle_mod):
    b = _bench()
    lv, blobs = b.amr_hierarchy(mfmhd.PREC_FP64)
    assert sum(len(L["boxes"]) for L in lv) >= 2990
    st = b.amr_state(lv, blobs)
    ng, nc = 4, 23
    rng = np.random.default_rng(3)
    for li, (L, (ctx, buf, table, offs)) in enumerate(zip(lv, st)):
        print("amr level", li, "patches", len(table), flush=True)
        S = buf.numpy()
        d0 = min(ctx.compute_dt(f, bx, L["dx"])[0] for f, bx in table)
        dt = 0.1 * d0
        ld_e, mp_e, dts_e, cells_e = ctx.update_level(table, dt, L["dx"], 0.0)
        mfmhd.synchronize()
        Re = buf.numpy()
        g = mfmhd.Context(L["p"])
        g.set_graphs(True)
        for rep in range(2):   # capture, then replay
            buf.upload(S)
            assert ctx.compute_dt(*table[0], L["dx"]) == g.compute_dt(*table[0], L["dx"])
            res = g.update_level(table, dt, L["dx"], 0.0)
            mfmhd.synchronize()
            assert res[0] == ld_e and res[1] == mp_e and list(res[2]) == list(dts_e)
            assert np.array_equal(buf.numpy(), Re), (li, rep)
        g.close()
        for q in rng.choice(len(table), size=4, replace=False):
            vlo, vhi = L["boxes"][q]
            shp = (nc,) + tuple(vhi[d] - vlo[d] + 1 + 2 * ng for d in (2, 1, 0))
            sz = int(np.prod(shp))
            Uq = S[offs[q]:offs[q] + sz].reshape(shp)
            lo = tuple(v - ng for v in vlo)
            o = oracle_mod.patch_update(L["p"], Uq, lo, vlo, vhi, dt, L["dx"], 0.0,
                                        nthreads=THREADS)
            a = Re[offs[q]:offs[q] + sz].reshape(shp)
            err = rel_err_per_comp(valid_view(a, lo, vlo, vhi), valid_view(o["U"], lo, vlo, vhi))
            assert err.max() <= TOL64, (li, q, err)
            assert abs(dts_e[q] - o["dt"]) <= 1e-12 * o["dt"]
        ctx.close()
        buf.free()
        del S, Re
