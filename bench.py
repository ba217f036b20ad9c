"""Benchmark of the multi-fluid MHD patch update on one MI355X.

Workload (default): the 3-D multi-fluid heliospheric box, 256^3 cells on one
patch (+4 ghost cells), plasma + 3 neutral fluids (23 components), solar wind
from an inner sphere r = 0.1 into the LISM (26 km/s, 3 uG at 45 degrees),
charge exchange on, Roe 8-wave fluxes, minmod, RK2, 8-wave sources, dt
control of the reference deck (cfl 0.2, initial 0.1, growth 1.01).
Inputs live in HBM before the timed region; each step is one full RK2 patch
update (both stages, sources, floors and the CFL reduction) with the state
advanced in place; dt follows the CFL condition from the previous step.

Prints one JSON line (metric cell-updates/s, roofline and CPU baseline).
"""
import argparse
import math
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ms-flukss-multifluid_amd"))

import numpy as np  # noqa: E402

import mfmhd  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s

# dt control of the reference's 3-D heliospheric deck (main/inputs/
# helio03.inputs): cfl 0.2, initial_cfl 0.1, max_dt_growth 1.01
CFL, CFL0, DT_GROWTH = 0.2, 0.1, 1.01

# Solar wind of the heliospheric workload: 7 /cc at the inner sphere r = 0.1
# (AU).  The setter takes the density at 1 AU and scales it by r^-2.  (With
# 7 /cc at 1 AU, i.e. 700 /cc at the sphere, the explicit charge-exchange
# source of the third neutral population is unstable at this dt and the
# state overflows within a few steps.)
SUN_N_1AU = 0.07


def finite_dt(dtn, where):
    """The next dt is the minimum over all cells and propagates NaN: a
    non-finite value means the state itself has gone non-finite."""
    if not math.isfinite(dtn) or dtn <= 0.0:
        raise RuntimeError("%s: state went non-finite (next dt %r)" % (where, dtn))
    return dtn

def note(msg):
    """One progress line on stderr (the JSON result stays alone on stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


# HBM bytes per launch of the dominant kernel (average over the two RK2
# stages), measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE on the
# default workload and scaled by what the counters read for the library's
# own copy kernel over 1 GiB (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0); summary
# in profiles/k_update_helio256_traffic.md.  Valid for that workload and
# stage structure only.
TRAFFIC = {(256, "fp32flux", "separate", "k_update"): 1.4068e10,
           (256, "fp32flux", "summed", "k_update"): 1.0368e10}


def c3_params(n, prec):
    dx = 1.0 / n
    p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, lismT=6527.0, lismB=3e-6,
                           lismBdir=(2 ** -0.5, 2 ** -0.5, 0.0),
                           lismUdir=(-1.0, 0.0, 0.0), sunN=SUN_N_1AU, sunV=4.5e7,
                           sunXYZ=(0.5, 0.5, 0.5), R0=0.1, initR=0.3,
                           netN=0.1, dx=dx, num_cells=(n, n, n))
    p.nfluids = 4
    p.time_approx = mfmhd.TIME_RK2
    p.flux = mfmhd.FLUX_ROE8
    p.precision = prec
    return p, dx


def host_cpu():
    """nproc, the CPU model and the SMT state of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    smt = "unknown"
    try:
        with open("/sys/devices/system/cpu/smt/active") as fh:
            smt = "on" if fh.read().strip() == "1" else "off"
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model, "smt": smt}


def _oracle_rate(oracle, p, dx, n_sample, threads, warmup=3, timed=10):
    """Median over `timed` steps (after `warmup`) of the oracle's RK2 update
    of an n_sample^3 patch of the 256^3 heliospheric state around the sun."""
    n, ng = 256, 4
    c0 = n // 2 - n_sample // 2
    vlo = (c0, c0, c0)
    vhi = (c0 + n_sample - 1,) * 3
    lo = tuple(v - ng for v in vlo)
    hi = tuple(v + ng for v in vhi)
    shape = (23, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    U = oracle.init_problem(p, shape, lo, lo, hi, dx)
    dt0, _ = oracle.compute_dt(p, U, lo, vlo, vhi, dx)
    times = []
    for s in range(warmup + timed):
        t0 = time.perf_counter()
        r = oracle.patch_update(p, U, lo, vlo, vhi, 0.5 * dt0, dx, nthreads=threads)
        if s >= warmup:
            times.append(time.perf_counter() - t0)
        assert r["rc"] == 0
    return n_sample ** 3 / float(np.median(times)), len(times)


def cpu_baseline(n_sample, threads, n_single=32):
    """The CPU oracle (the FP64 C++ restatement of the reference's path,
    -O3 -march=x86-64-v3 -ffp-contract=off -fno-fast-math, OpenMP over z
    planes) on bounded samples of the same workload: an n_sample^3 patch on
    `threads` cores and an n_single^3 patch on 1 core; 3 warm-up steps, then
    the median of 10."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    p, dx = c3_params(256, mfmhd.PREC_FP64)
    v_all, k = _oracle_rate(oracle, p, dx, n_sample, threads)
    v_one, _ = _oracle_rate(oracle, p, dx, n_single, 1)
    return {"value": v_all, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": "oracle FP64 RK2 update of a %d^3 patch of the 256^3 4-fluid "
                      "heliospheric state (+4 ghosts) around the sun, 3 warm-up steps, "
                      "median of %d, OpenMP %d threads" % (n_sample, k, threads),
            "single_core": {"value": v_one, "cores": 1,
                            "sample": "%d^3 patch, same state, 3 warm-up, median of 10"
                                      % n_single},
            "host": host_cpu()}


def serve_bench(prec, n_req=512, steps=20, rate_hz=4096.0, seed=7, stage=None):
    """Ensemble serving of the 2-D axisymmetric SW/LISM problem (the
    reference's swlism2F deck: 1500 AU domain, sun at 600 AU on the axis,
    plasma + 1 neutral = 13 components, first order in time, cfl 0.25,
    initial_cfl 0.1, max_dt_growth 1.02): n_req independent requests with
    grids nx x ny, nx and ny ~ U[64, 512], solar-wind speed ~ U[350, 800] km/s and
    LISM field ~ U[2, 5] uG, arriving as a Poisson process of rate rate_hz,
    continuously batched by mfm_server_* under a 64 GiB HBM budget.  Fluxes:
    8-wave Roe on Roe-averaged states (the deck's arithmetic-mean Roe loses
    positivity at the wind / LISM interface of SWLISMINIT's state within a
    few steps).  Reports the p50/p99 per-request step latency (one server
    iteration advances every active request one step) beside the
    throughput.  stage: the server's stage structure (default: its own)."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(n_req):
        nx, ny = (int(v) for v in rng.integers(64, 513, size=2))
        sunV = float(rng.uniform(3.5e7, 8.0e7))
        lismB = float(rng.uniform(2e-6, 5e-6))
        dx = 1500.0 / nx
        p = mfmhd.swlism_params(sunV=sunV, lismB=lismB, sunXY=(600.0, 0.0),
                                R0=max(10.0, 2.6 * dx), initR=max(80.0, 8.0 * dx),
                                dx=dx, num_cells=(nx, ny), nfluids=2)
        p.time_approx = mfmhd.TIME_FIRST_ORDER
        p.flux = mfmhd.FLUX_ROE8_ROEAVG
        p.precision = prec
        reqs.append((p, dx, nx * ny))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, size=n_req))
    srv = mfmhd.Server(reqs[0][0], 64.0 * (1 << 30))
    if stage is not None:
        srv.set_stage_kernel(stage)
    t0 = time.perf_counter()
    nxt, ids, meta = 0, [], {}
    cells = 0.0
    while True:
        now = time.perf_counter() - t0
        while nxt < n_req and arrivals[nxt] <= now:
            p, dx, ncell = reqs[nxt]
            ids.append(srv.submit(p, dx, steps, 0.25, 0.1, 1.02))
            meta[ids[-1]] = (ncell, p.sun_v, p.lism_bx)
            cells += float(ncell) * steps
            nxt += 1
        na, npd, _ = srv.step()
        for r in list(ids):
            st = srv.info(r)["state"]
            if st == mfmhd.REQ_FAILED:
                raise RuntimeError("served request %s failed: %s" % (meta[r], srv.error(r)))
            if st == mfmhd.REQ_DONE:
                srv.release(r)
                ids.remove(r)
        if nxt == n_req and na == 0 and npd == 0 and not ids:
            break
        if na == 0 and npd == 0 and nxt < n_req:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return {"requests": n_req, "steps_per_request": steps,
            "workload": "SW/LISM 2-D axisymmetric, nx x ny, nx, ny ~ U[64, 512], sunV ~ "
                        "U[350, 800] km/s, lismB ~ U[2, 5] uG, 2 fluids (13 comps), "
                        "first order, Roe8 on Roe averages, minmod",
            "arrival": "Poisson %.0f /s" % rate_hz,
            "stage_kernels": {None: "server default (separate)", mfmhd.STAGE_SEPARATE: "separate",
                              mfmhd.STAGE_SUMMED: "summed"}.get(stage, str(stage)),
            "p50_step_ms": st["p50_ms"], "p99_step_ms": st["p99_ms"],
            "iterations": st["iterations"],
            "cell_updates": cells,
            "cell_updates_per_s": st["cell_updates"] / (st["busy_ms"] * 1e-3),
            "wall_s": wall, "wall_cell_updates_per_s": cells / wall}


def amr_hierarchy(prec, n_patches=3000, levels=3, seed=11, n_blobs=12):
    """The batched-AMR workload: `levels` nested levels (ratio 2) of the
    heliospheric problem, about n_patches patches in all (n_patches/levels
    per level) with edges drawn uniformly from {8, 16, 32}, each in its own
    32^3 slot of a cube of m^3 slots; level L+1 refines the central half of
    level L (domain 32 m 2^L cells per side), so the three levels cover
    nested regions around the sun.  Seeded Gaussian blobs of density and
    pressure (n_blobs, amplitude U[0.2, 1], width U[0.02, 0.08] of the
    domain) sit on the HELIOINIT background.  Returns (levels, blobs)."""
    rng = np.random.default_rng(seed)
    per_level = n_patches // levels
    m = int(np.ceil(per_level ** (1.0 / 3.0)))
    side = 32 * m
    blobs = [(tuple(rng.uniform(0.3, 0.7, size=3)), float(rng.uniform(0.2, 1.0)),
              float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.02, 0.08)))
             for _ in range(n_blobs)]
    lv = []
    for L in range(levels):
        n_dom = side * 2 ** L
        off = (n_dom - side) // 2
        dx = 1.0 / n_dom
        p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, lismT=6527.0, lismB=3e-6,
                               lismBdir=(2 ** -0.5, 2 ** -0.5, 0.0),
                               lismUdir=(-1.0, 0.0, 0.0), sunN=SUN_N_1AU, sunV=4.5e7,
                               sunXYZ=(0.5, 0.5, 0.5), R0=0.1, initR=0.3,
                               netN=0.1, dx=dx, num_cells=(n_dom,) * 3)
        p.nfluids = 4
        p.time_approx = mfmhd.TIME_RK2
        p.flux = mfmhd.FLUX_ROE8
        p.precision = prec
        slots = rng.permutation(m ** 3)[:per_level]
        edges = rng.choice([8, 16, 32], size=per_level)
        boxes = []
        for q in range(per_level):
            sx, sy, sz = slots[q] % m, (slots[q] // m) % m, slots[q] // (m * m)
            e = int(edges[q])
            vlo = (off + 32 * sx, off + 32 * sy, off + 32 * sz)
            boxes.append((vlo, tuple(v + e - 1 for v in vlo)))
        lv.append({"p": p, "dx": dx, "boxes": boxes, "n_dom": n_dom})
    return lv, blobs


def apply_blobs(U, lo, dx, blobs, gamma):
    """Multiply the plasma density and pressure of the conserved FAB U
    (nc, nz, ny, nx; low corner lo) by 1 + sum of Gaussian blobs, keeping
    the velocity and the field."""
    nc, nz, ny, nx = U.shape
    z, y, x = np.meshgrid((np.arange(nz) + lo[2] + 0.5) * dx, (np.arange(ny) + lo[1] + 0.5) * dx,
                          (np.arange(nx) + lo[0] + 0.5) * dx, indexing="ij")
    fr = np.ones_like(x)
    fp = np.ones_like(x)
    for (cx, cy, cz), ar, ap, w in blobs:
        g = np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (w * w))
        fr += ar * g
        fp += ap * g
    rho = U[0]
    ke = 0.5 * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2) / rho
    me = (U[5] ** 2 + U[6] ** 2 + U[7] ** 2) / (8.0 * math.pi)
    pr = (gamma - 1.0) * (U[4] - ke - me)
    U[0] = rho * fr
    U[1:4] *= fr
    U[4] = pr * fp / (gamma - 1.0) + ke * fr + me


def amr_state(lv, blobs, ng=4, nc=23):
    """Device buffers of the hierarchy: per level one DeviceArray holding
    every patch's FAB (valid + ng ghosts) back to back, HELIOINIT on the
    device, then the blobs.  Returns, per level, (ctx, buf, table, offsets)."""
    out = []
    for L in lv:
        offs, off = [], 0
        for vlo, vhi in L["boxes"]:
            offs.append(off)
            off += nc * int(np.prod([h - l + 1 + 2 * ng for l, h in zip(vlo, vhi)]))
        buf = mfmhd.DeviceArray((off,))
        ctx = mfmhd.Context(L["p"])
        table = []
        for (vlo, vhi), o in zip(L["boxes"], offs):
            lo = tuple(v - ng for v in vlo)
            hi = tuple(v + ng for v in vhi)
            f = mfmhd.fab(buf.ptr + 8 * o, lo, hi, nc)
            ctx.init_problem(f, mfmhd.Box.of(lo, hi), L["dx"])
            table.append((f, mfmhd.Box.of(vlo, vhi)))
        mfmhd.synchronize()
        host = buf.numpy()
        for (vlo, vhi), o in zip(L["boxes"], offs):
            shp = (nc,) + tuple(h - l + 1 + 2 * ng for l, h in zip(vlo[::-1], vhi[::-1]))
            U = host[o:o + int(np.prod(shp))].reshape(shp)
            apply_blobs(U, tuple(v - ng for v in vlo), L["dx"], blobs, L["p"].gamma)
        buf.upload(host)
        del host
        out.append((ctx, buf, table, offs))
    mfmhd.synchronize()
    return out


def amr_bench(prec, n_patches=3000, levels=3, steps=5, warmup=1, seed=11):
    """The batched-AMR workload (amr_hierarchy: 3 nested levels, about 3000
    patches of edge 8/16/32, Gaussian blobs on HELIOINIT), 4 ghost cells,
    4 fluids.  Each level is one mfm_level_update_dev call (every kernel
    launched once for all its patches) replayed from a captured hipGraph,
    with the work walk autotuned on the level's largest patch; a step
    updates every level once.  Returns cell-updates/s over all levels."""
    lv, blobs = amr_hierarchy(prec, n_patches, levels, seed)
    st = amr_state(lv, blobs)
    total_cells = sum(int(np.prod([h - l + 1 for l, h in zip(vlo, vhi)]))
                      for L in lv for vlo, vhi in L["boxes"])
    state = []
    for L, (ctx, buf, table, offs) in zip(lv, st):
        ctx.set_graphs(True)
        big = max(range(len(table)), key=lambda q: table[q][1].shape()[0])
        f, b = table[big]
        d0, _ = ctx.compute_dt(f, b, L["dx"])
        ctx.autotune(f, b, 0.1 * d0, L["dx"])
        state.append({"ctx": ctx, "buf": buf, "table": table, "dx": L["dx"], "dt": None,
                      "dtn": None, "t": 0.0})

    def step():
        for L in state:
            if L["dtn"] is None:
                L["dtn"] = min(L["ctx"].compute_dt(f, b, L["dx"])[0] for f, b in L["table"])
            dt = CFL0 * L["dtn"] if L["dt"] is None else min(CFL * L["dtn"], DT_GROWTH * L["dt"])
            ld, _, _, _ = L["ctx"].update_level(L["table"], dt, L["dx"], L["t"])
            ld = finite_dt(ld, "batched AMR level")
            L["dt"], L["dtn"] = dt, ld
            L["t"] += dt

    for _ in range(warmup):
        step()
    mfmhd.synchronize()
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    mfmhd.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    walks = [L["ctx"].walk() for L in state]
    for L in state:
        L["ctx"].close()
        L["buf"].free()
    return {"patches": sum(len(L["boxes"]) for L in lv), "levels": levels,
            "layout": "nested, ratio 2, level L+1 refines the central half of level L",
            "edges": "8/16/32 uniform (seed %d)" % seed, "blobs": len(blobs),
            "cells": int(total_cells), "graphs": True, "walks": walks,
            "ms_per_step": ms, "cell_updates_per_s": total_cells / (ms * 1e-3)}


def fill_bench(box=16, per_side=10, levels=3, reps=5):
    """Ghost fill of a nested 3-level hierarchy (ratio 2): each level is a
    cube of per_side^3 boxes of box^3 cells (1000 boxes of 16^3 by default,
    3000 patches in all), level L+1 covering the central half of level L's
    domain, 4 ghost cells, 4 fluids on the HELIOINIT state.  One fill is, per
    level, fillInterp from the coarser level (time coefficient 0.5, old and
    new coarse data) and the exchange -- the ghost fill that precedes the
    updates of a step (srcIdeal/AMRLevelIdealMHD.cpp:701-772).  Also times
    the coarse-fine synchronisation that follows them (flux register, reflux,
    average down; :1282-1307, 1636-1647)."""
    ng, nc, r = 4, 23, 2
    side = box * per_side
    lv = []
    for L in range(levels):
        n_dom = side * r ** L
        off0 = 0 if L == 0 else (n_dom - side) // 2
        dx = 1.0 / n_dom
        p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, lismT=6527.0, lismB=3e-6,
                               lismBdir=(2 ** -0.5, 2 ** -0.5, 0.0),
                               lismUdir=(-1.0, 0.0, 0.0), sunN=SUN_N_1AU, sunV=4.5e7,
                               sunXYZ=(0.5, 0.5, 0.5), R0=0.1, initR=0.3,
                               netN=0.1, dx=dx, num_cells=(n_dom,) * 3)
        p.nfluids = 4
        p.time_approx = mfmhd.TIME_RK2
        boxes = []
        for c in range(per_side):
            for b in range(per_side):
                for a in range(per_side):
                    lo = (off0 + a * box, off0 + b * box, off0 + c * box)
                    boxes.append((lo, tuple(v + box - 1 for v in lo)))
        fcells = (box + 2 * ng) ** 3
        buf = mfmhd.DeviceArray((len(boxes) * nc * fcells,))
        old = mfmhd.DeviceArray((len(boxes) * nc * fcells,))
        ctx = mfmhd.Context(p)
        fabs, ofabs = [], []
        for q, (lo, hi) in enumerate(boxes):
            glo = tuple(v - ng for v in lo)
            ghi = tuple(v + ng for v in hi)
            f = mfmhd.fab(buf.ptr + 8 * q * nc * fcells, glo, ghi, nc)
            ctx.init_problem(f, mfmhd.Box.of(glo, ghi), dx)
            fabs.append(f)
            ofabs.append(mfmhd.fab(old.ptr + 8 * q * nc * fcells, glo, ghi, nc))
        old.copy_from(buf)
        # the FAB tables of a level are built once, as a driver keeps its
        # LevelData between steps
        fabs, ofabs = mfmhd.fab_array(fabs), mfmhd.fab_array(ofabs)
        lv.append({"p": p, "ctx": ctx, "boxes": boxes, "fabs": fabs, "ofabs": ofabs,
                   "buf": buf, "old": old, "dx": dx, "n_dom": n_dom})
    mfmhd.synchronize()
    t0 = time.perf_counter()
    for L, d in enumerate(lv):
        if L == 0:
            d["plan"] = mfmhd.FillPlan(d["p"], d["boxes"], interp_radius=ng)
        else:
            c = lv[L - 1]
            d["plan"] = mfmhd.FillPlan(d["p"], d["boxes"], c["boxes"], r, ng, c["dx"])
    plan_s = (time.perf_counter() - t0) / levels

    def fill(which):
        for L, d in enumerate(lv):
            if L > 0 and which in ("all", "interp"):
                c = lv[L - 1]
                d["ctx"].fill_interp(d["plan"], d["fabs"], c["ofabs"], c["fabs"], 0.5)
            if which in ("all", "exchange"):
                d["ctx"].exchange(d["plan"], d["fabs"])

    res = {}
    for which in ("exchange", "interp", "all"):
        fill(which)
        mfmhd.synchronize()
        e0, e1 = mfmhd.Event(), mfmhd.Event()
        e0.record()
        for _ in range(reps):
            fill(which)
        e1.record()
        mfmhd.synchronize()
        res[which] = e0.elapsed_ms(e1) / reps
    # coarse-fine synchronisation of each level pair: register increments
    # from the coarse fluxes and two fine steps, reflux, average down
    fcells_face = (box + 1) * box * box
    for d in lv:
        d["F"] = mfmhd.DeviceArray((len(d["boxes"]) * 3 * nc * fcells_face,), zero=True)
        d["Ff"] = []
        for q, (lo, hi) in enumerate(d["boxes"]):
            for dd in range(3):
                fhi = list(hi)
                fhi[dd] += 1
                d["Ff"].append(mfmhd.fab(d["F"].ptr + 8 * (3 * q + dd) * nc * fcells_face,
                                         lo, fhi, nc))
    regs = [mfmhd.FluxRegister(lv[L]["p"], lv[L]["boxes"], lv[L - 1]["boxes"], r)
            for L in range(1, levels)]
    for d in lv:
        d["Ff"] = mfmhd.fab_array(d["Ff"])

    def sync():
        for L in range(1, levels):
            c, f, rg = lv[L - 1], lv[L], regs[L - 1]
            rg.set_zero(c["ctx"])
            rg.increment_coarse(c["ctx"], c["Ff"], 1e-3)
            for _ in range(r):
                rg.increment_fine(f["ctx"], f["Ff"], 5e-4)
            rg.reflux(c["ctx"], c["fabs"], -1.0, c["dx"])
            rg.average_down(c["ctx"], c["fabs"], f["fabs"], f["dx"])

    sync()
    mfmhd.synchronize()
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    for _ in range(reps):
        sync()
    e1.record()
    mfmhd.synchronize()
    res["sync"] = e0.elapsed_ms(e1) / reps
    rinfo = [rg.info() for rg in regs]
    for rg in regs:
        rg.close()
    info = [d["plan"].info() for d in lv]
    xc = sum(i["exchange_cells"] for i in info)
    fcl = sum(i["interp_fine_cells"] for i in info)
    ccl = sum(i["interp_coarse_cells"] for i in info)
    # algorithmic bytes: exchange reads and writes each ghost cell once; the
    # interpolation reads the coarse cells it gathers twice (old, new) and
    # writes each fine cell once (the gathered ring is not counted)
    xb = 2 * 8 * nc * xc
    ib = 2 * 8 * nc * ccl + 8 * nc * fcl
    for d in lv:
        d["plan"].close()
        d["ctx"].close()
        d["buf"].free()
        d["old"].free()
        d["F"].free()
    return {"levels": levels, "patches": per_side ** 3 * levels, "box": box,
            "ghost": ng, "components": nc, "plan_build_s_per_level": plan_s,
            "ms_per_fill": res["all"],
            "exchange": {"ms": res["exchange"], "ghost_cells": xc,
                         "GB_per_s": xb / (res["exchange"] * 1e-3) / 1e9},
            "interp": {"ms": res["interp"], "fine_cells": fcl, "coarse_cells": ccl,
                       "GB_per_s": ib / (res["interp"] * 1e-3) / 1e9},
            "sync": {"ms": res["sync"],
                     "register_faces": sum(i["faces"] for i in rinfo),
                     "average_cells": sum(i["average_cells"] for i in rinfo),
                     "what": "per level pair: register zero, coarse increment, 2 fine "
                             "increments, reflux, average down"}}


def hancock_bench(prec, n=256, steps=10, warmup=3, walk=None):
    """The same 256^3 4-fluid heliospheric workload with the Hancock
    predictor-corrector (time_approximation = 1, the scheme of the
    reference's 3-D multi-fluid deck main/inputs/helio03.inputs): one Riemann
    solve per step instead of two.  Same dt control."""
    p, dx = c3_params(n, prec)
    p.time_approx = mfmhd.TIME_HANCOCK
    ctx = mfmhd.Context(p)
    if walk:
        ctx.set_walk(walk[0], walk[1])
    nc, ng = ctx.ncomp, ctx.nghost
    lo, hi = (-ng,) * 3, (n - 1 + ng,) * 3
    m = n + 2 * ng
    U = mfmhd.DeviceArray((nc, m, m, m))
    Uf = mfmhd.fab(U.ptr, lo, hi, nc)
    ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1,) * 3)
    dtn, _ = ctx.compute_dt(Uf, valid, dx)
    t, dt = 0.0, None
    for _ in range(warmup):
        dt = CFL0 * dtn if dt is None else min(CFL * dtn, DT_GROWTH * dt)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "update")
        t += dt
    mfmhd.synchronize()
    ctx.set_timing(True)
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    for _ in range(steps):
        dt = min(CFL * dtn, DT_GROWTH * dt)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "update")
        t += dt
    e1.record()
    mfmhd.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    kern = {}
    for st, stage in ((0, "predictor"), (1, "corrector")):
        for kind, name in enumerate(mfmhd.KT_NAMES):
            kms, kn, kit = ctx.timing(st, kind)
            if kn:
                kern["%s/%s" % (name, stage)] = {"avg_ms": kms / kn, "launches": kn}
    ctx.close()
    U.free()
    return {"scheme": "Hancock predictor-corrector (time_approximation = 1)",
            "ghost": ng, "ms_per_step": ms, "cell_updates_per_s": float(n) ** 3 / (ms * 1e-3),
            "kernels": kern}


def brio_wu_bench(n=800, t_end=0.1, cfl=0.8):
    """The 1-D Brio-Wu shock tube at its stated size: n cells on [0, 1],
    gamma = 2, Roe 8-wave on arithmetic means, minmod, 8-wave sources, RK2,
    CFL 0.8, fixed-state boundaries, FP64, run to t = 0.1 (a 2-D patch one
    cell high, periodic in y).  Too small to fill the GPU: reported for
    latency (launch-bound), not for the roofline."""
    S4PI = math.sqrt(4.0 * math.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, S4PI, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -S4PI, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    p.spacedim = 2
    p.time_approx = mfmhd.TIME_RK2
    p.flux = mfmhd.FLUX_ROE8
    p.precision = mfmhd.PREC_FP64
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_MINMOD
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, 0, 0]
    p.periodic[:] = [0, 1, 0]
    dx = 1.0 / n
    ctx = mfmhd.Context(p)
    ng = ctx.nghost
    lo, hi = (-ng, -ng, 0), (n - 1 + ng, ng, 0)
    U = mfmhd.DeviceArray((8, 1, 1 + 2 * ng, n + 2 * ng))
    Uf = mfmhd.fab(U.ptr, lo, hi, 8)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1, 0, 0))
    ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
    ctx.fill_periodic(Uf, valid)
    dtn, _ = ctx.compute_dt(Uf, valid, dx)
    t, steps = 0.0, 0
    mfmhd.synchronize()
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    while t < t_end * (1 - 1e-12):
        dt = min(cfl * dtn, t_end - t)
        ctx.fill_periodic(Uf, valid)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "Brio-Wu")
        t += dt
        steps += 1
    e1.record()
    mfmhd.synchronize()
    ms = e0.elapsed_ms(e1)
    ctx.close()
    U.free()
    return {"workload": "Brio-Wu 1-D, %d cells, Roe8, minmod, RK2, CFL %.1f, to t=%.2g, FP64"
                        % (n, cfl, t_end),
            "steps": steps, "t": t, "total_ms": ms, "ms_per_step": ms / steps,
            "cell_updates_per_s": n * steps / (ms * 1e-3)}


def orszag_glm_bench(n=512, steps=1000, cfl=0.5, flux=mfmhd.FLUX_ROE8):
    """The 2-D Orszag-Tang vortex at its stated size: n^2 periodic on
    [0, 1]^2, gamma = 5/3, GLM (Dedner-Roe: flux 1 with divB_cleaning 5,
    8-wave off, Ch = dx / dt_raw and Cp = sqrt(Ch dx) each step as the level
    driver sets them, dt_raw = dx / c_max), minmod, RK2, FP64, CFL 0.5 (the
    reference's Orszag deck; 0.8 is unstable for this unsplit 2-D update),
    `steps` steps with the periodic ghost fill in each.  flux = FLUX_HLLC
    runs the HLLC flux with the same GLM block instead of the 7-wave Roe."""
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.spacedim = 2
    p.nfluids = 1
    p.dedner = 1
    p.eight_wave = 0
    p.flux = flux
    p.time_approx = mfmhd.TIME_RK2
    p.precision = mfmhd.PREC_FP64
    for v in range(8):
        p.limiter[v] = mfmhd.LIM_MINMOD
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [n - 1, n - 1, 0]
    p.periodic[:] = [1, 1, 0]
    dx = 1.0 / n
    ctx = mfmhd.Context(p)
    ng = ctx.nghost
    nc = ctx.ncomp
    lo, hi = (-ng, -ng, 0), (n - 1 + ng, n - 1 + ng, 0)
    U = mfmhd.DeviceArray((nc, 1, n + 2 * ng, n + 2 * ng))
    Uf = mfmhd.fab(U.ptr, lo, hi, nc)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1, n - 1, 0))
    ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
    ctx.fill_periodic(Uf, valid)
    dtn, _ = ctx.compute_dt(Uf, valid, dx)
    t = 0.0

    def step():
        nonlocal t, dtn
        dt = cfl * dtn
        ch = dx / dtn
        ctx.set_dedner(ch, math.sqrt(ch * dx))
        ctx.fill_periodic(Uf, valid)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "Orszag-Tang")
        t += dt

    for _ in range(5):
        step()
    mfmhd.synchronize()
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    for _ in range(steps):
        step()
    e1.record()
    mfmhd.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    ctx.close()
    U.free()
    rate = n * n / (ms * 1e-3)
    solver = "HLLC" if flux == mfmhd.FLUX_HLLC else "Dedner-Roe"
    return {"workload": "Orszag-Tang 2-D %d^2 periodic, GLM (%s), minmod, RK2, "
                        "CFL %.1f, FP64" % (n, solver, cfl),
            "steps": steps, "t_end": t, "ms_per_step": ms, "cell_updates_per_s": rate,
            "roofline_frac": rate * 5 * nc * 8 / HBM_PEAK,
            "alg_bytes_per_cell_update": 5 * nc * 8}


def orszag_bench(nxy=512, nz=32, steps=10, warmup=3):
    """The scheme of the reference's Orszag-Tang deck (main/inputs/
    orszag.inputs: Hancock, characteristic slopes with check_tvd = 1, van
    Leer on every variable, Roe 8-wave, cfl 0.5, max_dt_growth 1.1) on one
    periodic nxy x nxy x nz box in FP64 -- the deck's finest level (64 x 64
    x 4 refined three times by 2) as one patch, with the ghost exchange in
    every step -- and the same box with component-wise slopes."""
    out = {"box": [nxy, nxy, nz], "precision": "fp64"}
    for label, char in (("characteristic", 1), ("componentwise", 0)):
        p = mfmhd.orszag_params(1.66666666667)
        p.spacedim = 3
        p.nfluids = 1
        p.time_approx = mfmhd.TIME_HANCOCK
        p.char_reconstruction = char
        p.check_tvd = 1
        p.flux = mfmhd.FLUX_ROE8
        p.eight_wave = 1
        p.divb_method = 1
        p.precision = mfmhd.PREC_FP64
        for v in range(8):
            p.limiter[v] = mfmhd.LIM_VANLEER
        hi_v = (nxy - 1, nxy - 1, nz - 1)
        for d in range(3):
            p.domain_lo[d] = 0
            p.domain_hi[d] = hi_v[d]
            p.periodic[d] = 1
        dx = 1.0 / nxy
        ctx = mfmhd.Context(p)
        ng = ctx.nghost
        lo = (-ng,) * 3
        hi = tuple(h + ng for h in hi_v)
        U = mfmhd.DeviceArray((8, nz + 2 * ng, nxy + 2 * ng, nxy + 2 * ng))
        Uf = mfmhd.fab(U.ptr, lo, hi, 8)
        ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
        plan = mfmhd.FillPlan(p, [((0, 0, 0), hi_v)], interp_radius=ng)
        valid = mfmhd.Box.of((0, 0, 0), hi_v)
        ctx.exchange(plan, [Uf])
        dtn, _ = ctx.compute_dt(Uf, valid, dx)
        t, dt = 0.0, None

        def step():
            nonlocal t, dt, dtn
            dt = 0.5 * dtn if dt is None else min(0.5 * dtn, 1.1 * dt)
            ctx.exchange(plan, [Uf])
            dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "update")
            t += dt

        for _ in range(warmup):
            step()
        mfmhd.synchronize()
        e0, e1 = mfmhd.Event(), mfmhd.Event()
        e0.record()
        for _ in range(steps):
            step()
        e1.record()
        mfmhd.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        ctx.set_timing(True)   # per-kernel times from a separate pass
        for _ in range(3):
            step()
        mfmhd.synchronize()
        kern = {}
        for st, stage in ((0, "predictor"), (1, "corrector")):
            for kind, name in enumerate(mfmhd.KT_NAMES):
                kms, kn, kit = ctx.timing(st, kind)
                if kn:
                    kern["%s/%s" % (name, stage)] = {"avg_ms": kms / kn, "launches": kn}
        out[label] = {"ms_per_step": ms,
                      "cell_updates_per_s": float(nxy * nxy * nz) / (ms * 1e-3),
                      "kernels": kern}
        plan.close()
        ctx.close()
        U.free()
    return out


def l1_per_variable(a, b, nf, own=False):
    """Relative L1 error of a against b per variable: ||a_c - b_c||_1 over
    the sum of ||b_q||_1 of the vector the component belongs to (momentum,
    field; the group norm), or, with own, over the component's own
    ||b_c||_1 (which a component near zero everywhere, like B_z of a field
    in the x-y plane, divides by nearly nothing)."""
    groups = [[0], [1, 2, 3], [1, 2, 3], [1, 2, 3], [4], [5, 6, 7], [5, 6, 7], [5, 6, 7]]
    for f in range(1, nf):
        s0 = 8 + 5 * (f - 1)
        g = [s0 + 1, s0 + 2, s0 + 3]
        groups += [[s0], g, g, g, [s0 + 4]]
    if own:
        groups = [[c] for c in range(len(groups))]
    norms = {}
    out = []
    for c, g in enumerate(groups):
        key = tuple(g)
        if key not in norms:
            norms[key] = sum(float(np.abs(b[q]).sum()) for q in g)
        out.append(float(np.abs(a[c] - b[c]).sum()) / max(norms[key], 1e-300))
    return out


def fp64_bench(n, dts, warmup, walk, final32):
    """The same workload in the FP64 build with the FP32 run's dt sequence:
    time per step over the steps after the warm-up, and the FP32-flux
    run's relative L1 error per variable against it at the end (the
    heliospheric configuration's FP64 run beside the FP32 flux run)."""
    p, dx = c3_params(n, mfmhd.PREC_FP64)
    ctx = mfmhd.Context(p)
    ctx.set_walk(walk[0], walk[1])
    nc, ng = ctx.ncomp, ctx.nghost
    lo, hi = (-ng,) * 3, (n - 1 + ng,) * 3
    m = n + 2 * ng
    U = mfmhd.DeviceArray((nc, m, m, m))
    Uf = mfmhd.fab(U.ptr, lo, hi, nc)
    ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1,) * 3)
    t = 0.0
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    for q, dt in enumerate(dts):
        if q == warmup:
            mfmhd.synchronize()
            e0.record()
        finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "fp64 update")
        t += dt
    e1.record()
    mfmhd.synchronize()
    steps = len(dts) - warmup
    ms = e0.elapsed_ms(e1) / steps
    final64 = U.numpy()[:, ng:ng + n, ng:ng + n, ng:ng + n]
    ctx.close()
    U.free()
    err = l1_per_variable(final32, final64, p.nfluids)
    own = l1_per_variable(final32, final64, p.nfluids, own=True)
    return {"ms_per_step": ms, "cell_updates_per_s": float(n) ** 3 / (ms * 1e-3),
            "steps_compared": len(dts),
            "fp32_flux_l1_per_variable": err, "fp32_flux_l1_max": max(err),
            "fp32_flux_l1_norm": "group: ||a_c - b_c||_1 / sum of ||b_q||_1 over the "
                                 "component's vector (momentum, B); own: / ||b_c||_1",
            "fp32_flux_l1_per_variable_own_norm": own,
            "fp32_flux_l1_max_own_norm": max(own)}


def host_path_bench(ctx, U, lo, valid, dt, dx, t, reps=3):
    """mfm_patch_update_host on a host copy of the state (pageable memory):
    each call stages U host -> HBM -> host around one update, so its rate
    includes PCIe both ways.  Wall time per call, best of `reps`."""
    Uh = U.numpy()
    walls = []
    for _ in range(reps):
        w0 = time.perf_counter()
        finite_dt(ctx.update_state_host(Uh, lo, valid, dt, dx, t)[0], "host update")
        walls.append(time.perf_counter() - w0)
    ms = 1e3 * min(walls)
    cells = float(np.prod(valid.shape()))
    return {"ms_per_step": ms, "cell_updates_per_s": cells / (ms * 1e-3),
            "staged_bytes_per_step": 2 * Uh.nbytes, "calls": reps,
            "memory": "pageable host FAB"}


def copy_ceiling(gib=4, reps=5):
    """Device-to-device copy bandwidth (read + write bytes per second): the
    HBM rate a streaming kernel can reach on this device."""
    nb = gib << 30
    a = mfmhd.DeviceArray((nb // 8,), zero=True)
    b = mfmhd.DeviceArray((nb // 8,))
    b.copy_from(a)
    mfmhd.synchronize()
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    for _ in range(reps):
        b.copy_from(a)
    e1.record()
    mfmhd.synchronize()
    ms = e0.elapsed_ms(e1) / reps
    a.free()
    b.free()
    return 2.0 * nb / (ms * 1e-3)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--precision", choices=["fp32flux", "fp64"], default="fp32flux")
    ap.add_argument("--cfl", type=float, default=CFL)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the ensemble-serving measurement")
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default work walk instead of autotuning it")
    ap.add_argument("--no-amr", action="store_true",
                    help="skip the batched AMR patch-table measurement")
    ap.add_argument("--no-hancock", action="store_true",
                    help="skip the Hancock-scheme measurement of the same workload")
    ap.add_argument("--no-fp64", action="store_true",
                    help="skip the FP64 run of the same workload (and the FP32 error)")
    ap.add_argument("--no-orszag", action="store_true",
                    help="skip the Orszag-Tang deck scheme (characteristic slopes) timing")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) update measurement")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the Brio-Wu (800 cells to t=0.1) and Orszag-Tang GLM "
                         "(512^2, 1000 steps) runs")
    ap.add_argument("--no-fill", action="store_true",
                    help="skip the level ghost-fill measurement")
    ap.add_argument("--stage", choices=["separate", "summed", "fused"], default=None,
                    help="kernel structure of the RK2 stages (default: the library's)")
    ap.add_argument("--tz", type=int, default=8,
                    help="planes per z-chunk of the tile kernels (summed, fused)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the run, launch the library's copy kernel over MIB MiB "
                         "twice (calibrates HBM byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(a level's patch table would shard by patch, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    assert mfmhd.device_count() > 0, "bench.py needs a HIP device"
    prec = mfmhd.PREC_FP32_FLUX if args.precision == "fp32flux" else mfmhd.PREC_FP64
    n = args.n
    p, dx = c3_params(n, prec)
    ctx = mfmhd.Context(p)
    modes = {"separate": mfmhd.STAGE_SEPARATE, "summed": mfmhd.STAGE_SUMMED,
             "fused": mfmhd.STAGE_FUSED}
    if args.stage is not None:
        ctx.set_stage_kernel(modes[args.stage], args.tz)
    nc = ctx.ncomp
    ng = ctx.nghost
    lo, hi = (-ng,) * 3, (n - 1 + ng,) * 3
    m = n + 2 * ng
    U = mfmhd.DeviceArray((nc, m, m, m))
    Uf = mfmhd.fab(U.ptr, lo, hi, nc)
    ctx.init_problem(Uf, mfmhd.Box.of(lo, hi), dx)
    valid = mfmhd.Box.of((0, 0, 0), (n - 1,) * 3)
    dtn, _ = ctx.compute_dt(Uf, valid, dx)
    t = 0.0
    dt = None
    # autotuner: work walk of the kernels, timed on a copy of the state
    walk = ctx.autotune(Uf, valid, CFL0 * dtn, dx, reps=2) if not args.no_tune else ctx.walk()

    def next_dt(dtn, dt):
        # first step initial_cfl * dt_raw, then growth-limited cfl * dt_raw
        return CFL0 * dtn if dt is None else min(args.cfl * dtn, DT_GROWTH * dt)

    dts = []
    for _ in range(args.warmup):
        dt = next_dt(dtn, dt)
        dts.append(dt)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "update")
        t += dt
    mfmhd.synchronize()
    # per-kernel HIP events on the launch stream over the timed region
    ctx.set_timing(True)
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        dt = next_dt(dtn, dt)
        dts.append(dt)
        dtn = finite_dt(ctx.update_state(Uf, valid, dt, dx, t)[0], "update")
        t += dt
    e1.record()
    mfmhd.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_ms(e1) / args.steps
    cells = float(n) ** 3
    value = cells / (ms * 1e-3)
    fs = 4 if prec == mfmhd.PREC_FP32_FLUX else 8
    kern = {}
    for st in (0, 1):
        for kind, name in enumerate(mfmhd.KT_NAMES):
            kms, kn, kit = ctx.timing(st, kind)
            if kn:
                kern["%s/stage%d" % (name, st + 1)] = {
                    "avg_ms": kms / kn, "launches": kn, "items_per_launch": kit / kn}
    # Algorithmic bytes: per cell-update 5*nc*8
    # = 920 B: stage 1 reads U^n and writes U* (2*nc*8), stage 2 reads U* and
    # U^n and writes U^n+1 (3*nc*8).  The whole step's rate against 8 TB/s
    # is roofline.frac; the dominant kernel's own rate, with the same bytes
    # per stage, is roofline.kernel.
    step_bytes = 5 * nc * 8
    stage_bytes = {0: 2 * nc * 8, 1: 3 * nc * 8}
    # the stage structure the library ran (MFM_STAGE_AUTO chooses per call)
    used = lambda k: sum(ctx.timing(st, k)[1] for st in (0, 1))
    structure = ("fused" if used(mfmhd.KT_STAGE) else
                 ("summed" if used(mfmhd.KT_FSUM_PLASMA) else "separate"))
    kind = mfmhd.KT_STAGE if structure == "fused" else mfmhd.KT_UPDATE
    kname = mfmhd.KT_NAMES[kind]
    tkey = (n, args.precision, structure, kname)
    tot_bytes = tot_ms = 0.0
    launches = 0
    for st in (0, 1):
        kms, kn, kit = ctx.timing(st, kind)
        tot_bytes += stage_bytes[st] * kit
        tot_ms += kms
        launches += kn
    k_achieved = tot_bytes / (tot_ms * 1e-3)
    step_achieved = step_bytes * value
    out = {
        "metric": "cell-updates/s per MF-MHD step",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 state, %s fluxes" % ("f32" if prec else "f64"),
        "data": "synthetic (HELIOINIT state generated on the device)",
        "config": {"workload": "3-D multi-fluid heliospheric box %d^3, plasma + 3 "
                               "neutrals (23 comps), solar wind from an inner sphere r=0.1 "
                               "(7/cc, 450 km/s) into the LISM flow (26 km/s, 3 uG at 45 deg), "
                               "RK2, Roe8, minmod, 8-wave, charge exchange, single patch" % n,
                   "cells": int(cells), "components": nc, "precision": args.precision,
                   "dt_control": "cfl %.2g, initial_cfl %.2g, max_dt_growth %.2f" % (args.cfl, CFL0, DT_GROWTH),
                   "walk": {"row_tile": walk[0], "xcd_order": walk[1]},
                   "stage_kernels": structure,
                   "tile_kernels": ("%s column tiles, z-chunks of %s" %
                                    ("64 x 8" if structure == "summed" else "16 x 16",
                                     "%d planes" % args.tz if args.stage is not None
                                     else "8 planes (library default)")
                                    if structure in ("summed", "fused") else None)},
        "roofline": {"bound": "hbm", "scope": "whole RK2 step, %d B per cell-update" % step_bytes,
                     "achieved": step_achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": step_achieved / HBM_PEAK,
                     "traffic": TRAFFIC.get(tkey),
                     "kernel": {"name": kname,
                                "alg_bytes_per_cell": {"stage1": stage_bytes[0],
                                                       "stage2": stage_bytes[1]},
                                "alg_bytes_per_launch": tot_bytes / launches,
                                "avg_launch_ms": tot_ms / launches,
                                "achieved": k_achieved / 1e9,
                                "frac": k_achieved / HBM_PEAK,
                                "traffic_over_alg": (
                                    TRAFFIC[tkey] / (tot_bytes / launches)
                                    if tkey in TRAFFIC else None)}},
        "step_hbm_frac": step_achieved / HBM_PEAK,
        "kernels": kern,
        "wall_s_timed": wall,
    }
    if not args.no_host:
        out["host_path"] = host_path_bench(ctx, U, lo, valid, dt, dx, t)
    ctx.close()
    final32 = None
    if prec == mfmhd.PREC_FP32_FLUX and not args.no_fp64:
        final32 = U.numpy()[:, ng:ng + n, ng:ng + n, ng:ng + n].copy()
    U.free()
    if final32 is not None:
        out["fp64"] = fp64_bench(n, dts, args.warmup, walk, final32)
        note("FP64 run done")
        del final32
    if args.calib_copy > 0:
        # the library's copy kernel (k_copy: reads and writes MIB MiB per
        # launch, one warm-up launch and one timed): the reference dispatch
        # for scaling FETCH_SIZE / WRITE_SIZE to bytes
        mfmhd.hbm_ceiling(mfmhd.CEIL_COPY, args.calib_copy << 20, 1)
    note("timed step done: %.3f ms/step" % ms)
    if not args.no_serve:
        out["ensemble"] = serve_bench(prec, n_req=args.serve_requests)
        note("ensemble done")
    if not args.no_amr:
        out["amr_batched"] = amr_bench(prec)
        note("batched AMR done")
    if not args.no_hancock:
        out["hancock"] = hancock_bench(prec, n, args.steps, args.warmup, walk)
        note("Hancock done")
    if not args.no_fill:
        out["ghost_fill"] = fill_bench()
    if not args.no_orszag:
        out["orszag_char"] = orszag_bench()
    if not args.no_small:
        out["brio_wu_800"] = brio_wu_bench()
        out["orszag_glm_512"] = orszag_glm_bench()
        out["orszag_glm_512_hllc"] = orszag_glm_bench(flux=mfmhd.FLUX_HLLC)
        note("Brio-Wu and Orszag-Tang done")
    out["roofline"]["copy_ceiling_GBps"] = copy_ceiling() / 1e9
    # the library's own streaming kernels over 2 GiB arrays
    out["roofline"]["ceilings_GBps"] = {
        name: mfmhd.hbm_ceiling(kind, 2 << 30)[0]
        for name, kind in (("read", mfmhd.CEIL_READ), ("copy", mfmhd.CEIL_COPY),
                           ("triad", mfmhd.CEIL_TRIAD))}
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        note("CPU baseline (oracle, %d threads and 1) ..." % threads)
        cb = cpu_baseline(args.cpu_sample, threads)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
