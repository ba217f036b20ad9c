"""The tile kernels of a stage against the separate flux and update kernels,
and against the oracle, through the C ABI: the fused stage kernel (k_stage:
interior face fluxes of a tile solved into LDS and the cell update in one
launch) and the summed stages (k_fsum per fluid group: the face-flux
differences of each cell summed over the directions, then k_update).

The fused and separate structures run the same arithmetic per face and per
cell, so in the FP64 build the states, face fluxes, div(B), dt and dt cell
agree to the last bit; the summed stages exist in the FP32-flux build only
(the FP64 build runs the separate kernels in that mode, so it agrees to the
last bit too); in the FP32-flux build the compiler may contract a few
products differently in the kernels, and the summed stages add the three
directions' flux differences before applying them and store the sums in
FP32, so on smooth periodic states they agree to 1e-6 relative per
component; on the heliospheric box, whose FP32 error inside the solar-wind
sphere is 3e-4 of max |B| in every structure, each is held to the FP32 bar
against the FP64 oracle (relative L1 <= 1e-5).  Cases cover every equation system and solver family, every
problem's boundary faces (read by k_stage from the boundary kernel's face
arrays), partial tiles, z-chunks shorter than a tile column, 2-D, first
order and the caller's flux outputs."""
import numpy as np
import pytest

import mfmhd
from helpers import (gpu_update, grid_shape, l1_rel_vector_scaled, periodic_fill, smooth_state,
                     valid_view, vector_groups)
from test_gpu_parity import check64, helio_setup, periodic_params

pytestmark = pytest.mark.gpu


MODES = [mfmhd.STAGE_FUSED, mfmhd.STAGE_SUMMED]


def fp32_l1(g, o, lo, vlo, vhi, nf, ded=False):
    """The FP32-flux bar: relative L1 error per variable <= 1e-5 against
    the FP64 oracle (each vector's magnitude norm for its components)."""
    assert g["rc"] == 0 and o["rc"] == 0
    groups = vector_groups(nf, ded)
    if ded:
        groups.append((5, 6, 7, 8))
    err = l1_rel_vector_scaled(valid_view(g["U"], lo, vlo, vhi), valid_view(o["U"], lo, vlo, vhi),
                               groups)
    assert err.max() <= 1e-5, err


def both(p, U, lo, vlo, vhi, dt, dx, tz=16, time=0.0, ch=None, mode=mfmhd.STAGE_FUSED):
    out = []
    for m in (mode, mfmhd.STAGE_SEPARATE):
        ctx = mfmhd.Context(p)
        if ch is not None:
            ctx.set_dedner(*ch)
        ctx.set_stage_kernel(m, tz)
        out.append(gpu_update(ctx, U, lo, vlo, vhi, dt, dx, time, True, bool(p.eight_wave)))
        ctx.close()
    return out


def agree(a, b, fp64=True, mode=mfmhd.STAGE_FUSED):
    assert a["rc"] == 0 and b["rc"] == 0, (a["rc"], b["rc"])
    tol = 0.0 if fp64 else 1e-6
    for x, y in [(a["U"], b["U"])] + list(zip(a["F"], b["F"])):
        for c in range(y.shape[0]):
            den = max(np.max(np.abs(y[c])), 1e-300)
            assert np.max(np.abs(x[c] - y[c])) <= tol * den, c
    if a["divB"] is not None:
        den = max(np.max(np.abs(b["divB"])), 1e-300)
        assert np.max(np.abs(a["divB"] - b["divB"])) <= max(tol, 1e-14) * den
    if fp64:
        assert a["dt"] == b["dt"] and a["cell"] == b["cell"]
    else:
        assert abs(a["dt"] - b["dt"]) <= 1e-6 * b["dt"]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("prec", [mfmhd.PREC_FP64, mfmhd.PREC_FP32_FLUX])
@pytest.mark.parametrize("nf", [4, 3, 2, 1])
def test_fused_helio_matches_separate_kernels(oracle_mod, nf, prec, mode):
    """Heliospheric box: LISM inflow and characteristic boundary faces on
    every side, charge exchange, re-init; a 26^3 box is not a multiple of
    the 16 x 16 tile, and tz = 5 leaves a short last z-chunk."""
    print("tile kernels helio nf", nf, "prec", prec, "mode", mode, flush=True)
    n = 26
    p, dx = helio_setup(oracle_mod, n, nf, prec)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dt, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    a, b = both(p, U, lo, vlo, vhi, 0.5 * dt, dx, tz=5, mode=mode)
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0)
    if prec == mfmhd.PREC_FP64 or mode == mfmhd.STAGE_FUSED:
        agree(a, b, prec == mfmhd.PREC_FP64, mode)
    else:
        # the FP32-flux build's error in B inside the solar-wind sphere
        # (where B = 0 at t = 0) is 3e-4 of max |B| in every kernel
        # structure: summed and separate land on different roundings of it,
        # so they are held to the FP32 bar against the FP64 oracle instead
        fp32_l1(a, o, lo, vlo, vhi, nf)
        fp32_l1(b, o, lo, vlo, vhi, nf)
    if prec == mfmhd.PREC_FP64:
        check64(a, o, lo, vlo, vhi)
    # a sub-patch touching three boundaries
    vlo, vhi = (3, 0, 9), (25, 17, 25)
    a, b = both(p, U, lo, vlo, vhi, 0.5 * dt, dx, tz=16, mode=mode)
    if prec == mfmhd.PREC_FP64 or mode == mfmhd.STAGE_FUSED:
        agree(a, b, prec == mfmhd.PREC_FP64, mode)
    else:
        fp32_l1(a, oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0), lo, vlo, vhi, nf)


@pytest.mark.parametrize("case", [
    ("roe8-minmod-1f", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_MINMOD, False),
    ("rusanov-mc-2f", 2, mfmhd.FLUX_RUSANOV, mfmhd.LIM_MC, False),
    ("roe8avg-vanleer-4f", 4, mfmhd.FLUX_ROE8_ROEAVG, mfmhd.LIM_VANLEER, False),
    ("glm-superbee", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_SUPERBEE, True),
    ("roe7-mc-2f", 2, mfmhd.FLUX_ROE7, mfmhd.LIM_MC, False),
    ("roe7avg-minmod-1f", 1, mfmhd.FLUX_ROE7_ROEAVG, mfmhd.LIM_MINMOD, False),
    ("hllc-minmod-2f", 2, mfmhd.FLUX_HLLC, mfmhd.LIM_MINMOD, False),
    ("hllc-glm", 1, mfmhd.FLUX_HLLC, mfmhd.LIM_MINMOD, True),
], ids=lambda c: c[0])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_FIRST_ORDER])
@pytest.mark.parametrize("mode", MODES)
def test_fused_periodic_matches_separate_kernels(oracle_mod, case, time, mode):
    name, nf, flux, lim, ded = case
    print("tile kernels periodic", name, time, mode, flush=True)
    p = periodic_params(nf=nf, flux=flux, lim=lim, dedner=ded, eight=not ded, time=time)
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (19, 33, 11)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = periodic_fill(smooth_state(p, shape, lo, seed=7), ng)
    dx = 1.0 / 32
    ch = (1.3, 0.9) if ded else None
    a, b = both(p, U, lo, vlo, vhi, 2e-3, dx, tz=7, ch=ch, mode=mode)
    agree(a, b, True, mode)
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 2e-3, dx, 0.0,
                                **({"ch": 1.3, "cp": 0.9} if ded else {}))
    check64(a, o, lo, vlo, vhi)


@pytest.mark.parametrize("mode", MODES)
def test_fused_2d_and_brio_wu(oracle_mod, mode):
    """2-D periodic (Orszag-Tang with GLM) and the Brio-Wu tube with its
    fixed-state boundary faces, RK2."""
    p = periodic_params(nf=1, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_MINMOD, dims=2,
                        dedner=True, eight=False)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (40, 23, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = periodic_fill(smooth_state(p, shape, lo, seed=3), 4, dims=2)
    a, b = both(p, U, lo, vlo, vhi, 1e-3, 1.0 / 41, ch=(1.0, 1.0), mode=mode)
    agree(a, b, True, mode)
    S4PI = np.sqrt(4 * np.pi)
    p = mfmhd.riemann_params(2.0, [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, S4PI, 0],
                             [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -S4PI, 0], 0.5)
    p.spacedim = 2
    p.time_approx = mfmhd.TIME_RK2
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [99, 3, 0]
    p.periodic[:] = [0, 1, 0]
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (99, 3, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    dx = 1.0 / 100
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    a, b = both(p, U, lo, vlo, vhi, 1e-4, dx, mode=mode)
    agree(a, b, True, mode)


@pytest.mark.parametrize("case", [
    ("roe8-minmod-1f", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_MINMOD, False, 3),
    ("rusanov-mc-2f", 2, mfmhd.FLUX_RUSANOV, mfmhd.LIM_MC, False, 3),
    ("roe8avg-vanleer-4f", 4, mfmhd.FLUX_ROE8_ROEAVG, mfmhd.LIM_VANLEER, False, 3),
    ("glm-superbee", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_SUPERBEE, True, 3),
    ("glm-2d", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_MINMOD, True, 2),
    ("roe8-3f-2d", 3, mfmhd.FLUX_ROE8, mfmhd.LIM_MINMOD, False, 2),
    ("roe7-2f", 2, mfmhd.FLUX_ROE7, mfmhd.LIM_MINMOD, False, 3),
    ("roe7avg-1f-2d", 1, mfmhd.FLUX_ROE7_ROEAVG, mfmhd.LIM_MC, False, 2),
    ("hllc-3f", 3, mfmhd.FLUX_HLLC, mfmhd.LIM_MINMOD, False, 3),
    ("hllc-glm-2d", 1, mfmhd.FLUX_HLLC, mfmhd.LIM_MINMOD, True, 2),
], ids=lambda c: c[0])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_FIRST_ORDER])
def test_summed_fp32_periodic(oracle_mod, case, time):
    """Summed stages of the FP32-flux build against its separate kernels
    (1e-6 per component) and against the FP64 oracle (relative L1 <= 1e-5
    per variable, vector norm), every solver family, 2-D and 3-D."""
    name, nf, flux, lim, ded, dims = case
    print("summed fp32 periodic", name, time, flush=True)
    p = periodic_params(nf=nf, flux=flux, lim=lim, dedner=ded, eight=not ded, time=time,
                        dims=dims, prec=mfmhd.PREC_FP32_FLUX)
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (19, 33, 11 if dims == 3 else 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng, dims)
    U = periodic_fill(smooth_state(p, shape, lo, seed=11), ng, dims)
    dx = 1.0 / 32
    ch = (1.3, 0.9) if ded else None
    a, b = both(p, U, lo, vlo, vhi, 2e-3, dx, tz=7, ch=ch, mode=mfmhd.STAGE_SUMMED)
    agree(a, b, False, mfmhd.STAGE_SUMMED)
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 2e-3, dx, 0.0,
                                **({"ch": 1.3, "cp": 0.9} if ded else {}))
    fp32_l1(a, o, lo, vlo, vhi, nf, ded)


@pytest.mark.parametrize("nf", [4, 2])
def test_summed_fp32_helio_and_brio_wu(oracle_mod, nf):
    """The heliospheric box (every boundary face read from the boundary
    kernel's arrays, charge exchange, re-init) and the Brio-Wu tube with its
    fixed-state faces, summed stages of the FP32-flux build against the
    FP64 oracle (relative L1 <= 1e-5 per variable)."""
    n = 26
    p, dx = helio_setup(oracle_mod, n, nf, mfmhd.PREC_FP32_FLUX)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dt, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    a, b = both(p, U, lo, vlo, vhi, 0.5 * dt, dx, tz=5, mode=mfmhd.STAGE_SUMMED)
    fp32_l1(a, oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0), lo, vlo, vhi, nf)
    sv = ((3, 0, 9), (25, 17, 25))
    a, b = both(p, U, lo, *sv, 0.5 * dt, dx, mode=mfmhd.STAGE_SUMMED)
    fp32_l1(a, oracle_mod.patch_update(p, U, lo, *sv, 0.5 * dt, dx, 0.0), lo, *sv, nf)
    S4PI = np.sqrt(4 * np.pi)
    p = mfmhd.riemann_params(2.0, [1.0, 0, 0, 0, 1.0, 0.75 * S4PI, S4PI, 0],
                             [0.125, 0, 0, 0, 0.1, 0.75 * S4PI, -S4PI, 0], 0.5)
    p.spacedim = 2
    p.time_approx = mfmhd.TIME_RK2
    p.precision = mfmhd.PREC_FP32_FLUX
    p.domain_lo[:] = [0, 0, 0]
    p.domain_hi[:] = [99, 3, 0]
    p.periodic[:] = [0, 1, 0]
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (99, 3, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    dx = 1.0 / 100
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    a, b = both(p, U, lo, vlo, vhi, 1e-4, dx, mode=mfmhd.STAGE_SUMMED)
    fp32_l1(a, oracle_mod.patch_update(p, U, lo, vlo, vhi, 1e-4, dx, 0.0), lo, vlo, vhi, 1)


def test_auto_mode_choice(oracle_mod):
    """MFM_STAGE_AUTO, the default: the summed stages where the cells fill
    at least 3/4 of their 64 x 8 column tiles in the FP32-flux build, the
    separate kernels otherwise (a 20-wide box, the FP64 build, the
    allocation-free entry point); each choice meets its bar against the
    oracle."""
    def kinds(p, vlo, vhi, noalloc=False):
        nc = mfmhd.num_components(p)
        lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
        U = periodic_fill(smooth_state(p, shape, lo, seed=5), 4)
        ctx = mfmhd.Context(p)
        ctx.set_timing(True)
        g = gpu_update(ctx, U, lo, vlo, vhi, 2e-3, 1.0 / 32, noalloc=noalloc)
        used = {k: ctx.timing(s, k)[1] for s in (0, 1)
                for k in (mfmhd.KT_FSUM_PLASMA, mfmhd.KT_FLUX_PLASMA)}
        ctx.close()
        o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 2e-3, 1.0 / 32, 0.0)
        return g, o, lo, used
    wide = ((0, 0, 0), (63, 15, 5))
    p = periodic_params(nf=2, prec=mfmhd.PREC_FP32_FLUX)
    g, o, lo, used = kinds(p, *wide)
    assert used[mfmhd.KT_FSUM_PLASMA] > 0 and used[mfmhd.KT_FLUX_PLASMA] == 0
    fp32_l1(g, o, lo, *wide, 2)
    g, o, lo, used = kinds(p, (0, 0, 0), (19, 15, 5))
    assert used[mfmhd.KT_FSUM_PLASMA] == 0 and used[mfmhd.KT_FLUX_PLASMA] > 0
    fp32_l1(g, o, lo, (0, 0, 0), (19, 15, 5), 2)
    g, o, lo, used = kinds(p, *wide, noalloc=True)
    assert used[mfmhd.KT_FSUM_PLASMA] == 0 and used[mfmhd.KT_FLUX_PLASMA] > 0
    fp32_l1(g, o, lo, *wide, 2)
    p = periodic_params(nf=2)
    g, o, lo, used = kinds(p, *wide)
    assert used[mfmhd.KT_FSUM_PLASMA] == 0 and used[mfmhd.KT_FLUX_PLASMA] > 0
    check64(g, o, lo, *wide)
