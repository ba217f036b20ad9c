"""The fused stage kernel (k_stage: interior face fluxes of a tile solved into
LDS and the cell update in one launch) against the separate flux and update
kernels, and against the oracle, through the C ABI.

Both kernel structures run the same arithmetic per face and per cell, so in
the FP64 build the states, face fluxes, div(B), dt and dt cell agree to the
last bit; in the FP32-flux build the compiler may contract a few products
differently in the two kernels, so they agree to 1e-6 relative per
component.  Cases cover every equation system and solver family, every
problem's boundary faces (read by k_stage from the boundary kernel's face
arrays), partial tiles, z-chunks shorter than a tile column, 2-D, first
order and the caller's flux outputs."""
import numpy as np
import pytest

import mfmhd
from helpers import gpu_update, grid_shape, periodic_fill, smooth_state, valid_view
from test_gpu_parity import check64, helio_setup, periodic_params

pytestmark = pytest.mark.gpu


def both(p, U, lo, vlo, vhi, dt, dx, tz=16, time=0.0, ch=None):
    out = []
    for fused in (True, False):
        ctx = mfmhd.Context(p)
        if ch is not None:
            ctx.set_dedner(*ch)
        ctx.set_stage_kernel(fused, tz)
        out.append(gpu_update(ctx, U, lo, vlo, vhi, dt, dx, time, True, bool(p.eight_wave)))
        ctx.close()
    return out


def agree(a, b, fp64=True):
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


@pytest.mark.parametrize("prec", [mfmhd.PREC_FP64, mfmhd.PREC_FP32_FLUX])
@pytest.mark.parametrize("nf", [4, 3, 2, 1])
def test_fused_helio_matches_separate_kernels(oracle_mod, nf, prec):
    """Heliospheric box: LISM inflow and characteristic boundary faces on
    every side, charge exchange, re-init; a 26^3 box is not a multiple of
    the 16 x 16 tile, and tz = 5 leaves a short last z-chunk."""
    print("fused helio nf", nf, "prec", prec, flush=True)
    n = 26
    p, dx = helio_setup(oracle_mod, n, nf, prec)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dt, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    a, b = both(p, U, lo, vlo, vhi, 0.5 * dt, dx, tz=5)
    agree(a, b, prec == mfmhd.PREC_FP64)
    if prec == mfmhd.PREC_FP64:
        o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0)
        check64(a, o, lo, vlo, vhi)
    # a sub-patch touching three boundaries
    a, b = both(p, U, lo, (3, 0, 9), (25, 17, 25), 0.5 * dt, dx, tz=16)
    agree(a, b, prec == mfmhd.PREC_FP64)


@pytest.mark.parametrize("case", [
    ("roe8-minmod-1f", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_MINMOD, False),
    ("rusanov-mc-2f", 2, mfmhd.FLUX_RUSANOV, mfmhd.LIM_MC, False),
    ("roe8avg-vanleer-4f", 4, mfmhd.FLUX_ROE8_ROEAVG, mfmhd.LIM_VANLEER, False),
    ("glm-superbee", 1, mfmhd.FLUX_ROE8, mfmhd.LIM_SUPERBEE, True),
], ids=lambda c: c[0])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_FIRST_ORDER])
def test_fused_periodic_matches_separate_kernels(oracle_mod, case, time):
    name, nf, flux, lim, ded = case
    print("fused periodic", name, time, flush=True)
    p = periodic_params(nf=nf, flux=flux, lim=lim, dedner=ded, eight=not ded, time=time)
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (19, 33, 11)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = periodic_fill(smooth_state(p, shape, lo, seed=7), ng)
    dx = 1.0 / 32
    ch = (1.3, 0.9) if ded else None
    a, b = both(p, U, lo, vlo, vhi, 2e-3, dx, tz=7, ch=ch)
    agree(a, b)
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, 2e-3, dx, 0.0,
                                **({"ch": 1.3, "cp": 0.9} if ded else {}))
    check64(a, o, lo, vlo, vhi)


def test_fused_2d_and_brio_wu(oracle_mod):
    """2-D periodic (Orszag-Tang with GLM) and the Brio-Wu tube with its
    fixed-state boundary faces, RK2."""
    p = periodic_params(nf=1, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_MINMOD, dims=2,
                        dedner=True, eight=False)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (40, 23, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = periodic_fill(smooth_state(p, shape, lo, seed=3), 4, dims=2)
    a, b = both(p, U, lo, vlo, vhi, 1e-3, 1.0 / 41, ch=(1.0, 1.0))
    agree(a, b)
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
    a, b = both(p, U, lo, vlo, vhi, 1e-4, dx)
    agree(a, b)
