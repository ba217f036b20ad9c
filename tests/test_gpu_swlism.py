"""GPU parity of the 2-D axisymmetric SW/LISM problem (swlism/
SWLISMProblem.cpp, the ensemble workload) against the oracle, through the C
ABI: LISM ghost states and the axis mirror, the BFLUX / BFLUXGAS boundary
faces, the geometric sources of coord_sys = 1, the axisymmetric div(B)
term, the re-init inside R0, regions and CHARGE_EXCHANGE_MF (4F for four
fluids), floors and the dt exclusion.  FP64 build <= 1e-12 relative per
component and step; FP32-flux build <= 1e-5 per-variable relative L1 (the
norm of each vector's magnitude for its components; the component's own
norm is printed beside it)."""
import numpy as np
import pytest

import mfmhd
from helpers import (gpu_update, grid_shape, l1_rel_per_comp, l1_rel_vector_scaled,
                     valid_view, vector_groups)
from test_gpu_parity import TOL32_L1, check64

pytestmark = pytest.mark.gpu


def swlism_setup(nf, nx=48, ny=32, time=mfmhd.TIME_RK2, prec=mfmhd.PREC_FP64,
                 flux=mfmhd.FLUX_ROE8_ROEAVG):
    """A 1500 x 1000 AU box (the deck's domain length) with the sun at
    0.4 of it on the axis; Roe-averaged 8-wave fluxes, which stay positive
    at the wind / LISM interface of the initial state."""
    dx = 1500.0 / nx
    p = mfmhd.swlism_params(sunXY=(0.4 * nx * dx, 0.0), R0=3.1 * dx, initR=8.0 * dx,
                            dx=dx, num_cells=(nx, ny), nfluids=nf, lismB=3e-6)
    p.time_approx = time
    p.precision = prec
    p.flux = flux
    return p, dx


def pair(oracle_mod, p, U, lo, vlo, vhi, dt, dx, fused=mfmhd.STAGE_SEPARATE):
    ctx = mfmhd.Context(p)
    ctx.set_stage_kernel(fused)
    g = gpu_update(ctx, U, lo, vlo, vhi, dt, dx, 0.0, True, bool(p.eight_wave))
    ctx.close()
    o = oracle_mod.patch_update(p, U, lo, vlo, vhi, dt, dx, 0.0, want_F=True,
                                want_divB=bool(p.eight_wave))
    return g, o


@pytest.mark.parametrize("fused", [mfmhd.STAGE_SEPARATE, mfmhd.STAGE_FUSED, mfmhd.STAGE_SUMMED])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_FIRST_ORDER])
@pytest.mark.parametrize("nf", [1, 2, 3, 4])
def test_swlism_patch_fp64(oracle_mod, nf, time, fused):
    print("swlism nf", nf, "time", time, "fused", fused, flush=True)
    nx, ny = 48, 32
    p, dx = swlism_setup(nf, nx, ny, time)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (nx - 1, ny - 1, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    dt = 0.1 * dtn   # the deck's initial_cfl, then its cfl
    for step in range(3):
        g, o = pair(oracle_mod, p, U, lo, vlo, vhi, dt, dx, fused)
        check64(g, o, lo, vlo, vhi)
        for d in range(2):
            assert np.max(np.abs(g["F"][d] - o["F"][d])) <= \
                1e-11 * max(np.max(np.abs(o["F"][d])), 1e-300), d
        U = o["U"]
        dt = 0.25 * o["dt"]
    # sub-patches: on the axis, on the right boundary, interior around the sun
    for a, b in (((0, 0, 0), (15, 9, 0)), ((30, 12, 0), (47, 31, 0)),
                 ((14, 3, 0), (27, 17, 0))):
        g, o = pair(oracle_mod, p, U, lo, a, b, dt, dx, fused)
        check64(g, o, lo, a, b)


@pytest.mark.parametrize("flux", [mfmhd.FLUX_ROE8, mfmhd.FLUX_RUSANOV, mfmhd.FLUX_ROE7,
                                  mfmhd.FLUX_ROE7_ROEAVG, mfmhd.FLUX_HLLC])
def test_swlism_other_solvers_fp64(oracle_mod, flux):
    """The deck's own solver (8-wave Roe on arithmetic means), Rusanov, the
    7-wave Roe solver and the HLLC flux (at the boundary faces too, before
    SWLISMBC), one step from the initial state at the deck's initial_cfl."""
    p, dx = swlism_setup(2, 48, 32, mfmhd.TIME_FIRST_ORDER, flux=flux)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (47, 31, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    g, o = pair(oracle_mod, p, U, lo, vlo, vhi, 0.1 * dtn, dx)
    check64(g, o, lo, vlo, vhi)


@pytest.mark.parametrize("nf", [1, 2, 4])
def test_swlism_fp32_flux_error(oracle_mod, nf):
    nx, ny = 64, 40
    p, dx = swlism_setup(nf, nx, ny, prec=mfmhd.PREC_FP32_FLUX)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (nx - 1, ny - 1, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dtn, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    g, o = pair(oracle_mod, p, U, lo, vlo, vhi, 0.25 * dtn, dx)
    assert g["rc"] == 0 and o["rc"] == 0
    a = valid_view(g["U"], lo, vlo, vhi)
    b = valid_view(o["U"], lo, vlo, vhi)
    own = l1_rel_per_comp(a, b)
    err = l1_rel_vector_scaled(a, b, vector_groups(nf))
    print("fp32 per-variable L1 (own norm):", " ".join("%.2e" % e for e in own))
    print("fp32 per-variable L1 (vector norm):", " ".join("%.2e" % e for e in err))
    assert err.max() <= TOL32_L1, err
    assert abs(g["dt"] - o["dt"]) <= 1e-5 * o["dt"]


def test_axisymmetric_caller_ghosts_fp64(oracle_mod):
    """coord_sys = 1 without a problem (caller-filled ghosts): the geometric
    sources and the axisymmetric div(B) term on a perturbed state."""
    p = mfmhd.orszag_params(5.0 / 3.0)
    p.spacedim = 2
    p.coord_sys = mfmhd.COORD_AXISYMMETRIC
    p.nfluids = 2
    p.time_approx = mfmhd.TIME_RK2
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 2, 0), (29, 21, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4, dims=2)
    rng = np.random.default_rng(17)
    W = np.zeros(shape)
    W[0] = 1 + 0.2 * rng.uniform(size=shape[1:])
    W[1:3] = 0.3 * rng.uniform(-1, 1, size=(2,) + shape[1:])
    W[4] = 0.8 + 0.2 * rng.uniform(size=shape[1:])
    W[5:7] = 0.5 * rng.uniform(-1, 1, size=(2,) + shape[1:])
    W[8] = 0.5 + 0.1 * rng.uniform(size=shape[1:])
    W[9:11] = 0.3 * rng.uniform(-1, 1, size=(2,) + shape[1:])
    W[12] = 0.4 + 0.1 * rng.uniform(size=shape[1:])
    U = W.copy()
    U[1:4] *= W[0]
    U[4] = W[4] / (2 / 3) + 0.5 * W[0] * (W[1] ** 2 + W[2] ** 2) + \
        (W[5] ** 2 + W[6] ** 2) / (8 * np.pi)
    U[9:12] *= W[8]
    U[12] = W[12] / (2 / 3) + 0.5 * W[8] * (W[9] ** 2 + W[10] ** 2)
    g, o = pair(oracle_mod, p, U, lo, vlo, vhi, 2e-3, 0.05)
    check64(g, o, lo, vlo, vhi)
