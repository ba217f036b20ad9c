"""Edge cases of the device update through the C ABI: an empty table is
refused on the host before any device work (oversized FABs are refused by
the same host checks, tested without a device through mfm_level_check in
test_capi_cpu.py), and degenerate
patches (one cell, one-cell-thick slabs, a table mixing them with a large
patch) match the oracle / their single-patch updates."""
import numpy as np
import pytest

import mfmhd
from helpers import GpuPatch, gpu_update, grid_shape, periodic_fill, smooth_state, valid_view
from test_gpu_parity import check64, periodic_params, run_pair

pytestmark = pytest.mark.gpu


def test_empty_table_refused():
    ctx = mfmhd.Context(periodic_params(nf=1))
    with pytest.raises(mfmhd.MfmError) as e:
        ctx.update_level([], 0.01, 0.05)
    assert e.value.code == mfmhd.ERR_ARG
    ctx.close()


@pytest.mark.parametrize("vhi", [(0, 0, 0), (0, 11, 9), (9, 0, 11), (11, 9, 0), (13, 0, 0)])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_HANCOCK])
def test_degenerate_patches(oracle_mod, vhi, time):
    print("degenerate patch", vhi, time, flush=True)
    p = periodic_params(nf=2, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_VANLEER, time=time)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo = (0, 0, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=sum(vhi) + time)
    g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.005, 0.05)
    check64(g, o, lo, vlo, vhi, U_in=U)


def test_table_mixing_tiny_and_large_patches():
    p = periodic_params(nf=4, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_MINMOD)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    boxes = [((0, 0, 0), (0, 0, 0)), ((5, 5, 5), (36, 27, 20)), ((-3, 7, 2), (-3, 7, 9)),
             ((40, 0, 0), (41, 1, 0))]
    states = []
    for q, (vlo, vhi) in enumerate(boxes):
        lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
        states.append((smooth_state(p, shape, lo, seed=70 + q), lo, vlo, vhi))
    ctx = mfmhd.Context(p)
    singles = [gpu_update(ctx, U, lo, vlo, vhi, 0.004, 0.05) for U, lo, vlo, vhi in states]
    pats = [GpuPatch(U, lo) for U, lo, _, _ in states]
    ctx.update_level([(g.fab, mfmhd.Box.of(s[2], s[3])) for g, s in zip(pats, states)],
                     0.004, 0.05)
    mfmhd.synchronize()
    for q, (g, s) in enumerate(zip(pats, singles)):
        assert np.array_equal(g.numpy(), s["U"]), q
    ctx.close()



def trough_state(p, n, ng):
    """A periodic state whose plasma is scaled by a deep, steep trough along
    x (0.01, 0.2, 3.0 at x = 5, 6, 7; pressure scales with it): hyperbee
    slopes then give non-positive face densities."""
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = smooth_state(p, shape, lo, seed=3)
    f = np.ones(n)
    f[5], f[6], f[7] = 0.01, 0.2, 3.0
    V = U[:, ng:ng + n, ng:ng + n, ng:ng + n]
    b2 = (V[5] ** 2 + V[6] ** 2 + V[7] ** 2) / (8 * np.pi)
    V[0:4] *= f
    V[4] = f * (V[4] - b2) + b2
    periodic_fill(U, ng)
    return U, lo, vlo, vhi


@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_HANCOCK])
def test_check_positivity_has_no_effect(oracle_mod, time):
    """mhdam.check_positivity is inert in the reference (main/MHDAM.cpp:978-987
    only queries slopes::checkPositivity(), src/slopes.cpp:84-87, and the flag
    stays false), so REDUCEORDER never runs there.  On a state with
    non-positive face densities the device update is the same bit for bit
    with the option on or off, and matches the oracle."""
    print("check_positivity", time, flush=True)
    p = periodic_params(nf=2, flux=mfmhd.FLUX_ROE8, lim=mfmhd.LIM_HYPERBEE, time=time)
    ng, n, dx = mfmhd.num_ghost_cells(p), 12, 0.05
    U, lo, vlo, vhi = trough_state(p, n, ng)
    WM, WP = oracle_mod.face_values(p, U, lo, 0, (-1, -1, -1), (n, n, n))
    faces = np.concatenate([valid_view(W, lo, (-1,) * 3, (n,) * 3)[0].ravel() for W in (WM, WP)])
    assert (faces <= 0).sum() > 0
    dt, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
    out = []
    for cp in (0, 1):
        p.check_positivity = cp
        g, o = run_pair(oracle_mod, p, U, lo, vlo, vhi, 0.3 * dt, dx)
        check64(g, o, lo, vlo, vhi, U_in=U)
        out.append(g)
    assert np.array_equal(out[0]["U"], out[1]["U"])
    assert out[0]["dt"] == out[1]["dt"] and out[0]["cell"] == out[1]["cell"]
