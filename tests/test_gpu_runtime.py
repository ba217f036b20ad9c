"""Runtime entry points of the C ABI beside the update: the periodic ghost
fill of a one-box level (mfm_fill_periodic_dev), per-kernel timing
(mfm_ctx_set_timing / mfm_ctx_get_timing), the work walk setters, the
event helpers, and the box checks of mfm_patch_compute_dt_dev.  None of them
may change the update's results."""
import numpy as np
import pytest

import mfmhd
from helpers import GpuPatch, grid_shape, gpu_update, smooth_state
from test_gpu_parity import helio_setup, periodic_params

pytestmark = pytest.mark.gpu


def periodic_reference(U, lo, vlo, vhi, dims):
    """Every cell copies the valid cell congruent to it modulo the valid
    extent, per periodic direction (several wraps when the ghost depth
    exceeds the extent)."""
    idx = []
    for d, axis in ((0, 3), (1, 2), (2, 1)):
        i = np.arange(U.shape[axis]) + lo[d]
        if d < dims:
            n = vhi[d] - vlo[d] + 1
            i = vlo[d] + np.mod(i - vlo[d], n)
        idx.append(i - lo[d])
    return U[:, idx[2]][:, :, idx[1]][:, :, :, idx[0]]


@pytest.mark.parametrize("dims,vhi", [(3, (9, 6, 4)), (3, (4, 2, 7)), (2, (7, 5, 0))])
def test_fill_periodic_matches_modular_copy(dims, vhi):
    p = periodic_params(nf=2, dims=dims)
    nc = mfmhd.num_components(p)
    ng = 4
    vlo = (0, 0, 0)
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng, dims)
    rng = np.random.default_rng(sum(vhi) + dims)
    U = rng.standard_normal(shape)
    ctx = mfmhd.Context(p)
    g = GpuPatch(U, lo)
    ctx.fill_periodic(g.fab, mfmhd.Box.of(vlo, vhi))
    got = g.numpy()
    assert np.array_equal(got, periodic_reference(U, lo, vlo, vhi, dims))
    ctx.close()


def test_fill_periodic_refuses_bad_boxes():
    p = periodic_params(nf=1)
    ctx = mfmhd.Context(p)
    U = np.arange(8 * 6 * 6 * 6, dtype=np.float64).reshape(8, 6, 6, 6)
    g = GpuPatch(U, (0, 0, 0))
    for vlo, vhi in (((2, 2, 2), (1, 3, 3)),    # empty
                     ((1, 1, 1), (6, 4, 4))):   # outside U
        with pytest.raises(mfmhd.MfmError) as e:
            ctx.fill_periodic(g.fab, mfmhd.Box.of(vlo, vhi))
        assert e.value.code == mfmhd.ERR_SHAPE
    assert np.array_equal(g.numpy(), U)
    ctx.close()


@pytest.mark.parametrize("mode", [mfmhd.STAGE_SEPARATE, mfmhd.STAGE_FUSED, mfmhd.STAGE_SUMMED])
def test_kernel_timing_counts_launches_and_keeps_results(oracle_mod, mode):
    n = 16
    p, dx = helio_setup(oracle_mod, n, 4, mfmhd.PREC_FP32_FLUX)
    nc = mfmhd.num_components(p)
    ng = mfmhd.num_ghost_cells(p)
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    dt = 0.5 * oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)[0]
    ctx = mfmhd.Context(p)
    ctx.set_stage_kernel(mode)
    plain = gpu_update(ctx, U, lo, vlo, vhi, dt, dx)
    ctx.set_timing(True)
    timed = gpu_update(ctx, U, lo, vlo, vhi, dt, dx)
    assert np.array_equal(plain["U"], timed["U"]) and plain["dt"] == timed["dt"]
    # one RK2 step: one update per stage, the corrector's over the valid box
    # (the fused stage kernel, or k_update after the separate flux kernels or
    # after the face-sum kernels: one plasma and three neutral launches)
    fused = mode == mfmhd.STAGE_FUSED
    upd, other = (mfmhd.KT_STAGE, mfmhd.KT_UPDATE) if fused else (mfmhd.KT_UPDATE, mfmhd.KT_STAGE)
    sums = mode == mfmhd.STAGE_SUMMED
    for st in (0, 1):
        ms, launches, items = ctx.timing(st, upd)
        assert launches == 1 and ms > 0.0, (st, ms, launches)
        assert items >= n ** 3
        assert ctx.timing(st, other)[1] == 0
        assert ctx.timing(st, mfmhd.KT_FLUX_PLASMA)[1] == (3 if mode == mfmhd.STAGE_SEPARATE else 0)
        assert ctx.timing(st, mfmhd.KT_FSUM_PLASMA)[1] == (1 if sums else 0)
        assert ctx.timing(st, mfmhd.KT_FSUM_NEUTRAL)[1] == (3 if sums else 0)
        assert ctx.timing(st, mfmhd.KT_FLUX_BC)[1] == 3   # boundary faces of x, y, z
    assert ctx.timing(1, upd)[2] == n ** 3
    assert ctx.timing(0, mfmhd.KT_PRIM)[1] == 1
    assert ctx.timing(0, mfmhd.KT_PREDICT)[1] == 0    # no Hancock predictor in RK2
    # switching timing on again starts new sums
    ctx.set_timing(True)
    assert ctx.timing(0, mfmhd.KT_UPDATE) == (0.0, 0, 0.0)
    with pytest.raises(mfmhd.MfmError):
        ctx.timing(2, mfmhd.KT_UPDATE)
    with pytest.raises(mfmhd.MfmError):
        ctx.timing(0, len(mfmhd.KT_NAMES))
    ctx.set_timing(False)
    ctx.close()


def test_walk_setters_and_events():
    p = periodic_params(nf=1)
    ctx = mfmhd.Context(p)
    ctx.set_walk(4, 0)
    assert ctx.walk() == (4, 0)
    ctx.set_walk(16, 5)               # any non-zero XCD flag means on
    assert ctx.walk() == (16, 1)
    for bad in (0, 65, -3):
        with pytest.raises(mfmhd.MfmError) as e:
            ctx.set_walk(bad, 1)
        assert e.value.code == mfmhd.ERR_ARG
    assert ctx.walk() == (16, 1)
    # events bracket device work in stream order
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (11, 11, 11)
    lo, hi, shape = grid_shape(nc, vlo, vhi, mfmhd.num_ghost_cells(p))
    U = smooth_state(p, shape, lo, seed=4)
    g = GpuPatch(U, lo)
    e0, e1 = mfmhd.Event(), mfmhd.Event()
    e0.record()
    ctx.update_state(g.fab, mfmhd.Box.of(vlo, vhi), 1e-3, 0.05)
    e1.record()
    mfmhd.synchronize()
    assert e0.elapsed_ms(e1) > 0.0
    ctx.close()


def test_compute_dt_refuses_empty_or_outside_box():
    """Refused on the host before any device work, like the update."""
    p = periodic_params(nf=1)
    ctx = mfmhd.Context(p)
    U = smooth_state(p, (8, 6, 6, 6), (0, 0, 0), seed=1)
    g = GpuPatch(U, (0, 0, 0))
    for vlo, vhi in (((2, 2, 2), (1, 3, 3)), ((1, 1, 1), (6, 4, 4))):
        with pytest.raises(mfmhd.MfmError) as e:
            ctx.compute_dt(g.fab, mfmhd.Box.of(vlo, vhi), 0.05)
        assert e.value.code == mfmhd.ERR_SHAPE
    dt, cell = ctx.compute_dt(g.fab, mfmhd.Box.of((1, 1, 1), (4, 4, 4)), 0.05)
    assert dt > 0.0 and all(1 <= c <= 4 for c in cell)
    ctx.close()
