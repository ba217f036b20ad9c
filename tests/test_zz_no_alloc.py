"""The allocation-free update (mfm_patch_update_noalloc_dev) against the
allocating entry point (mfm_patch_update_dev) through the C ABI: same
state, face fluxes, div(B), dt and dt cell, bit for bit, on two calls in a
row whose second patch is smaller (the fixed scratch keeps the first call's
contents, which must not leak into the second), in both precisions; and a
patch whose scratch exceeds the fixed array is refused before any device
work, leaving U untouched."""
import numpy as np
import pytest

import mfmhd
from helpers import gpu_update, grid_shape, GpuPatch

pytestmark = pytest.mark.gpu


def helio(n, nf, prec):
    dx = 1.0 / n
    p = mfmhd.helio_params(lismB=3e-6, lismBdir=(np.cos(np.pi / 4), np.sin(np.pi / 4), 0),
                           lismV=2.6e6, sunN=7.0, sunV=4.5e7,
                           sunXYZ=(0.5, 0.5, 0.5), R0=0.12, initR=0.3,
                           netN=0.1, dx=dx, num_cells=(n, n, n))
    p.nfluids = nf
    p.time_approx = mfmhd.TIME_RK2
    p.precision = prec
    return p, dx


def same(a, b):
    assert a["rc"] == 0 and b["rc"] == 0, (a["rc"], b["rc"])
    assert np.array_equal(a["U"], b["U"])
    for fa, fb in zip(a["F"], b["F"]):
        assert np.array_equal(fa, fb)
    assert np.array_equal(a["divB"], b["divB"])
    assert a["dt"] == b["dt"] and a["cell"] == b["cell"]


@pytest.mark.parametrize("prec", [mfmhd.PREC_FP64, mfmhd.PREC_FP32_FLUX])
def test_noalloc_matches_allocating_entry_point(oracle_mod, prec):
    for n in (24, 12):  # the second call is smaller
        print("noalloc helio n", n, "prec", prec, flush=True)
        p, dx = helio(n, 4, prec)
        nc = mfmhd.num_components(p)
        vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
        lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
        U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
        dt, _ = oracle_mod.compute_dt(p, U, lo, vlo, vhi, dx)
        ctx = mfmhd.Context(p)
        a = gpu_update(ctx, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0, True, True, noalloc=True)
        b = gpu_update(ctx, U, lo, vlo, vhi, 0.5 * dt, dx, 0.0, True, True, noalloc=False)
        ctx.close()
        same(a, b)
        # and a sub-patch touching two boundaries, again smaller
        a = gpu_update(mfmhd.Context(p), U, lo, (0, 2, 3), (n // 2, n - 1, n - 2),
                       0.5 * dt, dx, 0.0, True, True, noalloc=True)
        b = gpu_update(mfmhd.Context(p), U, lo, (0, 2, 3), (n // 2, n - 1, n - 2),
                       0.5 * dt, dx, 0.0, True, True, noalloc=False)
        same(a, b)


def test_noalloc_refuses_patches_beyond_the_fixed_scratch(oracle_mod):
    # 64^3 cells + 4 ghosts, 23 components: the first-stage state alone is
    # 69 MB, the whole FP64 scratch about 200 MB, beyond the 128 MiB array
    n = 64
    p, dx = helio(n, 4, mfmhd.PREC_FP64)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1, n - 1, n - 1)
    lo, hi, shape = grid_shape(nc, vlo, vhi, 4)
    U = np.full(shape, 1.0)
    g = GpuPatch(U, lo)
    ctx = mfmhd.Context(p)
    rc, _, _ = ctx.update_state(g.fab, mfmhd.Box.of(vlo, vhi), 1e-3, dx, check=False,
                                noalloc=True)
    assert rc == mfmhd.ERR_NOMEM
    assert "fixed scratch" in ctx.last_error()
    assert np.array_equal(g.numpy(), U)
    ctx.close()
