"""The device against properties of the heliospheric boundary conditions
that fix them independently of the oracle: a uniform interstellar state
(plasma and the first neutral population at its temperature) is a steady
solution of the box with all its boundaries (HELIOGS ghost states, the
fixed inflow flux, BFLUX / BFLUXR upwind branches, charge exchange in
equilibrium), in both builds and both time schemes."""
import numpy as np
import pytest

import mfmhd
from helpers import gpu_update, valid_view
from test_helio_pins import lism_params, uniform_lism

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [(mfmhd.PREC_FP64, 1e-12), (mfmhd.PREC_FP32_FLUX, 2e-6)])
@pytest.mark.parametrize("time", [mfmhd.TIME_RK2, mfmhd.TIME_HANCOCK])
@pytest.mark.parametrize("nf", [1, 2])
def test_device_lism_free_stream_is_steady(prec, tol, time, nf):
    print("free stream", prec, time, nf, flush=True)
    n = 12
    p = lism_params(n, 8.0e6, (0.6, 0.8, 0.0), (-0.8, 0.36, 0.48), nf, 1e-6)
    p.time_approx = time
    p.precision = prec
    ng = mfmhd.num_ghost_cells(p)
    lo = (-ng,) * 3
    U = uniform_lism(p, (n + 2 * ng,) * 3)
    ctx = mfmhd.Context(p)
    g = gpu_update(ctx, U, lo, (0, 0, 0), (n - 1,) * 3, 0.02, 1.0 / n)
    ctx.close()
    assert g["rc"] == 0
    a = valid_view(g["U"], lo, (0, 0, 0), (n - 1,) * 3)
    b = valid_view(U, lo, (0, 0, 0), (n - 1,) * 3)
    groups = [[0], [1, 2, 3], [1, 2, 3], [1, 2, 3], [4], [5, 6, 7], [5, 6, 7], [5, 6, 7]]
    if nf == 2:
        groups += [[8], [9, 10, 11], [9, 10, 11], [9, 10, 11], [12]]
    err = np.array([np.abs(a[c] - b[c]).max() / np.abs(b[gr]).max() for c, gr in enumerate(groups)])
    assert err.max() <= tol, err
