"""mfm_patch_update_host: the update on host FABs, staged through HBM by the
library (the entry point a caller without device buffers binds in place of
PatchMHDAM::updateState, src/PatchMHDAM.H:252-260).  It gives the device
entry point's results bit for bit (state, ghosts, face fluxes, divB, dt and
its cell), and its argument errors leave the host state untouched."""
import numpy as np
import pytest

import mfmhd
from helpers import grid_shape, gpu_update
from test_gpu_parity import helio_setup

pytestmark = pytest.mark.gpu


def face_boxes(vlo, vhi):
    out = []
    for d in range(3):
        hi = list(vhi)
        hi[d] += 1
        out.append((tuple(vlo), tuple(hi)))
    return out


@pytest.mark.parametrize("prec", [mfmhd.PREC_FP64, mfmhd.PREC_FP32_FLUX])
def test_host_path_matches_device_path(oracle_mod, prec):
    print("host path prec", prec, flush=True)
    n = 16
    p, dx = helio_setup(oracle_mod, n, 4, prec)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U0 = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    ctx = mfmhd.Context(p)
    dtn, _ = oracle_mod.compute_dt(p, U0, lo, vlo, vhi, dx)
    dt = 0.5 * dtn
    g = gpu_update(ctx, U0, lo, vlo, vhi, dt, dx, 0.0, want_F=True, want_divB=True)
    assert g["rc"] == 0
    Uh = U0.copy()
    F = []
    for flo, fhi in face_boxes(vlo, vhi):
        F.append((np.zeros((nc, fhi[2] - flo[2] + 1, fhi[1] - flo[1] + 1,
                            fhi[0] - flo[0] + 1)), flo))
    D = (np.zeros((1, n, n, n)), vlo)
    nd, cell = ctx.update_state_host(Uh, lo, mfmhd.Box.of(vlo, vhi), dt, dx, 0.0, F=F,
                                     divB=D)
    assert np.array_equal(Uh, g["U"])
    for d in range(3):
        assert np.array_equal(F[d][0], g["F"][d]), d
    assert np.array_equal(D[0], g["divB"])
    assert nd == g["dt"] and cell == g["cell"]
    # a second call on the host result continues the same run
    g2 = gpu_update(ctx, g["U"], lo, vlo, vhi, 0.5 * nd, dx, dt)
    nd2, cell2 = ctx.update_state_host(Uh, lo, mfmhd.Box.of(vlo, vhi), 0.5 * nd, dx, dt)
    assert np.array_equal(Uh, g2["U"])
    assert nd2 == g2["dt"] and cell2 == g2["cell"]
    ctx.close()


def test_host_path_errors_leave_state_untouched(oracle_mod):
    n = 8
    p, dx = helio_setup(oracle_mod, n, 2, mfmhd.PREC_FP64)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    vlo, vhi = (0, 0, 0), (n - 1,) * 3
    lo, hi, shape = grid_shape(nc, vlo, vhi, ng)
    U0 = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    ctx = mfmhd.Context(p)
    Uh = U0.copy()
    # the valid box grown by the ghost depth does not fit in U
    with pytest.raises(mfmhd.MfmError) as e:
        ctx.update_state_host(Uh, lo, mfmhd.Box.of((0, 0, 0), (n, n - 1, n - 1)), 1e-3, dx)
    assert e.value.code == mfmhd.ERR_SHAPE
    # a non-positive time step
    with pytest.raises(mfmhd.MfmError) as e:
        ctx.update_state_host(Uh, lo, mfmhd.Box.of(vlo, vhi), -1e-3, dx)
    assert e.value.code == mfmhd.ERR_ARG
    assert np.array_equal(Uh, U0)
    # arrays the binding cannot hand over as FABs are refused before the call
    with pytest.raises(mfmhd.MfmError):
        ctx.update_state_host(Uh.astype(np.float32), lo, mfmhd.Box.of(vlo, vhi), 1e-3, dx)
    with pytest.raises(mfmhd.MfmError):
        ctx.update_state_host(np.asfortranarray(Uh), lo, mfmhd.Box.of(vlo, vhi), 1e-3, dx)
    ctx.close()
