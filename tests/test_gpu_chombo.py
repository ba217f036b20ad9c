"""A level read from a Chombo HDF5 checkpoint (mfm_chombo_*) goes through the
device update unchanged: boxes and FABs from the file, one
mfm_level_update_dev call, each box equal to the oracle's update of the
same FAB (FP64, <= 1e-12)."""
import numpy as np
import pytest

import mfmhd
from chombo_h5 import write_chombo
from helpers import GpuPatch, rel_err_per_comp, valid_view
from test_gpu_parity import helio_setup

pytestmark = pytest.mark.gpu


def test_restart_level_through_the_device(oracle_mod, tmp_path):
    n, ng = 16, 4
    p, dx = helio_setup(oracle_mod, n, 4)
    nc = mfmhd.num_components(p)
    lo, hi = (-ng,) * 3, (n - 1 + ng,) * 3
    shape = (nc, n + 2 * ng, n + 2 * ng, n + 2 * ng)
    U = oracle_mod.init_problem(p, shape, lo, lo, hi, dx)
    boxes = [((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 7)), ((8, 0, 8), (15, 15, 15))]
    fabs = []
    for blo, bhi in boxes:
        sl = tuple(slice(blo[d] - ng - lo[d], bhi[d] + ng - lo[d] + 1) for d in (2, 1, 0))
        fabs.append(U[(slice(None),) + sl].copy())
    path = str(tmp_path / "chk0100.3d.hdf5")
    names = ["c%d" % c for c in range(nc)]
    write_chombo(path, [{"dx": dx, "dt": 0.0, "ref_ratio": 2, "domain": ((0, 0, 0), (n - 1,) * 3),
                         "boxes": boxes, "fabs": fabs}], names, time=0.0, ghost=(ng,) * 3)
    f = mfmhd.ChomboFile(path)
    assert f.boxes(0) == boxes
    read = [f.read_box(0, b) for b in range(len(boxes))]
    f.close()
    dtn = min(oracle_mod.compute_dt(p, a, glo, b[0], b[1], dx)[0] for (a, glo), b in zip(read, boxes))
    dt = 0.2 * dtn
    ctx = mfmhd.Context(p)
    pats = [GpuPatch(a, glo) for a, glo in read]
    ctx.update_level([(g.fab, mfmhd.Box.of(*b)) for g, b in zip(pats, boxes)], dt, dx)
    mfmhd.synchronize()
    for (a, glo), g, b in zip(read, pats, boxes):
        o = oracle_mod.patch_update(p, a, glo, b[0], b[1], dt, dx)
        assert o["rc"] == 0
        err = rel_err_per_comp(valid_view(g.numpy(), glo, *b), valid_view(o["U"], glo, *b)).max()
        print("box", b, "err", err, flush=True)
        assert err <= 1e-12, (b, err)
    ctx.close()
