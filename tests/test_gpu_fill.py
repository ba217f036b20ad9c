"""Device ghost fill of a level (mfm_level_exchange_dev,
mfm_level_fill_interp_dev) against the CPU oracle's restatement of
LevelData::exchange and PiecewiseLinearFillPatchMHDAM::fillInterp.

Both are FP64 with the same operation order (no FMA contraction), so the
bar is bit-for-bit equality; NaN marks ghost cells nothing writes, and must
stay NaN on both sides.
"""
import numpy as np
import pytest

import mfmhd
from fill_layouts import (fab_shape, fine_fabs, level_params, rough_data,
                          split)

pytestmark = pytest.mark.gpu

N3 = (32, 32, 32)
FINE3 = [((8, 8, 8), (15, 15, 15)), ((16, 8, 8), (23, 15, 15)),
         ((8, 16, 8), (15, 23, 15)), ((0, 0, 24), (7, 7, 31))]
N2 = (48, 32, 1)
FINE2 = [((8, 8, 0), (23, 15, 0)), ((24, 8, 0), (31, 23, 0)),
         ((40, 0, 0), (47, 7, 0))]


def to_dev(arrs, los, nc):
    devs, fabs = [], []
    for a, lo in zip(arrs, los):
        d = mfmhd.DeviceArray(a.shape, data=a)
        hi = (lo[0] + a.shape[3] - 1, lo[1] + a.shape[2] - 1, lo[2] + a.shape[1] - 1)
        devs.append(d)
        fabs.append(mfmhd.fab(d.ptr, lo, hi, nc))
    return devs, fabs


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("dims,periodic", [(3, (0, 0, 0)), (3, (1, 1, 1)), (2, (1, 0, 0))])
def test_exchange_matches_oracle(oracle_mod, dims, periodic):
    n, boxes = (N3, FINE3) if dims == 3 else (N2, FINE2)
    print("exchange", dims, periodic, flush=True)
    p = level_params(n, dims, periodic)
    nc, ng = mfmhd.num_components(p), 4
    arrs, los = fine_fabs(boxes, nc, ng, dims, 1.0 / n[0], seed=5)
    devs, fabs = to_dev(arrs, los, nc)
    ctx = mfmhd.Context(p)
    plan = mfmhd.FillPlan(p, boxes, interp_radius=ng)
    ctx.exchange(plan, fabs)
    oracle_mod.level_exchange(p, boxes, arrs, los, ng)
    for d, a in zip(devs, arrs):
        assert same(d.numpy(), a)


CASES = [
    # id, dims, periodic, r, ng, alpha, old, new, nfluids, dedner
    ("3d-r2-g4-mix", 3, (0, 0, 0), 2, 4, 0.37, True, True, 1, False),
    ("3d-per-r2-g4-mix", 3, (1, 1, 1), 2, 4, 0.37, True, True, 1, False),
    ("3d-r4-g4-new", 3, (0, 0, 0), 4, 4, 1.0, False, True, 1, False),
    ("3d-r2-g3-old", 3, (0, 1, 0), 2, 3, 0.0, True, False, 1, False),
    ("3d-r2-g2-mix-4f", 3, (0, 0, 0), 2, 2, 0.61, True, True, 4, False),
    ("3d-dedner-mix", 3, (0, 0, 0), 2, 4, 0.37, True, True, 1, True),
    ("2d-perx-dedner-mix", 2, (1, 0, 0), 2, 4, 0.5, True, True, 1, True),
]


@pytest.mark.parametrize("name,dims,periodic,r,ng,alpha,use_old,use_new,nf,dedner",
                         CASES, ids=[c[0] for c in CASES])
def test_fill_interp_matches_oracle(oracle_mod, name, dims, periodic, r, ng, alpha,
                                    use_old, use_new, nf, dedner):
    print("fill_interp", name, flush=True)
    n, boxes = (N3, FINE3) if dims == 3 else (N2, FINE2)
    p = level_params(n, dims, periodic, nfluids=nf, dedner=dedner)
    nc = mfmhd.num_components(p)
    dxc = 1.0 / (n[0] // r)
    cn_ = tuple(n[d] // r if d < dims else 1 for d in range(3))
    cb = split(cn_, (2, 2, 2) if dims == 3 else (3, 2, 1), dims)
    old = [rough_data(b, nc, 900 + i, dxc) for i, b in enumerate(cb)] if use_old else None
    new = [rough_data(b, nc, 700 + i, dxc) for i, b in enumerate(cb)] if use_new else None
    clos = [b[0] for b in cb]
    arrs, los = fine_fabs(boxes, nc, ng, dims, dxc / r, seed=3)
    devs, fabs = to_dev(arrs, los, nc)
    dold, fold = to_dev(old, clos, nc) if old is not None else (None, None)
    dnew, fnew = to_dev(new, clos, nc) if new is not None else (None, None)
    ctx = mfmhd.Context(p)
    plan = mfmhd.FillPlan(p, boxes, cb, r, ng, dxc)
    ctx.fill_interp(plan, fabs, fold, fnew, alpha)
    oracle_mod.fill_interp(p, boxes, arrs, los, cb, old, new, clos, r, ng, dxc, alpha)
    nfill = 0
    for d, a in zip(devs, arrs):
        g = d.numpy()
        assert same(g, a), np.nanmax(np.abs(g - a))
        nfill += int(np.count_nonzero(~np.isnan(a[0])))
    assert nfill > 0


def test_fill_then_exchange_twice(oracle_mod):
    """The step's order (fillInterp, then exchange) twice on one plan with
    new data each time: per-call tables are restaged correctly."""
    p = level_params(N3, 3, (0, 0, 0), nfluids=2)
    nc, r, ng = mfmhd.num_components(p), 2, 4
    dxc = 1.0 / 16
    cb = split((16, 16, 16), (2, 2, 2), 3)
    clos = [b[0] for b in cb]
    ctx = mfmhd.Context(p)
    plan = mfmhd.FillPlan(p, FINE3, cb, r, ng, dxc)
    for it in range(2):
        print("fill+exchange round", it, flush=True)
        old = [rough_data(b, nc, 40 * it + i, dxc) for i, b in enumerate(cb)]
        new = [rough_data(b, nc, 40 * it + 20 + i, dxc) for i, b in enumerate(cb)]
        arrs, los = fine_fabs(FINE3, nc, ng, 3, dxc / r, seed=11 + it)
        devs, fabs = to_dev(arrs, los, nc)
        dold, fold = to_dev(old, clos, nc)
        dnew, fnew = to_dev(new, clos, nc)
        alpha = 0.25 + 0.5 * it
        ctx.fill_interp(plan, fabs, fold, fnew, alpha)
        ctx.exchange(plan, fabs)
        oracle_mod.fill_interp(p, FINE3, arrs, los, cb, old, new, clos, r, ng, dxc, alpha)
        oracle_mod.level_exchange(p, FINE3, arrs, los, ng)
        for d, a in zip(devs, arrs):
            assert same(d.numpy(), a)


def test_fill_shape_errors_leave_state_untouched():
    p = level_params(N3)
    nc, ng = mfmhd.num_components(p), 4
    arrs, los = fine_fabs(FINE3[:1], nc, 2, 3, 1.0 / 32)  # only 2 ghost layers
    devs, fabs = to_dev(arrs, los, nc)
    ctx = mfmhd.Context(p)
    plan = mfmhd.FillPlan(p, FINE3[:1] + [((16, 8, 8), (23, 15, 15))], interp_radius=ng)
    other = mfmhd.DeviceArray(fab_shape(((12, 4, 4), (27, 19, 19)), nc), zero=True)
    with pytest.raises(mfmhd.MfmError) as e:
        ctx.exchange(plan, fabs + [mfmhd.fab(other.ptr, (12, 4, 4), (27, 19, 19), nc)])
    assert e.value.code == mfmhd.ERR_SHAPE
    assert same(devs[0].numpy(), arrs[0])
