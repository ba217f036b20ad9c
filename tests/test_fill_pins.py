"""Pins of the ghost-fill oracle (exchange + PiecewiseLinearFillPatchMHDAM)
and the host side of the fill plans, on the CPU.

The oracle restates LevelData::exchange and
srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp / FineInterpMHDAMF.ChF.  What
fixes it here, independently of its own code:
  - exchange: every ghost cell that lies in a box of the level (or in a
    periodic image) holds that box's value of a known global function, the
    others are untouched;
  - interpolation: the set of fine cells written equals the definition
    (ghost cells whose coarse parent lies under no coarsened fine box), data
    linear in space is reproduced exactly (centred and one-sided slopes,
    coefficient signs, the limiter leaving linear data alone), the time
    coefficient is linear, and a hand-computed limited case;
  - interp_divB_0: hand cases, and the corrected slopes are divergence-free.
"""
import numpy as np
import pytest

import mfmhd
from fill_layouts import (GHOST, fab_shape, fine_fabs, grown, interp_cells,
                          level_params, linear_data, owner, split)

N3 = (32, 32, 32)
FINE3 = [((8, 8, 8), (15, 15, 15)), ((16, 8, 8), (23, 15, 15)),
         ((8, 16, 8), (15, 23, 15)), ((0, 0, 24), (7, 7, 31))]
N2 = (48, 32, 1)
FINE2 = [((8, 8, 0), (23, 15, 0)), ((24, 8, 0), (31, 23, 0)),
         ((40, 0, 0), (47, 7, 0))]


def coarse_level(n, r, parts, dims):
    nc_ = tuple(n[d] // r if d < dims else 1 for d in range(3))
    return split(nc_, parts, dims)


@pytest.mark.parametrize("dims,periodic", [(3, (0, 0, 0)), (3, (1, 1, 1)),
                                           (3, (1, 0, 1)), (2, (1, 0, 0))])
def test_exchange_copies_owner_values(oracle_mod, dims, periodic):
    n, boxes = (N3, FINE3) if dims == 3 else (N2, FINE2)
    p = level_params(n, dims, periodic)
    nc, ng = 3, 4

    def f(cell, c):
        return 1000.0 * c + cell[0] + 37.0 * cell[1] + 1031.0 * cell[2]

    fabs, los = [], []
    for b in boxes:
        g = grown(b, ng, dims)
        a = np.full(fab_shape(g, nc), GHOST)
        for k in range(b[0][2], b[1][2] + 1):
            for j in range(b[0][1], b[1][1] + 1):
                for i in range(b[0][0], b[1][0] + 1):
                    for c in range(nc):
                        a[c, k - g[0][2], j - g[0][1], i - g[0][0]] = f((i, j, k), c)
        fabs.append(a)
        los.append(g[0])
    before = [a.copy() for a in fabs]
    oracle_mod.level_exchange(p, boxes, fabs, los, ng)
    written = 0
    for a, a0, lo in zip(fabs, before, los):
        nz, ny, nx = a.shape[1:]
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    cell = (lo[0] + i, lo[1] + j, lo[2] + k)
                    if not np.isnan(a0[0, k, j, i]):
                        assert np.array_equal(a[:, k, j, i], a0[:, k, j, i])
                        continue
                    own = owner(boxes, cell, n, periodic, dims)
                    if own is None:
                        assert np.isnan(a[:, k, j, i]).all(), cell
                    else:
                        want = [f(own[1], c) for c in range(nc)]
                        assert np.array_equal(a[:, k, j, i], want), cell
                        written += 1
    assert written > 0
    # the product's plan writes the same number of ghost cells
    plan = mfmhd.FillPlan(p, boxes, interp_radius=ng)
    assert plan.info()["exchange_cells"] == written


def _fill(oracle_mod, p, boxes, ng, r, dxc, cb, cold, cnew, alpha, nc, dims, seed=0):
    dxf = dxc / r
    fabs, los = fine_fabs(boxes, nc, ng, dims, dxf, seed)
    before = [a.copy() for a in fabs]
    clos = [b[0] for b in cb]
    oracle_mod.fill_interp(p, boxes, fabs, los, cb, cold, cnew, clos, r, ng, dxc, alpha)
    return fabs, before, los


@pytest.mark.parametrize("dims,r,ng", [(3, 2, 4), (3, 4, 4), (3, 2, 3), (2, 2, 4)])
def test_interp_cells_and_linear_data(oracle_mod, dims, r, ng):
    n, boxes = (N3, FINE3) if dims == 3 else (N2, FINE2)
    periodic = (0, 0, 0)
    p = level_params(n, dims, periodic)
    nc = 4
    dxc = 1.0 / 8.0
    cb = coarse_level(n, r, (2, 2, 2) if dims == 3 else (3, 2, 1), dims)
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, nc)
    b = rng.uniform(-2, 2, (nc, 3))
    if dims == 2:
        b[:, 2] = 0.0
    cnew = [linear_data(c, nc, dxc, (a, b)) for c in cb]
    fabs, before, los = _fill(oracle_mod, p, boxes, ng, r, dxc, cb, None, cnew, 1.0, nc, dims)
    dxf = dxc / r
    total = 0
    for i, (F, F0, lo) in enumerate(zip(fabs, before, los)):
        want = interp_cells(boxes, i, ng, r, n, periodic, dims)
        changed = set()
        nz, ny, nx = F.shape[1:]
        for k in range(nz):
            for j in range(ny):
                for ii in range(nx):
                    if np.isnan(F0[0, k, j, ii]) and not np.isnan(F[0, k, j, ii]):
                        changed.add((lo[0] + ii, lo[1] + j, lo[2] + k))
        assert changed == want, (i, len(changed), len(want))
        total += len(want)
        for cell in want:
            k, j, ii = (cell[2] - lo[2], cell[1] - lo[1], cell[0] - lo[0])
            x = [dxf * (cell[d] + 0.5) for d in range(3)]
            exact = a + b @ np.array(x)
            assert np.allclose(F[:, k, j, ii], exact, rtol=0, atol=1e-12), cell
    plan = mfmhd.FillPlan(p, boxes, cb, r, ng, dxc)
    assert plan.info()["interp_fine_cells"] == total


def test_interp_time_coefficient(oracle_mod):
    p = level_params(N3)
    r, ng, nc, dxc = 2, 4, 2, 1.0 / 16.0
    cb = coarse_level(N3, r, (2, 2, 2), 3)
    rng = np.random.default_rng(3)
    A = (rng.uniform(-1, 1, nc), rng.uniform(-2, 2, (nc, 3)))
    B = (rng.uniform(-1, 1, nc), rng.uniform(-2, 2, (nc, 3)))
    old = [linear_data(c, nc, dxc, A) for c in cb]
    new = [linear_data(c, nc, dxc, B) for c in cb]
    dxf = dxc / r

    def check(alpha, cold, cnew, coef):
        fabs, before, los = _fill(oracle_mod, p, FINE3, ng, r, dxc, cb, cold, cnew, alpha, nc, 3)
        for i, (F, lo) in enumerate(zip(fabs, los)):
            for cell in interp_cells(FINE3, i, ng, r, N3, (0, 0, 0), 3):
                x = np.array([dxf * (cell[d] + 0.5) for d in range(3)])
                exact = coef[0] + coef[1] @ x
                got = F[:, cell[2] - lo[2], cell[1] - lo[1], cell[0] - lo[0]]
                assert np.allclose(got, exact, rtol=0, atol=1e-12), (alpha, cell)

    mix = (0.3 * B[0] + 0.7 * A[0], 0.3 * B[1] + 0.7 * A[1])
    check(0.3, old, new, mix)
    check(0.0, old, new, A)
    check(1.0, old, new, B)
    check(0.3, None, new, B)     # old data absent: new only
    check(0.3, old, None, A)     # new data absent: old only


def test_interp_limiter_hand_case(oracle_mod):
    """Coarse data varying along x only: g = 0 (ic <= 2), 1 (ic = 3), 5
    (ic >= 4).  Fine box [8, 15]^3, r = 2, dx_c = 1/16.  Parent ic = 3 has
    centred slope (5 - 0)/(2 dx_c); its neighbourhood spans [0, 5], so
    eta = min(5 - 1, 1 - 0) / (|slope| dx_c / 2) = 1 / 1.25 and the limited
    slope is 2/dx_c: children at x_c -+ dx_c/4 get 0.5 and 1.5.  Parent
    ic = 2 sits at a flat minimum (eta = 0): children 0.  The x-high side is
    flat: children 5."""
    p = level_params(N3)
    r, ng, nc, dxc = 2, 4, 1, 1.0 / 16.0
    cb = coarse_level(N3, r, (2, 2, 2), 3)
    box = [FINE3[0]]

    def g(ic):
        return 0.0 if ic <= 2 else (1.0 if ic == 3 else 5.0)

    cnew = []
    for lo, hi in cb:
        a = np.zeros(fab_shape((lo, hi), nc))
        for i in range(lo[0], hi[0] + 1):
            a[0, :, :, i - lo[0]] = g(i)
        cnew.append(a)
    fabs, before, los = _fill(oracle_mod, p, box, ng, r, dxc, cb, None, cnew, 1.0, nc, 3)
    F, lo = fabs[0], los[0]
    want = {4: 0.0, 5: 0.0, 6: 0.5, 7: 1.5, 16: 5.0, 17: 5.0, 18: 5.0, 19: 5.0}
    for fi, v in want.items():
        # a ghost cell straight across the x face (y, z inside the box)
        got = F[0, 12 - lo[2], 12 - lo[1], fi - lo[0]]
        assert abs(got - v) <= 1e-12, (fi, got, v)


def test_divb_correction_hand_cases(oracle_mod):
    assert np.allclose(oracle_mod.interp_divb([3.0, 0.0, 0.0]), [0, 0, 0], atol=1e-15)
    assert np.allclose(oracle_mod.interp_divb([2.0, -1.0, 0.0]), [1, -1, 0], atol=1e-15)
    assert np.allclose(oracle_mod.interp_divb([1.0, 1.0, 1.0]), [0, 0, 0], atol=1e-15)
    assert np.allclose(oracle_mod.interp_divb([2.0, 1.0, 0.0], 2), [0, 0, 0], atol=1e-15)
    # |div B| / 3 below 1e-8: untouched
    s = [1e-9, 2e-9, -1e-9]
    assert np.array_equal(oracle_mod.interp_divb(s), s)


def test_divb_correction_removes_divergence(oracle_mod):
    rng = np.random.default_rng(11)
    for _ in range(2000):
        s = rng.standard_normal(3) * 10.0 ** rng.integers(-3, 3)
        if abs(s.sum()) / 3 <= 1e-8:
            continue
        t = oracle_mod.interp_divb(s)
        assert abs(t.sum()) <= 1e-12 * np.abs(s).sum(), (s, t)
        # each slope moves toward cancelling the divergence, never past zero
        # of its own sign change beyond the divergence
        assert np.all(np.abs(t - s) <= abs(s.sum()) + 1e-12 * np.abs(s).sum())
    for _ in range(500):
        s = rng.standard_normal(3)
        s[2] = 0.0
        if abs(s[:2].sum()) / 2 <= 1e-8:
            continue
        t = oracle_mod.interp_divb(s, 2)
        assert abs(t[:2].sum()) <= 1e-12 * np.abs(s).sum()


def test_fill_time_coef_c_abi():
    """mfm_fill_time_coef restates srcIdeal/AMRLevelIdealMHD.cpp:710-745."""
    a = mfmhd.fill_time_coef(time=1.25, t_coarse_new=2.0, dt_coarse=1.0, dt_fine=0.5,
                             ref_ratio=2)
    assert a == 0.25
    # within 0.04 dt_fine / ref_ratio of 0 or 1: snapped
    assert mfmhd.fill_time_coef(1.0 + 1e-3, 2.0, 1.0, 0.5, 2) == 0.0
    assert mfmhd.fill_time_coef(2.0 - 1e-3, 2.0, 1.0, 0.5, 2) == 1.0
    with pytest.raises(mfmhd.MfmError):
        mfmhd.fill_time_coef(0.5, 2.0, 1.0, 0.5, 2)
    with pytest.raises(mfmhd.MfmError):
        mfmhd.fill_time_coef(2.5, 2.0, 1.0, 0.5, 2)


def test_plan_rejects_improper_nesting():
    """A coarse level that leaves a hole under a fine ghost cell cannot feed
    the interpolation: the plan refuses it instead of reading garbage."""
    p = level_params(N3)
    cb = [b for b in coarse_level(N3, 2, (2, 2, 2), 3) if b[0] != (0, 0, 0)]
    with pytest.raises(mfmhd.MfmError) as e:
        mfmhd.FillPlan(p, [FINE3[0]], cb, 2, 4, 1.0 / 16)
    assert e.value.code == mfmhd.ERR_UNSUPPORTED
    # overlapping fine boxes are a shape error
    with pytest.raises(mfmhd.MfmError) as e:
        mfmhd.FillPlan(p, [FINE3[0], ((12, 8, 8), (19, 15, 15))], interp_radius=4)
    assert e.value.code == mfmhd.ERR_SHAPE
