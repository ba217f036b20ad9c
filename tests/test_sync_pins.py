"""Pins of the oracle's coarse-fine synchronisation (flux register, reflux,
average down) on the CPU.

What fixes it, independently of its code:
  - conservation across levels: a coarse step and r fine steps with
    arbitrary face fluxes, then reflux (scale -1, as the reference calls it)
    and the average down, leave the composite integral of every component
    unchanged to round-off on a periodic domain; any wrong sign, side, face
    choice or fine-face average breaks it;
  - without reflux the composite integral does change (the test has teeth);
  - the average of data linear in space is its value at the coarse centre;
  - the register faces counted by the product's plan equal an independent
    count of the coarse-fine interface.
"""
import numpy as np
import pytest

import mfmhd
from fill_layouts import centres, level_params, linear_data, split
from sync_layouts import (COARSE, FINE, N_FINE, R, cell_fabs, divergence, face_fabs,
                          params)

NC = 3


def _composite(Uc, Uf, dxc, dxf):
    """Integral over the composite grid: uncovered coarse cells + fine cells."""
    cov = np.zeros(Uc.shape[1:], dtype=bool)
    for lo, hi in FINE:
        cov[lo[2] // R:hi[2] // R + 1, lo[1] // R:hi[1] // R + 1, lo[0] // R:hi[0] // R + 1] = True
    tot = Uc[:, ~cov].sum(axis=1) * dxc ** 3
    for lo, hi in FINE:
        tot = tot + Uf[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].sum(axis=(1, 2, 3)) * dxf ** 3
    return tot


def _run_cycle(oracle_mod, reflux=True, seed=0):
    rng = np.random.default_rng(seed)
    p = params()
    nc_, nf_ = 16, 32
    dxc, dxf = 1.0 / nc_, 1.0 / nf_
    dtc = 0.01
    dtf = dtc / R
    Uc = rng.uniform(1, 2, (NC, nc_, nc_, nc_))
    Uf = rng.uniform(1, 2, (NC, nf_, nf_, nf_))
    # consistent start: covered coarse cells hold the mean of their fine cells
    reg = oracle_mod.FluxRegister(p, FINE, COARSE, R, NC)
    cfab, clos = cell_fabs(COARSE, Uc)
    ffab, flos = cell_fabs(FINE, Uf)
    reg.average_down(cfab, clos, ffab, flos, dxf)
    for (lo, hi), a in zip(COARSE, cfab):
        Uc[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a
    before = _composite(Uc, Uf, dxc, dxf)
    # coarse step with arbitrary face fluxes
    Gc = [rng.standard_normal((NC, nc_, nc_, nc_)) for _ in range(3)]
    Uc = Uc - dtc * divergence(Gc, dxc)
    fc, fclos = face_fabs(COARSE, Gc)
    reg.increment_coarse(fc, fclos, dtc)
    # r fine steps, each with its own fluxes
    for _ in range(R):
        Gf = [rng.standard_normal((NC, nf_, nf_, nf_)) for _ in range(3)]
        Uf = Uf - dtf * divergence(Gf, dxf)
        ff, fflos = face_fabs(FINE, Gf)
        reg.increment_fine(ff, fflos, dtf)
    cfab, clos = cell_fabs(COARSE, Uc)
    ffab, flos = cell_fabs(FINE, Uf)
    if reflux:
        reg.reflux(cfab, clos, -1.0, dxc)
    reg.average_down(cfab, clos, ffab, flos, dxf)
    for (lo, hi), a in zip(COARSE, cfab):
        Uc[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a
    return before, _composite(Uc, Uf, dxc, dxf)


def test_reflux_conserves_across_levels(oracle_mod):
    for seed in range(3):
        before, after = _run_cycle(oracle_mod, True, seed)
        assert np.allclose(after, before, rtol=1e-13, atol=0), (before, after)


def test_without_reflux_conservation_fails(oracle_mod):
    before, after = _run_cycle(oracle_mod, False, 0)
    assert not np.allclose(after, before, rtol=1e-9, atol=0)


def test_average_down_of_linear_data(oracle_mod):
    p = params()
    dxf, dxc = 1.0 / 32, 1.0 / 16
    rng = np.random.default_rng(5)
    a = rng.uniform(-1, 1, NC)
    b = rng.uniform(-2, 2, (NC, 3))
    ffab = [linear_data(bx, NC, dxf, (a, b)) for bx in FINE]
    flos = [bx[0] for bx in FINE]
    cfab = [np.full((NC, 8, 8, 8), np.nan) for _ in COARSE]
    clos = [bx[0] for bx in COARSE]
    oracle_mod.FluxRegister(p, FINE, COARSE, R, NC).average_down(cfab, clos, ffab, flos, dxf)
    seen = 0
    for (lo, hi), c in zip(COARSE, cfab):
        x, y, z = centres((lo, hi), dxc)
        want = np.stack([a[q] + b[q][0] * x + b[q][1] * y + b[q][2] * z for q in range(NC)])
        m = ~np.isnan(c[0])
        seen += int(m.sum())
        assert np.allclose(c[:, m], want[:, m], rtol=0, atol=1e-13)
    assert seen == 3 * 64   # every coarse cell under the three fine boxes


def _interface_faces():
    """Coarse faces between a cell under no fine box and a cell under one
    (periodic), counted directly on the coarse grid."""
    n = 16
    cov = np.zeros((n, n, n), dtype=bool)
    for lo, hi in FINE:
        cov[lo[2] // R:hi[2] // R + 1, lo[1] // R:hi[1] // R + 1, lo[0] // R:hi[0] // R + 1] = True
    count = 0
    for ax in (1, 2, 0):
        for s in (1, -1):
            count += int(np.sum(~cov & np.roll(cov, s, axis=ax)))
    return count


def test_register_faces_match_product_plan():
    p = params()
    reg = mfmhd.FluxRegister(p, FINE, COARSE, R)
    info = reg.info()
    assert info["faces"] == _interface_faces()
    assert info["average_cells"] == 3 * 64


def test_register_plan_errors():
    p = params()
    with pytest.raises(mfmhd.MfmError) as e:   # not coarsenable by 2
        mfmhd.FluxRegister(p, [((9, 8, 8), (15, 15, 15))], COARSE, R)
    assert e.value.code == mfmhd.ERR_SHAPE
    with pytest.raises(mfmhd.MfmError) as e:   # hole in the coarse level
        mfmhd.FluxRegister(p, FINE, COARSE[1:], R)
    assert e.value.code == mfmhd.ERR_UNSUPPORTED
