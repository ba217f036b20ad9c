"""Coarse-fine synchronisation on the device (mfm_flux_register_*,
mfm_level_average_down_dev) against the oracle, and one whole two-level AMR
cycle on the device (ghost fill, level updates with face fluxes, flux
register, reflux, average down) that must conserve mass across the levels.

The register arithmetic is FP64 in the same order as the oracle's, so the
parity bar is bit-for-bit.
"""
import numpy as np
import pytest

import mfmhd
from fill_layouts import centres, fab_shape, grown, level_params
from sync_layouts import COARSE, FINE, R, cell_fabs, face_fabs, params

pytestmark = pytest.mark.gpu


def to_dev(arrs, los, nc):
    devs, fabs = [], []
    for a, lo in zip(arrs, los):
        d = mfmhd.DeviceArray(a.shape, data=a)
        hi = (lo[0] + a.shape[3] - 1, lo[1] + a.shape[2] - 1, lo[2] + a.shape[1] - 1)
        devs.append(d)
        fabs.append(mfmhd.fab(d.ptr, lo, hi, nc))
    return devs, fabs


def test_register_ops_match_oracle(oracle_mod):
    p = params()
    nc = mfmhd.num_components(p)
    rng = np.random.default_rng(21)
    dxc, dxf, dtc = 1.0 / 16, 1.0 / 32, 0.013
    ctx = mfmhd.Context(p)
    reg = mfmhd.FluxRegister(p, FINE, COARSE, R)
    oreg = oracle_mod.FluxRegister(p, FINE, COARSE, R, nc)
    Uc = rng.uniform(1, 2, (nc, 16, 16, 16))
    Uf = rng.uniform(1, 2, (nc, 32, 32, 32))
    cfab, clos = cell_fabs(COARSE, Uc)
    ffab, flos = cell_fabs(FINE, Uf)
    dUc, fUc = to_dev(cfab, clos, nc)
    dUf, fUf = to_dev(ffab, flos, nc)
    for rnd in range(2):          # a second coarse step re-zeroes the register
        print("register round", rnd, flush=True)
        reg.set_zero(ctx)
        oreg.R[:] = 0.0
        Gc = [rng.standard_normal((nc, 16, 16, 16)) for _ in range(3)]
        fc, fclos = face_fabs(COARSE, Gc)
        dF, fF = to_dev(fc, fclos, nc)
        reg.increment_coarse(ctx, fF, dtc)
        oreg.increment_coarse(fc, fclos, dtc)
        for _ in range(R):
            Gf = [rng.standard_normal((nc, 32, 32, 32)) for _ in range(3)]
            ff, fflos = face_fabs(FINE, Gf)
            dG, fG = to_dev(ff, fflos, nc)
            reg.increment_fine(ctx, fG, dtc / R)
            oreg.increment_fine(ff, fflos, dtc / R)
        reg.reflux(ctx, fUc, -1.0, dxc)
        oreg.reflux(cfab, clos, -1.0, dxc)
        for d, a in zip(dUc, cfab):
            assert np.array_equal(d.numpy(), a)
        reg.average_down(ctx, fUc, fUf, dxf)
        oreg.average_down(cfab, clos, ffab, flos, dxf)
        for d, a in zip(dUc, cfab):
            assert np.array_equal(d.numpy(), a)


def _state(box, nc, dx, ng):
    """Smooth single-fluid MHD state (conserved) on box grown by ng."""
    g = grown(box, ng)
    x, y, z = centres(g, dx)
    t = 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(t * x) * np.cos(t * y) + 0.1 * np.sin(t * z)
    vx, vy, vz = 0.3 * np.sin(t * y), 0.3 * np.sin(t * z), 0.3 * np.sin(t * x)
    bx, by, bz = 0.5 + 0.1 * np.cos(t * z), 0.4 + 0.1 * np.cos(t * x), 0.3 + 0.1 * np.cos(t * y)
    pr = 1.0 + 0.1 * np.cos(t * (x + y))
    U = np.zeros(fab_shape(g, nc))
    U[0] = rho
    U[1], U[2], U[3] = rho * vx, rho * vy, rho * vz
    U[4] = pr / (5.0 / 3.0 - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz) + \
        (bx * bx + by * by + bz * bz) / (8 * np.pi)
    U[5], U[6], U[7] = bx, by, bz
    return U


@pytest.mark.parametrize("time,char", [(mfmhd.TIME_RK2, 0), (mfmhd.TIME_HANCOCK, 1)])
def test_amr_cycle_conserves_mass(time, char):
    """Coarse 16^3 periodic level of 8 boxes, fine level (ratio 2) of three
    boxes, one wrapping around the domain; single-fluid MHD with RK2, and with
    the Orszag-Tang deck's scheme (Hancock, characteristic slopes with
    check_tvd, van Leer).  One coarse step and two fine steps with the device
    ghost fill, then reflux and the average down: the composite mass is
    unchanged to round-off."""
    print("amr cycle", time, char, flush=True)
    ng = 4
    pc = level_params((16, 16, 16), 3, (1, 1, 1))
    pf = level_params((32, 32, 32), 3, (1, 1, 1))
    for p in (pc, pf):
        p.time_approx = time
        p.char_reconstruction = char
        if char:
            p.check_tvd = 1
            for v in range(8):
                p.limiter[v] = mfmhd.LIM_VANLEER
    nc = mfmhd.num_components(pc)
    dxc, dxf = 1.0 / 16, 1.0 / 32
    cctx, fctx = mfmhd.Context(pc), mfmhd.Context(pf)

    def level(boxes, dx):
        devs, fabs, fdevs, ffabs = [], [], [], []
        for b in boxes:
            U = _state(b, nc, dx, ng)
            g = grown(b, ng)
            d = mfmhd.DeviceArray(U.shape, data=U)
            devs.append(d)
            fabs.append(mfmhd.fab(d.ptr, g[0], g[1], nc))
            Fl = []
            for dd in range(3):
                hi = list(b[1])
                hi[dd] += 1
                fd = mfmhd.DeviceArray(fab_shape((b[0], tuple(hi)), nc), zero=True)
                fdevs.append(fd)
                Fl.append(mfmhd.fab(fd.ptr, b[0], hi, nc))
            ffabs.append(Fl)
        return devs, fabs, fdevs, ffabs

    cdev, cfab, cfd, cF = level(COARSE, dxc)
    fdev, ffab, ffd, fF = level(FINE, dxf)
    old = [mfmhd.DeviceArray(d.shape) for d in cdev]
    ofab = [mfmhd.fab(o.ptr, f.box.lo, f.box.hi, nc) for o, f in zip(old, cfab)]
    cplan = mfmhd.FillPlan(pc, COARSE, interp_radius=ng)
    fplan = mfmhd.FillPlan(pf, FINE, COARSE, R, ng, dxc)
    reg = mfmhd.FluxRegister(pf, FINE, COARSE, R)
    cvalid = [mfmhd.Box.of(*b) for b in COARSE]
    fvalid = [mfmhd.Box.of(*b) for b in FINE]

    def mass():
        cov = np.zeros((16, 16, 16), dtype=bool)
        for lo, hi in FINE:
            cov[lo[2] // R:hi[2] // R + 1, lo[1] // R:hi[1] // R + 1, lo[0] // R:hi[0] // R + 1] = True
        tot = 0.0
        for d, (lo, hi) in zip(cdev, COARSE):
            a = d.numpy()[0, ng:-ng, ng:-ng, ng:-ng]
            m = cov[lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            tot += a[~m].sum() * dxc ** 3
        for d in fdev:
            tot += d.numpy()[0, ng:-ng, ng:-ng, ng:-ng].sum() * dxf ** 3
        return tot

    # consistent start: coarse cells under the fine level = mean of the fine
    reg.average_down(cctx, cfab, [f for f in ffab], dxf)
    m0 = mass()
    # coarse step
    cctx.exchange(cplan, cfab)
    dtc = 0.3 * min(cctx.compute_dt(f, v, dxc)[0] for f, v in zip(cfab, cvalid))
    for o, d in zip(old, cdev):
        o.copy_from(d)
    reg.set_zero(cctx)
    cctx.update_level([(f, v, F) for f, v, F in zip(cfab, cvalid, cF)], dtc, dxc, 0.0)
    reg.increment_coarse(cctx, [x for F in cF for x in F], dtc)
    # two fine steps with the coarse data interpolated in time
    dtf = dtc / R
    for s in range(R):
        print("fine step", s, flush=True)
        fctx.fill_interp(fplan, ffab, ofab, cfab, s / R)
        fctx.exchange(fplan, ffab)
        fctx.update_level([(f, v, F) for f, v, F in zip(ffab, fvalid, fF)], dtf, dxf, s * dtf)
        reg.increment_fine(fctx, [x for F in fF for x in F], dtf)
    reg.reflux(cctx, cfab, -1.0, dxc)
    reg.average_down(cctx, cfab, ffab, dxf)
    m1 = mass()
    assert abs(m1 - m0) <= 1e-13 * abs(m0), (m0, m1)
    # and the step did move mass between the levels' cells
    assert any(not np.array_equal(o.numpy(), d.numpy()) for o, d in zip(old, cdev))
