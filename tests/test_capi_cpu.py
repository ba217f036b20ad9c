"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/mfm.h declares, struct layouts agree, the problem setters restate
the reference's nondimensionalisation, and context creation fails cleanly
without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mfmhd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mfm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = mfmhd.load()
    syms = declared_symbols()
    assert len(syms) >= 17
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(mfmhd.EXPORTS) == syms


def test_params_layout():
    lib = mfmhd.load()
    assert lib.mfm_params_sizeof() == C.sizeof(mfmhd.Params)


def test_defaults_follow_reference_deck_defaults():
    p = mfmhd.default_params()
    assert p.flux == 1 and list(p.limiter) == [1] * 8
    assert p.time_approx == mfmhd.TIME_FIRST_ORDER
    assert p.eight_wave == 1 and p.divb_method == 1
    assert p.entropy_fix == 0.3 and p.small_b == 1e-9
    assert mfmhd.num_ghost_cells(p) == 2
    p.time_approx = mfmhd.TIME_RK2
    assert mfmhd.num_ghost_cells(p) == 4
    p.time_approx = mfmhd.TIME_HANCOCK
    assert mfmhd.num_ghost_cells(p) == 3
    for nf, dedner, nc in ((1, 0, 8), (1, 1, 9), (2, 0, 13), (4, 0, 23)):
        p.nfluids, p.dedner = nf, dedner
        assert mfmhd.num_components(p) == nc


def test_riemann_setter():
    s4 = np.sqrt(4 * np.pi)
    L = [1.0, 0, 0, 0, 1.0, 0.75 * s4, s4, 0]
    R = [0.125, 0, 0, 0, 0.1, 0.75 * s4, -s4, 0]
    p = mfmhd.riemann_params(2.0, L, R, 0.5)
    r0, p0 = 0.125, 0.1
    umag2 = 2.0 * p0 / r0
    assert p.problem == mfmhd.PROBLEM_RIEMANN
    assert p.smallr == pytest.approx(1e-8 * r0, rel=1e-15)
    assert p.smallp == pytest.approx(1e-8 * r0 * 0.5 * umag2, rel=1e-14)


def test_helio_setter_nondimensionalisation():
    """Closed forms of SETHELIOS / HeliosphericProblem::input with the
    reference's physical constants (srcIdeal/EosCommon.fh:28-30)."""
    mp, kB = 1.67262171e-24, 1.3806505e-16
    n = 64
    dx = 1.0 / n
    p = mfmhd.helio_params(lismN=0.06, lismV=2.6e6, lismT=6527.0, lismB=3e-6,
                           lismBdir=(0.6, 0.8, 0.0), sunN=7.0, sunV=4.5e7,
                           sunT=51100.0, sunXYZ=(0.5003, 0.49, 0.51), R0=0.1,
                           netN=0.1, dx=dx, num_cells=(n, n, n))
    pref = 0.06 * mp * 2.6e6 ** 2
    lismP = 2 * kB * 0.06 * 6527.0 / pref
    assert p.lism_p == pytest.approx(lismP, rel=1e-14)
    assert p.lism_m == pytest.approx(1 / np.sqrt(5 / 3 * lismP), rel=1e-14)
    b = 3e-6 / np.sqrt(pref)
    assert p.lism_bx == pytest.approx(0.6 * b, rel=1e-14)
    assert p.lism_by == pytest.approx(0.8 * b, rel=1e-14)
    assert p.sun_rho == pytest.approx(7.0 / 0.06, rel=1e-15)
    assert p.sun_v == pytest.approx(4.5e7 / 2.6e6, rel=1e-15)
    assert p.sun_p == pytest.approx(2 * kB * 7.0 * 51100.0 / pref, rel=1e-14)
    assert p.net_rho == pytest.approx(0.1 / 0.06, rel=1e-15)
    assert p.net_p == pytest.approx(kB * 0.1 * 6527.0 / pref, rel=1e-14)
    # sun snapped to the nearest node (HeliosphericProblem.cpp:233-239)
    assert p.sun_xc == pytest.approx(32 * dx) and p.sun_yc == pytest.approx(31 * dx)
    assert p.sun_zc == pytest.approx(33 * dx)
    # floors: smallr = 1e-8, smallp = 1e-8 * 0.5 * (gamma p0 / r0)
    p0 = 1 / (5 / 3 * p.lism_m ** 2)
    assert p.smallr == 1e-8
    assert p.smallp == pytest.approx(1e-8 * 0.5 * 5 / 3 * p0, rel=1e-14)
    assert p.ce_scale_vel == 2.6e6
    assert p.ce_sigma_lindsay == pytest.approx(1e-16 * 1.5e13 * 0.06, rel=1e-15)
    assert p.omega == pytest.approx(2 * np.pi / (25.38 * 24 * 3600), rel=1e-15)


def test_ctx_validation_without_device():
    """Invalid parameters are rejected before any device work; with no GPU
    the library reports MFM_ERR_DEVICE instead of crashing."""
    lib = mfmhd.load()
    h = C.c_void_p()
    p = mfmhd.default_params()
    p.nfluids = 7
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.default_params()
    p.dedner = 1
    p.eight_wave = 1
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_ARG
    p = mfmhd.default_params()
    p.time_approx = mfmhd.TIME_HANCOCK   # Hancock without GLM cleaning only
    p.dedner = 1
    p.eight_wave = 0
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.default_params()
    p.time_approx = 3
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.default_params()            # divB_method 0: the reference reads an unset field
    p.divb_method = 0
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p.divb_method = 3
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_ARG
    p = mfmhd.default_params()            # Cartesian coordinates only
    p.coord_sys = mfmhd.COORD_AXISYMMETRIC
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.default_params()            # characteristic slopes: FP64 build only
    p.char_reconstruction = 1
    p.precision = mfmhd.PREC_FP32_FLUX
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.default_params()            # ... and a limiter on every plasma variable
    p.char_reconstruction = 1
    p.precision = mfmhd.PREC_FP64
    p.limiter[4] = mfmhd.LIM_NONE
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    p = mfmhd.helio_params(dx=1.0 / 16, num_cells=(16, 16, 16))   # GLM with neutrals:
    p.dedner = 1                                                  # caller ghosts only
    p.eight_wave = 0
    p.nfluids = 2
    assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_UNSUPPORTED
    for flux, rc in ((5, mfmhd.ERR_UNSUPPORTED), (101, mfmhd.ERR_ARG), (103, mfmhd.ERR_ARG),
                     (-1, mfmhd.ERR_ARG), (6, mfmhd.ERR_ARG), (12, mfmhd.ERR_ARG),
                     (20, mfmhd.ERR_ARG), (22, mfmhd.ERR_ARG)):
        # VFRoe, Dedner through the flux number (src/RiemannSolver.cpp:
        # 40-112) and numbers the factory has no case for: refused, not
        # silently mapped
        p = mfmhd.default_params()
        p.flux = flux
        assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == rc, flux
    for flux in (0, 2, 4, 11, 13):
        # GLM cleaning takes flux 1 or 3 (main/MHDAM.cpp:1045), or this
        # library's HLLC (21)
        p = mfmhd.orszag_params(5.0 / 3.0)
        p.dedner = 1
        p.eight_wave = 0
        p.flux = flux
        assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_ARG, flux
    if mfmhd.device_count() == 0:
        for flux in (0, 1, 2, 3, 4, 11, 13, 21):
            p = mfmhd.default_params()
            p.flux = flux
            assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_DEVICE, flux
        for flux in (1, 3, 21):
            p = mfmhd.orszag_params(5.0 / 3.0)
            p.dedner = 1
            p.eight_wave = 0
            p.flux = flux
            assert lib.mfm_ctx_create(C.byref(p), C.byref(h)) == mfmhd.ERR_DEVICE, flux


def test_level_check_refuses_without_a_device():
    """mfm_level_check runs the update's host checks of a patch table with
    no device and without reading data pointers: an oversized FAB (2^30
    cells per component or more; one cell fewer passes), missing ghosts, a wrong ncomp, a valid box
    outside the domain and bad dt / dx are refused with the code the update
    returns, naming the first bad patch; a good table passes."""
    p = mfmhd.orszag_params(5.0 / 3.0)
    ng = mfmhd.num_ghost_cells(p)
    nc = mfmhd.num_components(p)
    ok = (mfmhd.fab(0x1000, (-ng,) * 3, (15 + ng,) * 3, nc), mfmhd.Box.of((0,) * 3, (15,) * 3))
    assert mfmhd.level_check(p, [ok], 0.01, 0.05)[0] == mfmhd.OK
    n = 1024 - 2 * ng                          # a FAB of 1024^3 = 2^30 cells
    big = (mfmhd.fab(0x1000, (-ng,) * 3, (n - 1 + ng,) * 3, nc), mfmhd.Box.of((0,) * 3, (n - 1,) * 3))
    rc, bad, why = mfmhd.level_check(p, [ok, big], 0.01, 0.05)
    assert rc == mfmhd.ERR_SHAPE and bad == 1 and "2^30" in why
    # 1023 x 1024 x 1024 cells, one plane of x fewer: accepted
    fits = (mfmhd.fab(0x1000, (-ng,) * 3, (n - 2 + ng, n - 1 + ng, n - 1 + ng), nc),
            mfmhd.Box.of((0,) * 3, (n - 2, n - 1, n - 1)))
    assert mfmhd.level_check(p, [ok, fits], 0.01, 0.05)[0] == mfmhd.OK
    thin = (mfmhd.fab(0x1000, (-ng + 1,) * 3, (15 + ng,) * 3, nc), mfmhd.Box.of((0,) * 3, (15,) * 3))
    rc, bad, why = mfmhd.level_check(p, [thin, ok], 0.01, 0.05)
    assert rc == mfmhd.ERR_SHAPE and bad == 0 and "numGhostCells" in why
    wrong = (mfmhd.fab(0x1000, (-ng,) * 3, (15 + ng,) * 3, nc + 1), ok[1])
    assert mfmhd.level_check(p, [wrong], 0.01, 0.05)[:2] == (mfmhd.ERR_SHAPE, 0)
    assert mfmhd.level_check(p, [ok], 0.0, 0.05)[0] == mfmhd.ERR_ARG
    assert mfmhd.level_check(p, [ok], 0.01, -1.0)[0] == mfmhd.ERR_ARG
    q = mfmhd.helio_params(dx=1.0 / 16, num_cells=(16, 16, 16))
    out = (mfmhd.fab(0x1000, (-ng,) * 3, (19 + ng,) * 3, 8), mfmhd.Box.of((4,) * 3, (19,) * 3))
    q.nfluids = 1
    rc, bad, why = mfmhd.level_check(q, [out], 0.01, 1.0 / 16)
    assert rc == mfmhd.ERR_SHAPE and "domain" in why


def test_binding_constants_match_the_header():
    """Every MFM_* enumerator of include/mfm.h has the same value in the
    Python binding (MFM_LIM_MINMOD -> mfmhd.LIM_MINMOD, MFM_KT_* ->
    KT_*, MFM_REQ_* -> REQ_*, MFM_OK / MFM_ERR_* -> OK / ERR_*)."""
    src = open(os.path.join(ROOT, "include", "mfm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    found = 0
    for body in re.findall(r"enum\s*\{(.*?)\}", src, flags=re.S):
        for name, val in re.findall(r"\b(MFM_[A-Z0-9_]+)\s*=\s*(-?\d+)", body):
            py = name[len("MFM_"):]
            assert hasattr(mfmhd, py), name
            assert getattr(mfmhd, py) == int(val), (name, getattr(mfmhd, py), val)
            found += 1
    assert found >= 40
    abi = re.search(r"#define MFM_ABI_VERSION (\d+)", src)
    assert int(abi.group(1)) == mfmhd.ABI_VERSION
