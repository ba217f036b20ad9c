"""Python binding of the MI355X patch integrator (libmfmhd.so, include/mfm.h).

Host-side mirror of the reference's patch-integrator interface
(PatchMHDAM::updateState / computeDt / numGhostCells,
MHDAM-3.2.4/src/PatchMHDAM.H:252-280) over the C ABI, for tests, benchmarks
and Python callers.  Device memory, copies and timing go through the
library's own HIP runtime (DeviceArray, Event), so a process holds one HIP
runtime; every computation runs in the HIP kernels of libmfmhd.so.  Loading
fails loudly if the library has not been built: there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfmhd.so")

# ---- enums (include/mfm.h) ------------------------------------------------
OK, ERR_ARG, ERR_SHAPE, ERR_DEVICE, ERR_NUMERIC, ERR_NOMEM, ERR_UNSUPPORTED = \
    0, -1, -2, -3, -4, -5, -6
FLUX_ROE8, FLUX_RUSANOV, FLUX_ROE8_ROEAVG, FLUX_RUSANOV_ROEAVG = 1, 2, 3, 4
FLUX_DEFAULT, FLUX_ROE7, FLUX_ROE7_ROEAVG, FLUX_HLLC = 0, 11, 13, 21
(LIM_CD, LIM_MINMOD, LIM_SUPERBEE, LIM_VANALBADA, LIM_MC, LIM_HARMONIC,
 LIM_HYPERBEE, LIM_KOREN, LIM_VENKAT, LIM_VANLEER, LIM_WENO3, LIM_WENO3YC,
 LIM_NONE) = range(13)
TIME_FIRST_ORDER, TIME_HANCOCK, TIME_RK2 = 0, 1, 2
COORD_CARTESIAN, COORD_AXISYMMETRIC = 0, 1
PROBLEM_NONE, PROBLEM_RIEMANN, PROBLEM_HELIO, PROBLEM_SWLISM = 0, 1, 2, 3
PREC_FP64, PREC_FP32_FLUX = 0, 1
STAGE_SEPARATE, STAGE_FUSED, STAGE_SUMMED, STAGE_AUTO = 0, 1, 2, 3
ABI_VERSION = 2


class Box(C.Structure):
    _fields_ = [("lo", C.c_int * 3), ("hi", C.c_int * 3)]

    @staticmethod
    def of(lo, hi):
        b = Box()
        for d in range(3):
            b.lo[d] = int(lo[d])
            b.hi[d] = int(hi[d])
        return b

    def shape(self):
        """(nx, ny, nz)"""
        return tuple(self.hi[d] - self.lo[d] + 1 for d in range(3))


class Fab(C.Structure):
    _fields_ = [("data", C.c_void_p), ("box", Box), ("ncomp", C.c_int)]


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int), ("spacedim", C.c_int), ("nfluids", C.c_int),
        ("dedner", C.c_int), ("flux", C.c_int), ("limiter", C.c_int * 8),
        ("check_positivity", C.c_int), ("time_approx", C.c_int),
        ("divb_method", C.c_int), ("eight_wave", C.c_int),
        ("coord_sys", C.c_int), ("precision", C.c_int),
        ("entropy_fix", C.c_double), ("small_b", C.c_double),
        ("gamma", C.c_double), ("smallr", C.c_double), ("smallp", C.c_double),
        ("dedner_ch", C.c_double), ("dedner_cp", C.c_double),
        ("domain_lo", C.c_int * 3), ("domain_hi", C.c_int * 3),
        ("periodic", C.c_int * 3), ("problem", C.c_int),
        ("rp_left", C.c_double * 8), ("rp_right", C.c_double * 8),
        ("rp_start_x", C.c_double),
        ("sun_xc", C.c_double), ("sun_yc", C.c_double), ("sun_zc", C.c_double),
        ("sun_r0", C.c_double),
        ("sun_rho", C.c_double), ("sun_p", C.c_double), ("sun_v", C.c_double),
        ("sun_b", C.c_double), ("sun_tilt", C.c_double),
        ("sun_int_bc_radius", C.c_double), ("init_r", C.c_double),
        ("dir_br_n", C.c_double), ("omega", C.c_double),
        ("lism_v_dim", C.c_double), ("sun_b_monopolar", C.c_int),
        ("lism_ux", C.c_double), ("lism_uy", C.c_double),
        ("lism_uz", C.c_double), ("lism_p", C.c_double), ("lism_m", C.c_double),
        ("lism_t", C.c_double), ("lism_bx", C.c_double),
        ("lism_by", C.c_double), ("lism_bz", C.c_double),
        ("net_rho", C.c_double), ("net_p", C.c_double), ("tmlim", C.c_double),
        ("reg_safe_z_top", C.c_double), ("reg_safe_z_bot", C.c_double),
        ("r0_dt_ignore", C.c_double), ("ce_scale_vel", C.c_double),
        ("ce_sigma_lindsay", C.c_double), ("ce_single_literals", C.c_int),
        ("char_reconstruction", C.c_int), ("check_tvd", C.c_int),
    ]


class PatchRec(C.Structure):
    _fields_ = [("U", Fab), ("valid", Box), ("F", C.POINTER(Fab)),
                ("divB", C.POINTER(Fab))]


EXPORTS = [
    "mfm_params_sizeof", "mfm_params_default", "mfm_params_set_riemann",
    "mfm_params_set_orszag", "mfm_params_set_helio", "mfm_params_set_swlism",
    "mfm_num_components",
    "mfm_num_ghost_cells", "mfm_ctx_create", "mfm_ctx_destroy",
    "mfm_ctx_set_dedner", "mfm_ctx_last_error", "mfm_patch_update_dev",
    "mfm_patch_update_noalloc_dev",
    "mfm_patch_update_host", "mfm_patch_compute_dt_dev",
    "mfm_level_update_dev", "mfm_level_check", "mfm_init_problem_dev",
    "mfm_fill_periodic_dev",
    "mfm_dev_count", "mfm_dev_alloc", "mfm_dev_free", "mfm_dev_memcpy",
    "mfm_dev_memset", "mfm_dev_synchronize", "mfm_event_create",
    "mfm_event_destroy", "mfm_event_record", "mfm_event_elapsed_ms",
    "mfm_ctx_set_timing", "mfm_ctx_get_timing", "mfm_ctx_set_graphs",
    "mfm_ctx_set_walk", "mfm_ctx_get_walk", "mfm_ctx_autotune",
    "mfm_ctx_set_stage_kernel", "mfm_hbm_ceiling",
    "mfm_server_create", "mfm_server_destroy", "mfm_server_last_error",
    "mfm_server_set_stage_kernel",
    "mfm_server_submit", "mfm_server_step", "mfm_server_request_info",
    "mfm_server_fetch", "mfm_server_release", "mfm_server_stats",
    "mfm_server_request_error", "mfm_fill_plan_create",
    "mfm_fill_plan_destroy", "mfm_fill_plan_last_error", "mfm_fill_plan_info",
    "mfm_level_exchange_dev", "mfm_level_fill_interp_dev", "mfm_fill_time_coef",
    "mfm_flux_register_create", "mfm_flux_register_destroy",
    "mfm_flux_register_last_error", "mfm_flux_register_info",
    "mfm_flux_register_set_zero", "mfm_flux_register_increment_coarse",
    "mfm_flux_register_increment_fine", "mfm_flux_register_reflux",
    "mfm_level_average_down_dev",
    "mfm_chombo_open", "mfm_chombo_close", "mfm_chombo_last_error",
    "mfm_chombo_header", "mfm_chombo_component_name", "mfm_chombo_level",
    "mfm_chombo_boxes", "mfm_chombo_read_box",
    "mfm_h5_open", "mfm_h5_close", "mfm_h5_last_error", "mfm_h5_dataset_info",
    "mfm_h5_dataset_read", "mfm_h5_attribute_read", "mfm_h5_group_size",
    "mfm_h5_group_name",
]

REQ_PENDING, REQ_ACTIVE, REQ_DONE, REQ_FAILED = range(4)

# kernel kinds of mfm_ctx_get_timing
(KT_PRIM, KT_FLUX_PLASMA, KT_FLUX_NEUTRAL, KT_FLUX_BC, KT_UPDATE, KT_PREDICT, KT_STAGE,
 KT_FSUM_PLASMA, KT_FSUM_NEUTRAL) = range(9)
KT_NAMES = ("k_prim", "k_flux_plasma", "k_flux_neutral", "k_flux_bc", "k_update",
            "k_hancock_pred", "k_stage", "k_fsum_plasma", "k_fsum_neutral")

_lib = None


def load(path=LIB_PATH):
    """Load libmfmhd.so (raises if it is missing: no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            "libmfmhd.so not built (%s); run ms-flukss-multifluid_amd/build_lib.py"
            % path)
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    P = C.POINTER(Params)
    lib.mfm_params_sizeof.restype = C.c_size_t
    lib.mfm_params_default.argtypes = [P]
    lib.mfm_params_set_riemann.argtypes = [P, C.c_double, dp, dp, C.c_double]
    lib.mfm_params_set_orszag.argtypes = [P, C.c_double]
    lib.mfm_params_set_helio.argtypes = [
        P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp,
        dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
        C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
        ip]
    lib.mfm_params_set_swlism.argtypes = [
        P, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, dp,
        C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
        C.c_double, C.c_double, ip]
    lib.mfm_num_components.argtypes = [P]
    lib.mfm_hbm_ceiling.argtypes = [C.c_int, C.c_size_t, C.c_int, dp, dp]
    lib.mfm_level_check.argtypes = [P, C.c_void_p, C.c_int, C.c_double, C.c_double, ip,
                                    C.c_char_p, C.c_size_t]
    lib.mfm_num_ghost_cells.argtypes = [P]
    lib.mfm_ctx_create.argtypes = [P, C.POINTER(C.c_void_p)]
    lib.mfm_ctx_destroy.argtypes = [C.c_void_p]
    lib.mfm_ctx_set_dedner.argtypes = [C.c_void_p, C.c_double, C.c_double]
    lib.mfm_ctx_last_error.argtypes = [C.c_void_p]
    lib.mfm_ctx_set_timing.argtypes = [C.c_void_p, C.c_int]
    lib.mfm_ctx_set_graphs.argtypes = [C.c_void_p, C.c_int]
    lib.mfm_ctx_set_walk.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.mfm_ctx_get_walk.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.mfm_ctx_autotune.argtypes = [C.c_void_p, Fab, Box, C.c_double, C.c_double,
                                     C.c_double, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_double)]
    lib.mfm_ctx_get_timing.argtypes = [C.c_void_p, C.c_int, C.c_int,
                                       C.POINTER(C.c_double),
                                       C.POINTER(C.c_long),
                                       C.POINTER(C.c_double)]
    lib.mfm_ctx_last_error.restype = C.c_char_p
    lib.mfm_patch_update_dev.argtypes = [
        C.c_void_p, Fab, Box, C.c_double, C.c_double, C.c_double,
        C.POINTER(Fab), C.POINTER(Fab), dp, ip, C.c_void_p]
    lib.mfm_patch_update_noalloc_dev.argtypes = [
        C.c_void_p, Fab, Box, C.c_double, C.c_double, C.c_double,
        C.POINTER(Fab), C.POINTER(Fab), dp, ip, C.c_void_p]
    lib.mfm_patch_update_host.argtypes = [
        C.c_void_p, Fab, Box, C.c_double, C.c_double, C.c_double,
        C.POINTER(Fab), C.POINTER(Fab), dp, ip]
    lib.mfm_patch_compute_dt_dev.argtypes = [
        C.c_void_p, Fab, Box, C.c_double, dp, ip, C.c_void_p]
    lib.mfm_level_update_dev.argtypes = [
        C.c_void_p, C.POINTER(PatchRec), C.c_int, C.c_double, C.c_double,
        C.c_double, dp, ip, dp, ip, C.c_void_p]
    lib.mfm_init_problem_dev.argtypes = [C.c_void_p, Fab, Box, C.c_double,
                                         C.c_void_p]
    lib.mfm_fill_periodic_dev.argtypes = [C.c_void_p, Fab, Box, C.c_void_p]
    vp, dbl, lng, it = C.c_void_p, C.c_double, C.c_long, C.c_int
    lib.mfm_server_create.argtypes = [P, dbl, it, C.POINTER(vp)]
    lib.mfm_server_set_stage_kernel.argtypes = [vp, it]
    lib.mfm_server_destroy.argtypes = [vp]
    lib.mfm_server_last_error.argtypes = [vp]
    lib.mfm_server_last_error.restype = C.c_char_p
    lib.mfm_server_submit.argtypes = [vp, P, dbl, it, dbl, dbl, dbl, C.POINTER(lng)]
    lib.mfm_server_step.argtypes = [vp, C.POINTER(it), C.POINTER(it), C.POINTER(dbl)]
    lib.mfm_server_request_info.argtypes = [vp, lng, C.POINTER(it), C.POINTER(it),
                                            C.POINTER(dbl), C.POINTER(Box), C.POINTER(it)]
    lib.mfm_server_fetch.argtypes = [vp, lng, vp, C.c_size_t]
    lib.mfm_server_release.argtypes = [vp, lng]
    lib.mfm_server_request_error.argtypes = [vp, lng]
    lib.mfm_server_request_error.restype = C.c_char_p
    lib.mfm_server_stats.argtypes = [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(lng),
                                     C.POINTER(dbl), C.POINTER(dbl)]
    BP = C.POINTER(Box)
    FP = C.POINTER(Fab)
    lib.mfm_fill_plan_create.argtypes = [P, BP, it, BP, it, it, it, dbl, C.POINTER(vp)]
    lib.mfm_fill_plan_destroy.argtypes = [vp]
    lib.mfm_fill_plan_last_error.argtypes = [vp]
    lib.mfm_fill_plan_last_error.restype = C.c_char_p
    lib.mfm_fill_plan_info.argtypes = [vp, C.POINTER(lng), C.POINTER(lng), C.POINTER(lng),
                                       C.POINTER(C.c_size_t)]
    lib.mfm_level_exchange_dev.argtypes = [vp, vp, FP, vp]
    lib.mfm_level_fill_interp_dev.argtypes = [vp, vp, FP, FP, FP, dbl, vp]
    lib.mfm_fill_time_coef.argtypes = [dbl, dbl, dbl, dbl, it, C.POINTER(dbl)]
    lib.mfm_flux_register_create.argtypes = [P, BP, it, BP, it, it, C.POINTER(vp)]
    lib.mfm_flux_register_destroy.argtypes = [vp]
    lib.mfm_flux_register_last_error.argtypes = [vp]
    lib.mfm_flux_register_last_error.restype = C.c_char_p
    lib.mfm_flux_register_info.argtypes = [vp, C.POINTER(lng), C.POINTER(lng), C.POINTER(lng)]
    lib.mfm_flux_register_set_zero.argtypes = [vp, vp, vp]
    lib.mfm_flux_register_increment_coarse.argtypes = [vp, vp, FP, dbl, vp]
    lib.mfm_flux_register_increment_fine.argtypes = [vp, vp, FP, dbl, vp]
    lib.mfm_flux_register_reflux.argtypes = [vp, vp, FP, dbl, dbl, vp]
    lib.mfm_level_average_down_dev.argtypes = [vp, vp, FP, FP, dbl, vp]
    ip, sz = C.POINTER(it), C.c_size_t
    lib.mfm_chombo_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.mfm_chombo_close.argtypes = [vp]
    lib.mfm_chombo_last_error.argtypes = [vp]
    lib.mfm_chombo_last_error.restype = C.c_char_p
    lib.mfm_chombo_header.argtypes = [vp, ip, ip, ip, C.POINTER(dbl)]
    lib.mfm_chombo_component_name.argtypes = [vp, it]
    lib.mfm_chombo_component_name.restype = C.c_char_p
    lib.mfm_chombo_level.argtypes = [vp, it, ip, C.POINTER(dbl), C.POINTER(dbl), ip, BP, ip]
    lib.mfm_chombo_boxes.argtypes = [vp, it, BP, it]
    lib.mfm_chombo_read_box.argtypes = [vp, it, it, C.POINTER(dbl), sz]
    lib.mfm_h5_open.argtypes = [C.c_char_p, C.POINTER(vp)]
    lib.mfm_h5_close.argtypes = [vp]
    lib.mfm_h5_last_error.argtypes = [vp]
    lib.mfm_h5_last_error.restype = C.c_char_p
    lib.mfm_h5_dataset_info.argtypes = [vp, C.c_char_p, ip, C.POINTER(C.c_long), ip,
                                        C.POINTER(sz)]
    lib.mfm_h5_dataset_read.argtypes = [vp, C.c_char_p, vp, sz]
    lib.mfm_h5_attribute_read.argtypes = [vp, C.c_char_p, C.c_char_p, vp, sz, ip,
                                          C.POINTER(sz)]
    lib.mfm_h5_group_size.argtypes = [vp, C.c_char_p, ip]
    lib.mfm_h5_group_name.argtypes = [vp, C.c_char_p, it, C.c_char_p, sz]
    lib.mfm_dev_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
    lib.mfm_dev_free.argtypes = [C.c_void_p]
    lib.mfm_dev_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.mfm_dev_memset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    lib.mfm_event_create.argtypes = [C.POINTER(C.c_void_p)]
    lib.mfm_event_destroy.argtypes = [C.c_void_p]
    lib.mfm_event_record.argtypes = [C.c_void_p, C.c_void_p]
    lib.mfm_event_elapsed_ms.argtypes = [C.c_void_p, C.c_void_p,
                                         C.POINTER(C.c_float)]
    if lib.mfm_params_sizeof() != C.sizeof(Params):
        raise RuntimeError("mfm_params layout mismatch: C %d vs ctypes %d" %
                           (lib.mfm_params_sizeof(), C.sizeof(Params)))
    _lib = lib
    return lib


def _darr(vals):
    return (C.c_double * len(vals))(*[float(v) for v in vals])


def default_params():
    p = Params()
    load().mfm_params_default(C.byref(p))
    return p


def riemann_params(gamma, left, right, start_x):
    p = default_params()
    load().mfm_params_set_riemann(C.byref(p), gamma, _darr(left), _darr(right),
                                  start_x)
    return p


def orszag_params(gamma):
    p = default_params()
    load().mfm_params_set_orszag(C.byref(p), gamma)
    return p


def helio_params(gamma=5.0 / 3.0, lismN=0.06, lismV=2.64e6, lismT=6527.0,
                 lismB=1.5e-6, lismBdir=(1.0, 0.0, 0.0),
                 lismUdir=(-1.0, 0.0, 0.0), sunXYZ=(0.0, 0.0, 0.0), R0=0.1,
                 sunN=7.4, sunV=4.5e7, sunT=51100.0, sunB=0.0, sunTILT=0.0,
                 initR=80.0, netN=0.1, TMLIM=50000.0, regSafeDZ=0.0,
                 dx=1.0, num_cells=(1, 1, 1)):
    """HeliosphericProblem input with the reference's defaults
    (MHDAM-3.2.4/swlism/HeliosphericProblem.cpp:83-110)."""
    p = default_params()
    nc = (C.c_int * 3)(*[int(v) for v in num_cells])
    load().mfm_params_set_helio(
        C.byref(p), gamma, lismN, lismV, lismT, lismB, _darr(lismBdir),
        _darr(lismUdir), _darr(sunXYZ), R0, sunN, sunV, sunT, sunB, sunTILT,
        initR, netN, TMLIM, regSafeDZ, dx, nc)
    return p


def swlism_params(gamma=5.0 / 3.0, lismN=0.06, lismV=2.64e6, lismT=6527.0,
                  lismB=1.5e-6, sunXY=(600.0, 0.0), R0=10.0, sunN=7.4, sunV=4.5e7,
                  sunT=51100.0, initR=80.0, netN=0.1, TMLIM=50000.0, dx=15.625,
                  num_cells=(96, 64), nfluids=2):
    """SWLISMProblem input (MHDAM-3.2.4/swlism/SWLISMProblem.cpp:68-256)
    with the reference deck's values (main/inputs/swlism2F1.inputs):
    2-D axisymmetric, lengths in AU, dx = domain_length / num_cells[0]."""
    p = default_params()
    nc = (C.c_int * 2)(*[int(v) for v in num_cells])
    rc = load().mfm_params_set_swlism(
        C.byref(p), gamma, lismN, lismV, lismT, lismB, _darr(sunXY), R0, sunN, sunV,
        sunT, initR, netN, TMLIM, dx, nc)
    if rc != OK:
        raise MfmError(rc, "mfm_params_set_swlism failed")
    p.nfluids = nfluids
    return p


def num_components(p):
    return load().mfm_num_components(C.byref(p))


def num_ghost_cells(p):
    return load().mfm_num_ghost_cells(C.byref(p))


def fab(data_ptr, lo, hi, ncomp):
    return Fab(C.c_void_p(int(data_ptr)), Box.of(lo, hi), int(ncomp))


def fab_array(fabs):
    """A ctypes array of Fab from a list (an array passes through, so a
    caller can build a level's table once)."""
    if isinstance(fabs, C.Array):
        return fabs
    return (Fab * max(1, len(fabs)))(*fabs)


class MfmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mfm error %d: %s" % (code, msg))
        self.code = code


class Context:
    """A patch integrator bound to the current HIP device (replaces
    PatchMHDMF construction, define() and the COMMON-block setters)."""

    def __init__(self, params):
        self.lib = load()
        self.params = params
        h = C.c_void_p()
        rc = self.lib.mfm_ctx_create(C.byref(params), C.byref(h))
        if rc != OK:
            raise MfmError(rc, "mfm_ctx_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.mfm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise MfmError(rc, self.last_error())

    def last_error(self):
        msg = self.lib.mfm_ctx_last_error(self.h)
        return msg.decode() if msg else ""

    @property
    def ncomp(self):
        return num_components(self.params)

    @property
    def nghost(self):
        return num_ghost_cells(self.params)

    def set_dedner(self, ch, cp):
        self._check(self.lib.mfm_ctx_set_dedner(self.h, ch, cp))

    def set_timing(self, on):
        """Bracket every kernel launch with HIP events (resets the sums)."""
        self._check(self.lib.mfm_ctx_set_timing(self.h, 1 if on else 0))

    def set_graphs(self, on):
        """Replay the captured launch sequence while the table keeps its shape."""
        self._check(self.lib.mfm_ctx_set_graphs(self.h, 1 if on else 0))

    def set_walk(self, row_tile, xcd_order):
        self._check(self.lib.mfm_ctx_set_walk(self.h, int(row_tile), int(xcd_order)))

    def set_stage_kernel(self, mode, tz=8):
        """Kernel structure of RK2 / first-order stages: STAGE_SEPARATE,
        STAGE_FUSED (True), STAGE_SUMMED or STAGE_AUTO (the default: summed
        where the column tiles are well filled, FP32-flux build); tz =
        planes per z-chunk of the tile kernels."""
        if isinstance(mode, bool):
            mode = STAGE_FUSED if mode else STAGE_SEPARATE
        self._check(self.lib.mfm_ctx_set_stage_kernel(self.h, int(mode), int(tz)))

    def walk(self):
        t, x = C.c_int(), C.c_int()
        self._check(self.lib.mfm_ctx_get_walk(self.h, C.byref(t), C.byref(x)))
        return t.value, x.value

    def autotune(self, U, valid, dt, dx, time=0.0, reps=2):
        """Pick the fastest work walk on a copy of U -> (row_tile, xcd, ms)."""
        t, x, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.mfm_ctx_autotune(self.h, U, valid, dt, dx, time, int(reps),
                                              C.byref(t), C.byref(x), C.byref(ms)))
        return t.value, x.value, ms.value

    def timing(self, stage, kind):
        """(milliseconds, launches, work items) summed for one RK2 stage
        (0 predictor, 1 corrector / single stage) and kernel kind."""
        ms, n, it = C.c_double(), C.c_long(), C.c_double()
        self._check(self.lib.mfm_ctx_get_timing(self.h, stage, kind, C.byref(ms),
                                                C.byref(n), C.byref(it)))
        return ms.value, n.value, it.value

    def update_state(self, U, valid, dt, dx, time=0.0, F=None, divB=None,
                     stream=None, check=True, noalloc=False):
        """PatchMHDAM::updateState on a device FAB; returns (new_dt, cell).
        noalloc: go through mfm_patch_update_noalloc_dev (scratch in the
        library's fixed per-device array, no allocation in the call)."""
        nd = C.c_double()
        cell = (C.c_int * 3)()
        Fp = None
        if F is not None:
            arr = (Fab * 3)(*F)
            Fp = C.cast(arr, C.POINTER(Fab))
            self._Fkeep = arr
        dp = C.pointer(divB) if divB is not None else None
        fn = self.lib.mfm_patch_update_noalloc_dev if noalloc else self.lib.mfm_patch_update_dev
        rc = fn(self.h, U, valid, dt, dx, time, Fp, dp, C.byref(nd), cell,
                C.c_void_p(stream) if stream else None)
        if check:
            self._check(rc)
            return nd.value, tuple(cell)
        return rc, nd.value, tuple(cell)

    def update_state_host(self, U, lo, valid, dt, dx, time=0.0, F=None, divB=None):
        """mfm_patch_update_host: the same update on host arrays (C-ordered
        float64 (nc, nz, ny, nx) FAB data whose low corner is `lo`), staged
        through HBM by the library; U (and F, divB) are written back in
        place.  F: one host array per direction, (nc, nz, ny, nx) of the face
        box, with its low corner; divB likewise -> (new_dt, cell)."""
        import numpy as np

        def host_fab(a, alo):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.ndim == 4
                    and a.flags.c_contiguous and a.flags.writeable):
                raise MfmError(ERR_ARG, "host FAB must be a writeable C-ordered float64 "
                                        "(nc, nz, ny, nx) array")
            ahi = tuple(alo[d] + a.shape[3 - d] - 1 for d in range(3))
            return fab(a.ctypes.data, alo, ahi, a.shape[0])
        Uf = host_fab(U, lo)
        nd = C.c_double()
        cell = (C.c_int * 3)()
        Fp = None
        if F is not None:
            arr = (Fab * 3)(*[host_fab(a, alo) for a, alo in F])
            Fp = C.cast(arr, C.POINTER(Fab))
        dp = C.pointer(host_fab(*divB)) if divB is not None else None
        self._check(self.lib.mfm_patch_update_host(self.h, Uf, valid, dt, dx, time, Fp, dp,
                                                   C.byref(nd), cell))
        return nd.value, tuple(cell)

    def update_level(self, patches, dt, dx, time=0.0, stream=None):
        """One launch sequence over a patch table (list of (U, valid) or
        (U, valid, F) with F the patch's face-flux FABs, one per direction)
        -> (level_dt, patch index, per-patch dts, per-patch cells)."""
        n = len(patches)
        tab = (PatchRec * n)()
        keep = []
        for q, rec in enumerate(patches):
            tab[q].U = rec[0]
            tab[q].valid = rec[1]
            if len(rec) > 2 and rec[2] is not None:
                arr = (Fab * len(rec[2]))(*rec[2])
                keep.append(arr)
                tab[q].F = C.cast(arr, C.POINTER(Fab))
        nd = (C.c_double * n)()
        mc = (C.c_int * (3 * n))()
        ld = C.c_double()
        mp = C.c_int()
        rc = self.lib.mfm_level_update_dev(
            self.h, tab, n, dt, dx, time, nd, mc, C.byref(ld), C.byref(mp),
            C.c_void_p(stream) if stream else None)
        self._check(rc)
        return (ld.value, mp.value, list(nd),
                [tuple(mc[3 * q:3 * q + 3]) for q in range(n)])

    def compute_dt(self, U, box, dx, stream=None):
        nd = C.c_double()
        cell = (C.c_int * 3)()
        self._check(self.lib.mfm_patch_compute_dt_dev(
            self.h, U, box, dx, C.byref(nd), cell,
            C.c_void_p(stream) if stream else None))
        return nd.value, tuple(cell)

    def init_problem(self, U, box, dx, stream=None):
        self._check(self.lib.mfm_init_problem_dev(
            self.h, U, box, dx, C.c_void_p(stream) if stream else None))

    def fill_periodic(self, U, valid, stream=None):
        self._check(self.lib.mfm_fill_periodic_dev(
            self.h, U, valid, C.c_void_p(stream) if stream else None))

    def exchange(self, plan, fabs, stream=None):
        """LevelData::exchange of the level's FABs (in plan order)."""
        arr = fab_array(fabs)
        self._check(self.lib.mfm_level_exchange_dev(
            self.h, plan.h, arr, C.c_void_p(stream) if stream else None))

    def fill_interp(self, plan, fabs, coarse_old, coarse_new, alpha, stream=None):
        """PiecewiseLinearFillPatchMHDAM::fillInterp of the level's ghost
        cells from the coarser level (old and/or new data, either may be
        None)."""
        arr = fab_array(fabs)
        co = fab_array(coarse_old) if coarse_old is not None else None
        cn = fab_array(coarse_new) if coarse_new is not None else None
        self._check(self.lib.mfm_level_fill_interp_dev(
            self.h, plan.h, arr, co, cn, float(alpha),
            C.c_void_p(stream) if stream else None))


class FillPlan:
    """Ghost-fill plan of a level layout (mfm_fill_plan_create): exchange
    and, with coarse boxes, coarse-fine interpolation.  Host-only until first
    used by Context.exchange / Context.fill_interp."""

    def __init__(self, params, boxes, coarse_boxes=None, ref_ratio=2,
                 interp_radius=4, dx_coarse=1.0):
        self.lib = load()
        fb = (Box * max(1, len(boxes)))(*[Box.of(lo, hi) for lo, hi in boxes])
        cbl = coarse_boxes or []
        cb = (Box * max(1, len(cbl)))(*[Box.of(lo, hi) for lo, hi in cbl])
        h = C.c_void_p()
        rc = self.lib.mfm_fill_plan_create(
            C.byref(params), fb, len(boxes), cb if cbl else None, len(cbl),
            int(ref_ratio), int(interp_radius), float(dx_coarse), C.byref(h))
        self.h = h
        if rc != OK:
            msg = self.lib.mfm_fill_plan_last_error(h) if h else b""
            self.close()
            raise MfmError(rc, msg.decode() if msg else "")

    def info(self):
        xc, fc, cc, nb = C.c_long(), C.c_long(), C.c_long(), C.c_size_t()
        self.lib.mfm_fill_plan_info(self.h, C.byref(xc), C.byref(fc), C.byref(cc),
                                    C.byref(nb))
        return {"exchange_cells": xc.value, "interp_fine_cells": fc.value,
                "interp_coarse_cells": cc.value, "device_bytes": nb.value}

    def close(self):
        if self.h:
            self.lib.mfm_fill_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FluxRegister:
    """Coarse-fine synchronisation of a level pair (mfm_flux_register_*):
    the flux register of the interface and the average down.  Methods take
    the Context whose stream and error slot they use."""

    def __init__(self, params, fine_boxes, coarse_boxes, ref_ratio=2):
        self.lib = load()
        fb = (Box * max(1, len(fine_boxes)))(*[Box.of(lo, hi) for lo, hi in fine_boxes])
        cb = (Box * max(1, len(coarse_boxes)))(*[Box.of(lo, hi) for lo, hi in coarse_boxes])
        h = C.c_void_p()
        rc = self.lib.mfm_flux_register_create(C.byref(params), fb, len(fine_boxes), cb,
                                               len(coarse_boxes), int(ref_ratio), C.byref(h))
        self.h = h
        if rc != OK:
            msg = self.lib.mfm_flux_register_last_error(h) if h else b""
            self.close()
            raise MfmError(rc, msg.decode() if msg else "")

    def info(self):
        f, c, a = C.c_long(), C.c_long(), C.c_long()
        self.lib.mfm_flux_register_info(self.h, C.byref(f), C.byref(c), C.byref(a))
        return {"faces": f.value, "cells": c.value, "average_cells": a.value}

    @staticmethod
    def _arr(fabs):
        return fab_array(fabs)

    def set_zero(self, ctx, stream=None):
        ctx._check(self.lib.mfm_flux_register_set_zero(ctx.h, self.h,
                                                       C.c_void_p(stream) if stream else None))

    def increment_coarse(self, ctx, fluxes, scale, stream=None):
        """fluxes: per coarse box, per direction (flattened box-major)."""
        ctx._check(self.lib.mfm_flux_register_increment_coarse(
            ctx.h, self.h, self._arr(fluxes), float(scale), C.c_void_p(stream) if stream else None))

    def increment_fine(self, ctx, fluxes, scale, stream=None):
        ctx._check(self.lib.mfm_flux_register_increment_fine(
            ctx.h, self.h, self._arr(fluxes), float(scale), C.c_void_p(stream) if stream else None))

    def reflux(self, ctx, U_coarse, scale, dx_coarse, stream=None):
        ctx._check(self.lib.mfm_flux_register_reflux(
            ctx.h, self.h, self._arr(U_coarse), float(scale), float(dx_coarse),
            C.c_void_p(stream) if stream else None))

    def average_down(self, ctx, U_coarse, U_fine, dx_fine, stream=None):
        ctx._check(self.lib.mfm_level_average_down_dev(
            ctx.h, self.h, self._arr(U_coarse), self._arr(U_fine), float(dx_fine),
            C.c_void_p(stream) if stream else None))

    def close(self):
        if self.h:
            self.lib.mfm_flux_register_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fill_time_coef(time, t_coarse_new, dt_coarse, dt_fine, ref_ratio):
    """alpha of AMRLevelIdealMHD::step (srcIdeal/AMRLevelIdealMHD.cpp:710-745)."""
    a = C.c_double()
    rc = load().mfm_fill_time_coef(float(time), float(t_coarse_new), float(dt_coarse),
                                   float(dt_fine), int(ref_ratio), C.byref(a))
    if rc != OK:
        raise MfmError(rc, "time coefficient outside [0, 1]")
    return a.value


# ---- device memory and timing on the library's runtime ---------------------
def device_count():
    return load().mfm_dev_count()


def level_check(params, table, dt, dx):
    """mfm_level_check: the update's host checks of a patch table
    [(Fab, Box)], without a device -> (code, first refused patch, reason)."""
    arr = (PatchRec * len(table))()
    for q, (f, b) in enumerate(table):
        arr[q].U = f
        arr[q].valid = b
    bad = C.c_int()
    why = C.create_string_buffer(256)
    rc = load().mfm_level_check(C.byref(params), C.cast(arr, C.c_void_p), len(table),
                                float(dt), float(dx), C.byref(bad), why, 256)
    return rc, bad.value, why.value.decode()


CEIL_READ, CEIL_COPY, CEIL_TRIAD = range(3)


def hbm_ceiling(kind, nbytes=1 << 30, reps=5):
    """(GB/s, best ms) of the library's read / copy / triad kernel over
    arrays of nbytes each (bytes moved: 1x, 2x, 3x nbytes)."""
    g, ms = C.c_double(), C.c_double()
    rc = load().mfm_hbm_ceiling(int(kind), int(nbytes), int(reps), C.byref(g), C.byref(ms))
    if rc != OK:
        raise MfmError(rc, "mfm_hbm_ceiling failed")
    return g.value, ms.value


def synchronize():
    rc = load().mfm_dev_synchronize()
    if rc != OK:
        raise MfmError(rc, "device synchronize failed")


class DeviceArray:
    """A float64 device buffer holding a numpy-shaped array."""

    def __init__(self, shape, data=None, zero=False):
        import numpy as np
        self.lib = load()
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        rc = self.lib.mfm_dev_alloc(max(self.nbytes, 8), C.byref(p))
        if rc != OK:
            raise MfmError(rc, "device allocation of %d bytes failed" % self.nbytes)
        self.ptr = p.value
        if data is not None:
            self.upload(data)
        elif zero:
            self.lib.mfm_dev_memset(C.c_void_p(self.ptr), 0, self.nbytes)

    def upload(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        rc = self.lib.mfm_dev_memcpy(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data),
                                     self.nbytes, 0)
        if rc != OK:
            raise MfmError(rc, "host->device copy failed")

    def copy_from(self, other):
        """Device-to-device copy of an array of the same size."""
        assert other.nbytes == self.nbytes
        rc = self.lib.mfm_dev_memcpy(C.c_void_p(self.ptr), C.c_void_p(other.ptr),
                                     self.nbytes, 2)
        if rc != OK:
            raise MfmError(rc, "device->device copy failed")

    def numpy(self):
        import numpy as np
        a = np.empty(self.shape, dtype=np.float64)
        rc = self.lib.mfm_dev_memcpy(C.c_void_p(a.ctypes.data), C.c_void_p(self.ptr),
                                     self.nbytes, 1)
        if rc != OK:
            raise MfmError(rc, "device->host copy failed")
        return a

    def free(self):
        if self.ptr:
            self.lib.mfm_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """A HIP event for stream-ordered timing."""

    def __init__(self):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mfm_event_create(C.byref(h))
        if rc != OK:
            raise MfmError(rc, "event create failed")
        self.h = h

    def record(self, stream=None):
        self.lib.mfm_event_record(self.h, C.c_void_p(stream) if stream else None)

    def elapsed_ms(self, stop):
        ms = C.c_float()
        rc = self.lib.mfm_event_elapsed_ms(self.h, stop.h, C.byref(ms))
        if rc != OK:
            raise MfmError(rc, "event timing failed")
        return ms.value

    def __del__(self):
        try:
            self.lib.mfm_event_destroy(self.h)
        except Exception:
            pass


class Server:
    """Ensemble serving over mfm_server_* (continuous batching of
    independent solves of one equation class, admitted by an HBM budget)."""

    def __init__(self, cls_params, hbm_budget_bytes, max_active=1 << 20):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mfm_server_create(C.byref(cls_params), float(hbm_budget_bytes),
                                        int(max_active), C.byref(h))
        if rc != OK:
            raise MfmError(rc, "mfm_server_create failed")
        self.h = h

    def close(self):
        if self.h:
            self.lib.mfm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            msg = self.lib.mfm_server_last_error(self.h)
            raise MfmError(rc, msg.decode() if msg else "")

    def set_stage_kernel(self, mode):
        """Stage structure of every batch: STAGE_SEPARATE (default),
        STAGE_SUMMED or STAGE_FUSED; before the first request."""
        self._check(self.lib.mfm_server_set_stage_kernel(self.h, int(mode)))

    def submit(self, params, dx, nsteps, cfl, initial_cfl=None, max_dt_growth=1.0e30):
        """Queue a solve; dt control as the AMR driver's (first step
        initial_cfl * dt_raw, then min(cfl * dt_raw, growth * last dt))."""
        rid = C.c_long()
        c0 = cfl if initial_cfl is None else initial_cfl
        self._check(self.lib.mfm_server_submit(self.h, C.byref(params), float(dx),
                                               int(nsteps), float(cfl), float(c0),
                                               float(max_dt_growth), C.byref(rid)))
        return rid.value

    def step(self):
        """One iteration: (n_active, n_pending, iteration ms)."""
        na, npd, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.mfm_server_step(self.h, C.byref(na), C.byref(npd), C.byref(ms)))
        return na.value, npd.value, ms.value

    def info(self, rid):
        st, nd, t, nc = C.c_int(), C.c_int(), C.c_double(), C.c_int()
        b = Box()
        self._check(self.lib.mfm_server_request_info(self.h, rid, C.byref(st), C.byref(nd),
                                                     C.byref(t), C.byref(b), C.byref(nc)))
        lo, hi = tuple(b.lo), tuple(b.hi)
        return {"state": st.value, "steps": nd.value, "t": t.value, "lo": lo, "hi": hi,
                "ncomp": nc.value}

    def fetch(self, rid):
        """The request's state FAB as a numpy array (nc, nz, ny, nx)."""
        import numpy as np
        inf = self.info(rid)
        lo, hi = inf["lo"], inf["hi"]
        shape = (inf["ncomp"], hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
        a = np.empty(shape, dtype=np.float64)
        self._check(self.lib.mfm_server_fetch(self.h, rid, C.c_void_p(a.ctypes.data), a.nbytes))
        return a

    def error(self, rid):
        m = self.lib.mfm_server_request_error(self.h, rid)
        return m.decode() if m else ""

    def release(self, rid):
        self._check(self.lib.mfm_server_release(self.h, rid))

    def stats(self):
        p50, p99, cu, busy = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        it = C.c_long()
        self._check(self.lib.mfm_server_stats(self.h, C.byref(p50), C.byref(p99), C.byref(it),
                                              C.byref(cu), C.byref(busy)))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "iterations": it.value,
                "cell_updates": cu.value, "busy_ms": busy.value}


class H5File:
    """Read-only view of an HDF5 file through the library's format reader
    (mfm_h5_*): dataset shapes and values, attributes, group members."""

    CLASS_INT, CLASS_FLOAT, CLASS_STRING, CLASS_COMPOUND = 0, 1, 3, 6

    def __init__(self, path):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mfm_h5_open(os.fsencode(path), C.byref(h))
        if rc != OK:
            raise MfmError(rc, "cannot read %s as HDF5" % path)
        self.h = h

    def _check(self, rc):
        if rc != OK:
            raise MfmError(rc, self.lib.mfm_h5_last_error(self.h).decode())

    def members(self, path="/"):
        n = C.c_int()
        self._check(self.lib.mfm_h5_group_size(self.h, path.encode(), C.byref(n)))
        out = []
        for i in range(n.value):
            buf = C.create_string_buffer(4096)
            self._check(self.lib.mfm_h5_group_name(self.h, path.encode(), i, buf, 4096))
            out.append(buf.value.decode())
        return out

    def dataset_info(self, path):
        rank, cls, size = C.c_int(), C.c_int(), C.c_size_t()
        dims = (C.c_long * 8)()
        self._check(self.lib.mfm_h5_dataset_info(self.h, path.encode(), C.byref(rank), dims,
                                                 C.byref(cls), C.byref(size)))
        return tuple(dims[:rank.value]), cls.value, size.value

    def read(self, path, dtype):
        import numpy as np
        shape, cls, size = self.dataset_info(path)
        a = np.empty(shape, dtype=dtype)
        if a.itemsize != size:
            raise MfmError(ERR_SHAPE, "element size %d, dataset has %d" % (a.itemsize, size))
        self._check(self.lib.mfm_h5_dataset_read(self.h, path.encode(),
                                                 a.ctypes.data_as(C.c_void_p), a.nbytes))
        return a

    def attribute(self, obj_path, name):
        """(type class, raw bytes) of an attribute."""
        cls, nb = C.c_int(), C.c_size_t()
        self._check(self.lib.mfm_h5_attribute_read(self.h, obj_path.encode(), name.encode(),
                                                   None, 0, C.byref(cls), C.byref(nb)))
        buf = C.create_string_buffer(max(1, nb.value))
        self._check(self.lib.mfm_h5_attribute_read(self.h, obj_path.encode(), name.encode(),
                                                   buf, nb.value, C.byref(cls), C.byref(nb)))
        return cls.value, buf.raw[:nb.value]

    def close(self):
        if self.h:
            self.lib.mfm_h5_close(self.h)
            self.h = None


class ChomboFile:
    """A Chombo HDF5 checkpoint or plot file (mfm_chombo_*): header, levels,
    boxes and the FAB of each box in the device API's layout."""

    def __init__(self, path):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.mfm_chombo_open(os.fsencode(path), C.byref(h))
        if rc != OK:
            raise MfmError(rc, "cannot read %s as a Chombo file" % path)
        self.h = h
        nl, nc, sd, t = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._check(self.lib.mfm_chombo_header(self.h, C.byref(nl), C.byref(nc), C.byref(sd),
                                               C.byref(t)))
        self.num_levels, self.num_components = nl.value, nc.value
        self.space_dim, self.time = sd.value, t.value
        self.component_names = [self.lib.mfm_chombo_component_name(self.h, c).decode()
                                for c in range(self.num_components)]

    def _check(self, rc):
        if rc != OK:
            raise MfmError(rc, self.lib.mfm_chombo_last_error(self.h).decode())

    def level(self, lev):
        nb, dx, dt, rr = C.c_int(), C.c_double(), C.c_double(), C.c_int()
        dom = Box()
        g = (C.c_int * 3)()
        self._check(self.lib.mfm_chombo_level(self.h, lev, C.byref(nb), C.byref(dx), C.byref(dt),
                                              C.byref(rr), C.byref(dom), g))
        return {"nboxes": nb.value, "dx": dx.value, "dt": dt.value, "ref_ratio": rr.value,
                "prob_domain": (tuple(dom.lo), tuple(dom.hi)), "ghost": tuple(g)}

    def boxes(self, lev):
        n = self.level(lev)["nboxes"]
        arr = (Box * max(1, n))()
        self._check(self.lib.mfm_chombo_boxes(self.h, lev, arr, n))
        return [(tuple(arr[i].lo), tuple(arr[i].hi)) for i in range(n)]

    def read_box(self, lev, b):
        """(FAB values as (nc, nz, ny, nx), lo corner of the grown box)."""
        import numpy as np
        info = self.level(lev)
        lo, hi = self.boxes(lev)[b]
        g = info["ghost"]
        glo = tuple(lo[d] - g[d] for d in range(3))
        ghi = tuple(hi[d] + g[d] for d in range(3))
        shape = (self.num_components, ghi[2] - glo[2] + 1, ghi[1] - glo[1] + 1,
                 ghi[0] - glo[0] + 1)
        a = np.empty(shape)
        self._check(self.lib.mfm_chombo_read_box(self.h, lev, b,
                                                 a.ctypes.data_as(C.POINTER(C.c_double)),
                                                 a.size))
        return a, glo

    def close(self):
        if self.h:
            self.lib.mfm_chombo_close(self.h)
            self.h = None
