/*
 * mfm.h -- C ABI of the MI355X multi-fluid MHD patch integrator.
 *
 * The library advances one (or a table of) FArrayBox-shaped patches of
 * conserved variables by one Godunov step, exactly as the reference patch
 * integrator does:
 *
 *   PatchMHDAM::updateState   (MHDAM-3.2.4/src/PatchMHDAM.cpp:1194-1317,
 *                              declared at src/PatchMHDAM.H:252-260)
 *   PatchMHDMF::computeDt     (MHDAM-3.2.4/srcIdeal/PatchMHDMF.cpp:66-146)
 *   PatchMHDAM::numGhostCells (MHDAM-3.2.4/src/PatchMHDAM.cpp:280-286)
 *
 * Data layout is the Chombo FArrayBox layout: component-major slabs, x
 * fastest, index = (i-lo0) + n0*((j-lo1) + n1*((k-lo2) + n2*comp)).
 * Boxes are cell-centred and inclusive (Chombo Box smallEnd/bigEnd).
 * Every array pointer handed to an entry point that ends in _dev is a device
 * (HBM) pointer; entry points that end in _host take host memory and stage it.
 *
 * The Fortran COMMON blocks the reference mutates (/eos/, /RiemSol/,
 * /Dedner/, /chargeexcm/, /swlismcm/, /riemanncm/) are replaced by the
 * immutable mfm_params block below, copied into the context at creation.
 * Errors are returned as codes (the reference aborts through CH_assert /
 * MayDay::Error); on error the state arrays are left unmodified.
 */
#ifndef MFM_H
#define MFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFM_ABI_VERSION 2

/* ---- return codes ---------------------------------------------------- */
enum {
  MFM_OK = 0,
  MFM_ERR_ARG = -1,        /* bad argument / unsupported option          */
  MFM_ERR_SHAPE = -2,      /* box/ghost/ncomp inconsistent               */
  MFM_ERR_DEVICE = -3,     /* HIP runtime failure                        */
  MFM_ERR_NUMERIC = -4,    /* non-positive or non-finite dt (reference:  */
                           /* CH_assert(dt > 0), src/PhysProblem.cpp:316) */
  MFM_ERR_NOMEM = -5,
  MFM_ERR_UNSUPPORTED = -6
};

/* ---- enumerations (values equal the reference's input-deck integers) -- */
/* mhdam.flux (main/MHDAM.cpp:911-924, src/RiemannSolver.cpp:36-105)     */
/* 0 is the factory's default case, the 8-wave Roe flux on arithmetic     */
/* means like 1; 11 and 13 are the 7-wave Roe flux (no B_n wave, zero     */
/* normal-field flux) on arithmetic and Roe means; GLM cleaning takes 1   */
/* or 3 and runs the 7-wave flux with the (B_n, psi) block.  21 is the    */
/* HLLC flux of ideal MHD (S. Li, JCP 203, 2005), which the reference     */
/* does not have: pinned by its closed forms, not by reference parity;    */
/* with GLM cleaning the (B_n, psi) block replaces its normal-field flux. */
enum { MFM_FLUX_DEFAULT = 0, MFM_FLUX_ROE8 = 1, MFM_FLUX_RUSANOV = 2,
       MFM_FLUX_ROE8_ROEAVG = 3, MFM_FLUX_RUSANOV_ROEAVG = 4,
       MFM_FLUX_ROE7 = 11, MFM_FLUX_ROE7_ROEAVG = 13, MFM_FLUX_HLLC = 21 };
/* limiter1D::eLimiters (src/limiter1D.H:17-28)                           */
enum { MFM_LIM_CD = 0, MFM_LIM_MINMOD = 1, MFM_LIM_SUPERBEE = 2,
       MFM_LIM_VANALBADA = 3, MFM_LIM_MC = 4, MFM_LIM_HARMONIC = 5,
       MFM_LIM_HYPERBEE = 6, MFM_LIM_KOREN = 7, MFM_LIM_VENKAT = 8,
       MFM_LIM_VANLEER = 9, MFM_LIM_WENO3 = 10, MFM_LIM_WENO3YC = 11,
       MFM_LIM_NONE = 12 /* constant (first-order) face values */ };
/* mhdam.time_approximation (main/common.txt:118-122)                     */
enum { MFM_TIME_FIRST_ORDER = 0, MFM_TIME_HANCOCK = 1, MFM_TIME_RK2 = 2 };
/* mhdam.coord_sys (CoordinateSystemHandler::eCoordinateSystem)           */
enum { MFM_COORD_CARTESIAN = 0, MFM_COORD_AXISYMMETRIC = 1 };
/* physical problem supplying boundary conditions and problem sources     */
enum { MFM_PROBLEM_NONE = 0,      /* all ghosts supplied by the caller    */
       MFM_PROBLEM_RIEMANN = 1,   /* problems/RiemannProblem.cpp         */
       MFM_PROBLEM_HELIO = 2,     /* swlism/HeliosphericProblem.cpp      */
       MFM_PROBLEM_SWLISM = 3 };  /* swlism/SWLISMProblem.cpp: 2-D,      */
                                  /* coord_sys = MFM_COORD_AXISYMMETRIC   */
/* arithmetic of reconstruction + Riemann fluxes                          */
enum { MFM_PREC_FP64 = 0,         /* everything FP64 (parity build)       */
       MFM_PREC_FP32_FLUX = 1 };  /* FP32 faces/fluxes, FP64 state, FP64 */
                                  /* cons->prim, floors and accumulation  */

/* ---- geometry -------------------------------------------------------- */
typedef struct mfm_box { int lo[3]; int hi[3]; } mfm_box;

/* An FArrayBox view.  data points to ncomp slabs over box. */
typedef struct mfm_fab {
  double* data;
  mfm_box box;
  int ncomp;
} mfm_fab;

/* ---- parameters (replace the reference's COMMON blocks) ---------------- */
typedef struct mfm_params {
  int abi_version;          /* must be MFM_ABI_VERSION                      */
  int spacedim;             /* 2 or 3 (CH_SPACEDIM)                          */
  /* equation system: EqSysMHDMF(iDedner, nFluids, 0)                       */
  int nfluids;              /* 1..4 : plasma + (nfluids-1) neutral fluids  */
  int dedner;               /* 1 = GLM psi component (divB_cleaning = 5)   */
  /* numerics */
  int flux;                 /* MFM_FLUX_*                                   */
  int limiter[8];           /* limiter_RHO..limiter_BZ; all other slopes   */
                            /* (neutrals, psi) use minmod                  */
  int check_positivity;     /* mhdam.check_positivity; accepted, no effect:
                               in the reference it only queries the slopes'
                               flag, which stays false (main/MHDAM.cpp:978-987,
                               src/slopes.cpp:84-87), so its positivity
                               fallback never runs                           */
  int time_approx;          /* MFM_TIME_FIRST_ORDER, _HANCOCK or _RK2       */
  int divb_method;          /* mhdam.divB_method: 1 faces (COMPUTEBN), 2 CD
                               (COMPUTEBNCD).  0 is refused: the reference
                               then leaves the face field unset
                               (srcIdeal/PatchMHDMF.cpp:517) and reads it  */
  int eight_wave;           /* mhdam.8wave_use                              */
  int coord_sys;            /* MFM_COORD_*                                  */
  int precision;            /* MFM_PREC_*                                   */
  double entropy_fix;       /* del2 (mhdam.entropy_fix_coeff, default 0.3)  */
  double small_b;           /* smallB (1e-9, src/RiemannSolver.cpp:42)      */
  /* /eos/ */
  double gamma, smallr, smallp;
  /* /Dedner/ initial values; the level driver updates them per step      */
  double dedner_ch, dedner_cp;
  /* problem domain (ProblemDomain): cell range and periodicity            */
  int domain_lo[3], domain_hi[3], periodic[3];
  int problem;              /* MFM_PROBLEM_*                               */
  /* Riemann problem (SETRIEMANN): rho,u,v,w,p,bx,by,bz left / right       */
  double rp_left[8], rp_right[8];
  double rp_start_x;              /* initial discontinuity position (startX) */
  /* Heliospheric problem (SETHELIOS output, nondimensional)               */
  double sun_xc, sun_yc, sun_zc;  /* sun position (node-snapped)          */
  double sun_r0;                  /* inner sphere radius R0               */
  double sun_rho, sun_p, sun_v, sun_b, sun_tilt, sun_int_bc_radius;
  double init_r;                  /* initial solar-wind sphere (initR)    */
  double dir_br_n, omega, lism_v_dim;  /* lismV in cm/s for sunBeta       */
  int    sun_b_monopolar;
  double lism_ux, lism_uy, lism_uz, lism_p, lism_m, lism_t;
  double lism_bx, lism_by, lism_bz;
  double net_rho, net_p;
  double tmlim, reg_safe_z_top, reg_safe_z_bot;
  double r0_dt_ignore;            /* SW/LISM: cells within this distance  */
                                  /* of the sun do not limit dt            */
                                  /* (m_R0dtIgnore = R0 - dx of the        */
                                  /* coarsest level, SWLISMProblem.cpp:153);*/
                                  /* heliospheric: unused, the radius is   */
                                  /* R0 - 2 dx of the dx of each call      */
                                  /* (HeliosphericProblem::define :497)    */
  /* /chargeexcm/ (SETCHARGEEX_PARS, swlism/ChargeExchange2F.ChF:4-21)     */
  double ce_scale_vel;            /* lismV (cm/s)                         */
  double ce_sigma_lindsay;        /* 1e-16 * scaleLen * lismN             */
  int    ce_single_literals;      /* 1: Fortran default-real literals are */
                                  /* REAL*4 (13.493, 0.531, ...)          */
  /* mhdam.char_reconstruction (main/MHDAM.cpp:929-941): 1 = limit the    */
  /* plasma slopes in characteristic variables (Reconstruction::faceValues,*/
  /* src/Reconstruction.cpp:178-204)                                     */
  int    char_reconstruction;
  /* mhdam.check_tvd (main/MHDAM.cpp:942-966), read only with             */
  /* char_reconstruction: after the synthesis the face values of the     */
  /* flagged variables are clipped by CHECKTVD (src/reconstructionF.ChF: */
  /* 1248-1278).  1 = all plasma variables, 2 = density and pressure.    */
  int    check_tvd;
} mfm_params;

/* sizeof(mfm_params), for FFI bindings to check their struct layout. */
size_t mfm_params_sizeof(void);
/* Fill p with the reference defaults (main/MHDAM.cpp:647-1046). */
int mfm_params_default(mfm_params* p);
/* Problem setters, restating SETRIEMANN (problems/RiemannF.ChF:4-63) and
 * SETHELIOS + SETCHARGEEX_PARS (swlism/HeliosphericF.ChF:5-131,
 * swlism/HeliosphericProblem.cpp:81-454).  Dimensional inputs as in the
 * input decks; dx_level0 is the coarsest-level cell size. */
int mfm_params_set_riemann(mfm_params* p, double gamma,
                           const double left[8], const double right[8],
                           double start_x);
int mfm_params_set_orszag(mfm_params* p, double gamma);
int mfm_params_set_helio(mfm_params* p, double gamma,
                         double lismN, double lismV, double lismT, double lismB,
                         const double lismBdir[3], const double lismUdir[3],
                         const double sunXYZ[3], double R0, double sunN,
                         double sunV, double sunT, double sunB, double sunTILT,
                         double initR, double netN, double TMLIM,
                         double regSafeDZ, double dx_level0,
                         const int num_cells[3]);

/* SETSWLISM + SETCHARGEEX_PARS (swlism/SWLISMF.ChF:6-104,
 * swlism/SWLISMProblem.cpp:68-256): the 2-D axisymmetric solar-wind / LISM
 * problem (coord_sys = 1, the symmetry axis at y = 0 is the low y boundary),
 * num_cells[0] x num_cells[1] cells of dx_level0 (= domain_length /
 * num_cells[0]); lengths in AU, the sun node-snapped, flow along +x. */
int mfm_params_set_swlism(mfm_params* p, double gamma, double lismN, double lismV,
                          double lismT, double lismB, const double sunXY[2],
                          double R0, double sunN, double sunV, double sunT,
                          double initR, double netN, double TMLIM,
                          double dx_level0, const int num_cells[2]);

/* Number of state components for these parameters (EqSysMHDMF layout). */
int mfm_num_components(const mfm_params* p);
/* PatchMHDAM::numGhostCells: 4 (RK2), 3 (Hancock), 2 (first order). */
int mfm_num_ghost_cells(const mfm_params* p);

/* ---- context ------------------------------------------------------------ */
typedef struct mfm_ctx mfm_ctx;

/* Replaces PatchMHDMF construction + define() + the COMMON-block setters.
 * Binds to the current HIP device. */
int mfm_ctx_create(const mfm_params* p, mfm_ctx** out);
int mfm_ctx_destroy(mfm_ctx* ctx);
/* EqSysMHDMF::setCorrectionPotential -> SETDEDNERCONST (src/DednerF.ChF:4-13) */
int mfm_ctx_set_dedner(mfm_ctx* ctx, double ch, double cp);
/* Last error message of this context (static storage, never NULL). */
const char* mfm_ctx_last_error(const mfm_ctx* ctx);

/* Kernel timing (instrumentation; no counterpart in the reference, which
 * times whole steps with its own timers).  When on, HIP events bracket
 * every kernel launch of the updates on their stream; get_timing returns,
 * per RK2 stage (0 = predictor, 1 = corrector or the single first-order
 * stage) and kernel kind, the summed milliseconds, the number of launches
 * and the summed work items (cells, or faces for the flux kernels).
 * set_timing resets the sums. */
enum {
  MFM_KT_PRIM = 0,         /* cons->prim and physical ghost states       */
  MFM_KT_FLUX_PLASMA = 1,  /* interior faces, plasma Riemann flux        */
  MFM_KT_FLUX_NEUTRAL = 2, /* interior faces, neutral van Leer fluxes    */
  MFM_KT_FLUX_BC = 3,      /* faces on a physical boundary (fluxBC)      */
  MFM_KT_UPDATE = 4,       /* flux divergence, sources, floors, CFL      */
  MFM_KT_PREDICT = 5,      /* Hancock predictor (cell-local half step)   */
  MFM_KT_STAGE = 6,        /* fused stage: interior faces + update       */
  MFM_KT_FSUM_PLASMA = 7,  /* summed stage: plasma face sums of a tile   */
  MFM_KT_FSUM_NEUTRAL = 8  /* summed stage: one neutral fluid's sums     */
};
int mfm_ctx_set_timing(mfm_ctx* ctx, int on);
/* hipGraph replay (no counterpart in the reference): when on, the launch
 * sequence of an update (every kernel of both stages and the dt reduction)
 * is captured once and replayed while the patch table keeps its shape; the
 * per-call values (dt, time, constants) travel in the tables the library
 * copies before each replay.  Ignored while kernel timing is on.           */
int mfm_ctx_set_graphs(mfm_ctx* ctx, int on);
/* Work walk of the kernels (no counterpart in the reference): rows of a
 * region are visited in row_tile x row_tile tiles in (y, z), and with
 * xcd_order = 1 every XCD receives a contiguous range of blocks.  Any value
 * gives the same results; the walk only changes cache reuse.  Defaults 8, 1. */
int mfm_ctx_set_walk(mfm_ctx* ctx, int row_tile, int xcd_order);
int mfm_ctx_get_walk(const mfm_ctx* ctx, int* row_tile, int* xcd_order);
/* Autotuner: times `reps` updates of a device copy of U (U itself is not
 * modified) for row tiles {2, 4, 8, 16, 32} x xcd_order {0, 1}, keeps the
 * fastest walk in the context and returns it with its ms per update.       */
int mfm_ctx_autotune(mfm_ctx* ctx, mfm_fab U, mfm_box valid, double dt,
                     double dx, double time, int reps, int* row_tile,
                     int* xcd_order, double* best_ms);
int mfm_ctx_get_timing(mfm_ctx* ctx, int stage, int kind, double* ms,
                       long* launches, double* items);
/* Kernel structure of RK2 and first-order stages (no counterpart in the
 * reference; every mode gives the reference's results).
 * MFM_STAGE_SEPARATE: per-direction flux kernels writing the face
 * fluxes, then k_update.  MFM_STAGE_SUMMED: after k_prim and the
 * boundary-face kernel, one tile kernel per fluid group (the plasma with
 * psi and the normal B, then each neutral fluid) solves the interior faces
 * of a 64 x 8 column tile, z-chunks of tz planes (default 8), into LDS and
 * writes per cell the face-flux differences summed over the directions;
 * k_update applies those sums (FP32-flux build only; the FP64 build keeps
 * the reference's order, the directions one at a time, and runs the
 * separate kernels in this mode).  MFM_STAGE_FUSED: one tile kernel per stage
 * solves every face and updates the cells; the face fluxes never reach HBM,
 * but its register budget makes it slower on the 256^3 workload.
 * MFM_STAGE_AUTO (default): per batch, the summed stages in the FP32-flux
 * build when the batch's cells fill at least 3/4 of their 64 x 8 column
 * tiles (12.0 against 13.9-14.2 ms per step on the 256^3 workload), else the
 * separate kernels; the allocation-free entry point always takes the
 * separate kernels.  AUTO decides per call: in the FP32-flux build a patch
 * updated alone and in a level can then take different structures (results
 * equal within the FP32 bar, not bit for bit); a fixed mode gives the same
 * bits whatever the batch.  The Hancock scheme and characteristic slopes
 * always use the separate kernels.                                         */
enum { MFM_STAGE_SEPARATE = 0, MFM_STAGE_FUSED = 1, MFM_STAGE_SUMMED = 2, MFM_STAGE_AUTO = 3 };
int mfm_ctx_set_stage_kernel(mfm_ctx* ctx, int mode, int tz);

/* HBM ceilings of the current device, measured with the library's own
 * streaming kernels (no counterpart in the reference, a CPU code): read
 * (sum of an array), copy (b = a) and triad (a = b + s c) over arrays of
 * `bytes` each (a multiple of 16, >= 1 MiB), 16-byte accesses, 8 blocks of
 * 256 threads per CU.  One warm-up launch, then the best of `reps`;
 * gbps = bytes moved (1x, 2x or 3x `bytes`) / time.  Allocates its arrays
 * with hipMalloc and frees them before returning.                          */
enum { MFM_CEIL_READ = 0, MFM_CEIL_COPY = 1, MFM_CEIL_TRIAD = 2 };
int mfm_hbm_ceiling(int kind, size_t bytes, int reps, double* gbps, double* best_ms);

/* ---- one patch: PatchMHDAM::updateState ---------------------------------
 * U      : device FAB, box = valid grown by >= numGhostCells, ghosts filled
 *          (exchange / interpolation) as the reference requires; updated in
 *          place on `valid`.  Ghost cells are left unchanged.
 * F      : nullable; if given, spacedim device FABs on the face boxes of
 *          `valid` (box hi[d] = valid.hi[d]+1) with ncomp comps receive the
 *          time-averaged fluxes the reference returns for the flux registers.
 * divB   : nullable device FAB on `valid` (1 comp) -- written when 8-wave
 *          sources are on, as the reference does.
 * dt     : time step; dx : cell size of this level; time : current time.
 * new_dt : out, dx / max wave speed over `valid` (no CFL factor);
 * min_cell: out, cell attaining it.
 * stream : hipStream_t (NULL = default stream).  The call is asynchronous
 *          except that it synchronises to return new_dt.                   */
int mfm_patch_update_dev(mfm_ctx* ctx, mfm_fab U, mfm_box valid,
                         double dt, double dx, double time,
                         mfm_fab* F, mfm_fab* divB,
                         double* new_dt, int min_cell[3], void* stream);
/* Same contract with host FABs: stages U (and F, divB) through HBM.       */
int mfm_patch_update_host(mfm_ctx* ctx, mfm_fab U, mfm_box valid,
                          double dt, double dx, double time,
                          mfm_fab* F, mfm_fab* divB,
                          double* new_dt, int min_cell[3]);
/* Allocation-free variant of mfm_patch_update_dev.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, results and errors as
 * mfm_patch_update_dev, but all device scratch of the update (first-stage
 * state, primitives, face fluxes, tables, CFL partials) is taken from one
 * fixed array of MFM_NOALLOC_SCRATCH_BYTES per GPU, held from library load
 * to process end, and the staged tables from the context; the call makes no
 * device or pinned allocation.  A patch whose scratch exceeds the array is
 * refused with MFM_ERR_NOMEM before any device work (an RK2 step of a
 * 32^3 patch with 4 ghosts and 23 components needs about 75 MB in FP64).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other streams, other threads, other contexts) race.  Graph replay
 * (mfm_ctx_set_graphs) is not used by this entry point.                    */
#define MFM_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
int mfm_patch_update_noalloc_dev(mfm_ctx* ctx, mfm_fab U, mfm_box valid,
                                 double dt, double dx, double time,
                                 mfm_fab* F, mfm_fab* divB,
                                 double* new_dt, int min_cell[3], void* stream);
/* PatchMHDMF::computeDt over `box` (srcIdeal/PatchMHDMF.cpp:66-146). */
int mfm_patch_compute_dt_dev(mfm_ctx* ctx, mfm_fab U, mfm_box box, double dx,
                             double* dt, int min_cell[3], void* stream);

/* ---- a level: the loop of AMRLevelIdealMHD::step over its boxes --------
 * (srcIdeal/AMRLevelIdealMHD.cpp:1185-1325) as ONE launch sequence over a
 * patch table.  Patches may have different sizes.                          */
typedef struct mfm_patch {
  mfm_fab U;          /* device FAB with ghosts                            */
  mfm_box valid;      /* cells to advance                                  */
  mfm_fab* F;         /* nullable: spacedim face FABs (device)            */
  mfm_fab* divB;      /* nullable                                          */
} mfm_patch;

/* new_dt / min_cell: per-patch outputs (arrays of n / 3n), nullable.
 * level_dt: min over patches, min_patch: its index.                        */
int mfm_level_update_dev(mfm_ctx* ctx, const mfm_patch* table, int n,
                         double dt, double dx, double time,
                         double* new_dt, int* min_cell,
                         double* level_dt, int* min_patch, void* stream);

/* The host checks mfm_level_update_dev makes of a patch table (parameters,
 * ncomp, ghosts, unused dimensions, the 2^30-cell limit per component, the
 * domain, the optional F / divB FABs), without a device and without
 * reading any data pointer (no counterpart in the reference, whose patch
 * integrator asserts inside the update).  MFM_OK, or the code the update
 * would return; *bad_patch = the first refused patch (-1 if none) and why
 * (nullable, why_len bytes) the reason.                                    */
int mfm_level_check(const mfm_params* p, const mfm_patch* table, int n, double dt,
                    double dx, int* bad_patch, char* why, size_t why_len);

/* ---- initial data and ghost fill helpers (caller-side utilities) ------- */
/* HELIOINIT (swlism/HeliosphericF.ChF:1150-1361) / ORSZAGINIT
 * (problems/OrszagF.ChF:27-79) / RIEMANNINIT on a device FAB over `box`. */
int mfm_init_problem_dev(mfm_ctx* ctx, mfm_fab U, mfm_box box, double dx,
                         void* stream);
/* Periodic ghost fill of a single-box periodic level (the part of
 * LevelData::exchange that a one-box level needs).                         */
int mfm_fill_periodic_dev(mfm_ctx* ctx, mfm_fab U, mfm_box valid, void* stream);

/* ---- ghost fill of a level on the device --------------------------------
 * The step before updateState in AMRLevelIdealMHD::step
 * (srcIdeal/AMRLevelIdealMHD.cpp:701-772): coarse-fine interpolation of the
 * ghost cells that no box of the level covers, then the same-level exchange.
 *
 * A plan is built once per layout, on the host, from the level's boxes (its
 * problem domain and periodicity are those of `p`, the level's parameters)
 * and, for a refined level, the boxes of the next coarser level:
 *   - exchange: the Copier of LevelData<FArrayBox>::exchange for ghost
 *     depth interp_radius (Chombo 3.2.4, not part of the reference tree):
 *     every ghost cell of a box that lies in another box, or in a periodic
 *     image of any box, receives that box's valid value;
 *   - interpolation: PiecewiseLinearFillPatchMHDAM::define
 *     (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:118-378): the fine ghost
 *     cells to interpolate, the coarse cells under them and, per direction,
 *     whether their slope is centred or one-sided.
 * Creating a plan touches no device.  It fails with MFM_ERR_UNSUPPORTED when
 * the coarse boxes do not cover every coarse cell the interpolation reads
 * (the layouts are not properly nested).
 * coarse_boxes = NULL / ncoarse = 0 builds an exchange-only plan (level 0).
 * ref_ratio >= 2; interp_radius = the level's ghost depth (m_numGhost).
 * The calls on one plan go to one stream (the plan keeps the FAB tables of
 * its last calls on the device and reuses them while the FABs stay put).   */
typedef struct mfm_fill_plan mfm_fill_plan;
int mfm_fill_plan_create(const mfm_params* p, const mfm_box* boxes, int n,
                         const mfm_box* coarse_boxes, int ncoarse,
                         int ref_ratio, int interp_radius, double dx_coarse,
                         mfm_fill_plan** out);
int mfm_fill_plan_destroy(mfm_fill_plan* plan);
const char* mfm_fill_plan_last_error(const mfm_fill_plan* plan);
/* Sizes of the plan: ghost cells written by the exchange, fine cells and
 * coarse cells of the interpolation, device bytes it holds once used.      */
int mfm_fill_plan_info(const mfm_fill_plan* plan, long* exchange_cells,
                       long* interp_fine_cells, long* interp_coarse_cells,
                       size_t* device_bytes);
/* LevelData::exchange (srcIdeal/AMRLevelIdealMHD.cpp:772).  U: n device
 * FABs in plan order, each holding its box grown by interp_radius.         */
int mfm_level_exchange_dev(mfm_ctx* ctx, mfm_fill_plan* plan, const mfm_fab* U,
                           void* stream);
/* PiecewiseLinearFillPatchMHDAM::fillInterp
 * (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:383-432): coarse data at the
 * fine time, alpha*new + (1-alpha)*old (old only when alpha == 0 or new is
 * NULL, new only when alpha == 1 or old is NULL), piecewise constant, then
 * centred / one-sided slopes, the multidimensional limiter (interplimit_v,
 * srcIdeal/FineInterpMHDAMF.ChF:235-385) and, with GLM cleaning, the div(B)
 * slope correction (interp_divB_0, :449-555).  coarse_old / coarse_new:
 * ncoarse device FABs in plan order (valid boxes at least).  Writes only
 * the interpolated ghost cells of U.                                        */
int mfm_level_fill_interp_dev(mfm_ctx* ctx, mfm_fill_plan* plan,
                              const mfm_fab* U, const mfm_fab* coarse_old,
                              const mfm_fab* coarse_new, double alpha,
                              void* stream);
/* ---- coarse-fine synchronisation of a level pair ------------------------
 * The step after the updates: the flux register of the coarse-fine
 * interface and the average down (srcIdeal/AMRLevelIdealMHD.cpp:695,
 * 1282-1307 and 1636-1647).  The register restates Chombo 3.2.4's
 * LevelFluxRegister as MHDAM drives it (setToZero at the start of the coarse
 * step, incrementCoarse(F, dt), incrementFine(F, dt) per fine step,
 * reflux(U, -1, 1/dx)); the average is CoarseAverageExt::averageToCoarse
 * (src/CoarseAverageExt.cpp:126-194, averagewithvolumes,
 * src/CoarseAverageExtF.ChF:10-50).
 *
 * Built once per layout pair, on the host, from the fine level's boxes (its
 * domain and periodicity from `p`, the fine level's parameters) and the
 * coarse level's boxes.  Fine boxes must be coarsenable by ref_ratio.  A
 * register face is a coarse face on the boundary of the coarsened fine
 * region that is not on a non-periodic domain boundary; it belongs to the
 * coarse cell outside the fine region.  After the increments a face holds
 *   dF = sum over fine steps of dt_f <F_f> - dt_c F_c
 * (<F_f> the mean of the r^(d-1) fine faces over it); reflux adds
 *   scale * sigma * dF / dx_coarse
 * to its cell, sigma = +1 when the face is the cell's high face, so that
 * scale = -1 replaces the coarse flux through the interface by the fine
 * fluxes (conservation across levels).  The calls on one register go to one
 * stream.                                                                  */
typedef struct mfm_flux_register mfm_flux_register;
int mfm_flux_register_create(const mfm_params* p, const mfm_box* fine_boxes, int nfine,
                             const mfm_box* coarse_boxes, int ncoarse, int ref_ratio,
                             mfm_flux_register** out);
int mfm_flux_register_destroy(mfm_flux_register* reg);
const char* mfm_flux_register_last_error(const mfm_flux_register* reg);
/* faces of the register, coarse cells reflux touches, coarse cells the
 * average down writes */
int mfm_flux_register_info(const mfm_flux_register* reg, long* faces, long* cells,
                           long* average_cells);
int mfm_flux_register_set_zero(mfm_ctx* ctx, mfm_flux_register* reg, void* stream);
/* F: ncoarse * spacedim face FABs, F[box * spacedim + dir] = the fluxes of
 * that coarse box in that direction (the F output of the update).          */
int mfm_flux_register_increment_coarse(mfm_ctx* ctx, mfm_flux_register* reg,
                                       const mfm_fab* F, double scale, void* stream);
/* F: nfine * spacedim face FABs of the fine boxes, as above.               */
int mfm_flux_register_increment_fine(mfm_ctx* ctx, mfm_flux_register* reg,
                                     const mfm_fab* F, double scale, void* stream);
/* U: ncoarse device FABs of the coarse level.                              */
int mfm_flux_register_reflux(mfm_ctx* ctx, mfm_flux_register* reg, const mfm_fab* U,
                             double scale, double dx_coarse, void* stream);
/* Coarse cells under the fine boxes = volume-weighted mean of their fine
 * children.  U_coarse: ncoarse FABs, U_fine: nfine FABs.                   */
int mfm_level_average_down_dev(mfm_ctx* ctx, mfm_flux_register* reg,
                               const mfm_fab* U_coarse, const mfm_fab* U_fine,
                               double dx_fine, void* stream);

/* The time coefficient of the fill (srcIdeal/AMRLevelIdealMHD.cpp:710-745):
 * t_old = t_coarse_new - dt_coarse, alpha = (time - t_old) /
 * (t_coarse_new - t_old), snapped to 0 or 1 within 0.04 * dt_fine /
 * ref_ratio; MFM_ERR_ARG when it is outside [0, 1].                        */
int mfm_fill_time_coef(double time, double t_coarse_new, double dt_coarse,
                       double dt_fine, int ref_ratio, double* alpha);

/* ---- ensemble serving ----------------------------------------------------
 * No counterpart in the reference, which runs one simulation per process.
 * A server owns independent solves of one equation class (spacedim,
 * coordinate system, nfluids, dedner, Riemann solver family (Roe or
 * Rusanov), time scheme, precision and characteristic slopes of `cls`).  Each request brings
 * its own parameters (problem constants and domain = its grid), dx, step
 * count and CFL number.  mfm_server_step admits pending requests in arrival
 * order while their device bytes (state + update scratch) fit the HBM
 * budget and fewer than max_active run, generates their initial state on
 * the device (mfm_init_problem_dev), then advances every active request by
 * one step (its own dt, see mfm_server_submit) in the same batched launches
 * (continuous batching) and retires those that reached their step count.
 * A request's result is the same alone and in any mix (the server runs one
 * fixed stage structure whatever the batch).  A finished request
 * keeps its state (and its share of the budget) until released.            */
typedef struct mfm_server mfm_server;
enum { MFM_REQ_PENDING = 0, MFM_REQ_ACTIVE = 1, MFM_REQ_DONE = 2, MFM_REQ_FAILED = 3 };
int mfm_server_create(const mfm_params* cls, double hbm_budget_bytes,
                      int max_active, mfm_server** out);
int mfm_server_destroy(mfm_server* srv);
const char* mfm_server_last_error(const mfm_server* srv);
/* The stage structure of every batch (MFM_STAGE_SEPARATE, the default,
 * MFM_STAGE_SUMMED or MFM_STAGE_FUSED; see mfm_ctx_set_stage_kernel).  A
 * fixed structure keeps a request's result independent of its batch, so
 * MFM_STAGE_AUTO is refused; so is a change after the first request (its
 * HBM estimate depends on the structure).  No counterpart in the
 * reference.                                                              */
int mfm_server_set_stage_kernel(mfm_server* srv, int mode);
/* Step control as the AMR driver's (AMRLevelIdealMHD::computeNewDt,
 * computeInitialDt): first dt = initial_cfl * dt_raw, then
 * min(cfl * dt_raw, max_dt_growth * previous dt); dt_raw = dx / c_max. */
int mfm_server_submit(mfm_server* srv, const mfm_params* p, double dx,
                      int nsteps, double cfl, double initial_cfl,
                      double max_dt_growth, long* id);
/* One iteration; iter_ms = wall time of the batched step (0 if idle).      */
int mfm_server_step(mfm_server* srv, int* n_active, int* n_pending,
                    double* iter_ms);
/* state: MFM_REQ_*; fab_box/ncomp describe the state FAB (valid + ghosts). */
int mfm_server_request_info(mfm_server* srv, long id, int* state,
                            int* steps_done, double* t, mfm_box* fab_box,
                            int* ncomp);
/* Why a request failed ("" otherwise; valid until the request is released). */
const char* mfm_server_request_error(mfm_server* srv, long id);
/* Copy the whole state FAB (component-major) of a request to host memory. */
int mfm_server_fetch(mfm_server* srv, long id, double* host, size_t bytes);
int mfm_server_release(mfm_server* srv, long id);
/* Nearest-rank p50 / p99 of the iteration (= per-request step) latencies,
 * iterations run, cell-updates done and the summed iteration time.         */
int mfm_server_stats(mfm_server* srv, double* p50_ms, double* p99_ms,
                     long* iterations, double* cell_updates, double* busy_ms);

/* ---- device utilities for FFI callers without their own HIP runtime ----
 * (allocation, copies and stream-ordered timing on the library's runtime). */
int mfm_dev_count(void);
int mfm_dev_alloc(size_t bytes, void** ptr);
int mfm_dev_free(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device (synchronous)   */
int mfm_dev_memcpy(void* dst, const void* src, size_t bytes, int kind);
int mfm_dev_memset(void* ptr, int value, size_t bytes);
int mfm_dev_synchronize(void);
int mfm_event_create(void** ev);
int mfm_event_destroy(void* ev);
int mfm_event_record(void* ev, void* stream);
int mfm_event_elapsed_ms(void* start, void* stop, float* ms);

/* ------------------------------------------------ Chombo HDF5 files ---- */
/* Reader of the Chombo HDF5 checkpoint / plot files the reference restarts */
/* from (AMR_MHDAM::setupForRestart, srcIdeal/AMR_MHDAM.cpp:242, 584-747;   */
/* AMRLevelIdealMHD::readCheckpointHeader / readCheckpointLevel,            */
/* srcIdeal/AMRLevelIdealMHDIO.cpp:36-200), without the HDF5 library: the   */
/* file-format subset Chombo's writer produces with the library defaults    */
/* (contiguous or compact datasets, symbol-table or compact groups).         */
/* Chunked / compressed datasets and dense groups: MFM_ERR_UNSUPPORTED.      */
typedef struct mfm_chombo_file mfm_chombo_file;
int mfm_chombo_open(const char* path, mfm_chombo_file** out);
int mfm_chombo_close(mfm_chombo_file* f);
const char* mfm_chombo_last_error(const mfm_chombo_file* f);
/* root header: num_levels, num_components, Chombo_global/SpaceDim, time    */
int mfm_chombo_header(mfm_chombo_file* f, int* num_levels, int* num_components,
                      int* space_dim, double* time);
/* "component_<c>" (the state names readCheckpointHeader checks)           */
const char* mfm_chombo_component_name(const mfm_chombo_file* f, int comp);
/* level_<l>: box count, dx, dt, ref_ratio, prob_domain, outputGhost       */
int mfm_chombo_level(mfm_chombo_file* f, int level, int* nboxes, double* dx, double* dt,
                     int* ref_ratio, mfm_box* prob_domain, int ghost[3]);
int mfm_chombo_boxes(mfm_chombo_file* f, int level, mfm_box* boxes, int capacity);
/* values of box b: (box grown by ghost) cells x num_components doubles,   */
/* component-major, x fastest -- an mfm_fab's layout                       */
int mfm_chombo_read_box(mfm_chombo_file* f, int level, int box, double* dst, size_t count);

/* The HDF5 layer underneath, for inspection (paths like "/level_0/boxes") */
typedef struct mfm_h5_file mfm_h5_file;
int mfm_h5_open(const char* path, mfm_h5_file** out);
int mfm_h5_close(mfm_h5_file* f);
const char* mfm_h5_last_error(const mfm_h5_file* f);
/* type_class: 0 integer, 1 floating point, 3 string, 6 compound; dims[8]  */
int mfm_h5_dataset_info(mfm_h5_file* f, const char* path, int* rank, long* dims,
                        int* type_class, size_t* type_size);
int mfm_h5_dataset_read(mfm_h5_file* f, const char* path, void* dst, size_t bytes);
int mfm_h5_attribute_read(mfm_h5_file* f, const char* object_path, const char* name,
                          void* dst, size_t capacity, int* type_class, size_t* bytes);
int mfm_h5_group_size(mfm_h5_file* f, const char* path, int* count);
int mfm_h5_group_name(mfm_h5_file* f, const char* path, int index, char* name,
                      size_t capacity);

#ifdef __cplusplus
}
#endif
#endif /* MFM_H */
