// mfm_capi.cpp -- the C ABI (include/mfm.h): parameters, contexts, scratch
// management and the launch sequence of PatchMHDAM::updateState over one
// patch or a level's patch table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/mfm.h"
#include "mfm_device.hpp"
#include "mfm_host.hpp"
#include "mfm_kernels.hpp"

using namespace mfm;
using namespace mfm::host;

namespace mfm {
namespace host {

constexpr int kBlk = 256;
constexpr double kAutoTileFill = 0.75;  // MFM_STAGE_AUTO: cells / tile slots
const double eos_mp = 1.67262171e-24, eos_k = 1.3806505e-16;

int fail(mfm_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

int hipfail(mfm_ctx* c, hipError_t e, const char* what) {
  return fail(c, MFM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int ensure_arena(mfm_ctx* c, size_t bytes) {
  if (bytes <= c->arena_bytes) return MFM_OK;
  if (c->arena) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hipfail(c, e, "hipDeviceSynchronize");
    (void)hipFree(c->arena);
    c->arena = nullptr;
    c->arena_bytes = 0;
  }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc((void**)&c->arena, want);
  if (e != hipSuccess) {
    c->arena = nullptr;
    return fail(c, MFM_ERR_NOMEM, "hipMalloc of scratch failed");
  }
  c->arena_bytes = want;
  return MFM_OK;
}

int ensure_hbuf(mfm_ctx* c, size_t bytes) {
  if (bytes <= c->hbuf_bytes) return MFM_OK;
  if (c->hbuf) {
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hipfail(c, e, "hipDeviceSynchronize");
    (void)hipHostFree(c->hbuf);
    c->hbuf = nullptr;
    c->hbuf_bytes = 0;
  }
  size_t want = bytes + bytes / 4 + 4096;
  hipError_t e = hipHostMalloc((void**)&c->hbuf, want, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->hbuf = nullptr;
    return fail(c, MFM_ERR_NOMEM, "hipHostMalloc of staging failed");
  }
  c->hbuf_bytes = want;
  return MFM_OK;
}

// Box &= ProblemDomain (periodic directions are not clipped)
HBox clip(const Consts& C, HBox b) {
  for (int d = 0; d < C.dims; ++d)
    if (C.phys[d]) {
      b.lo[d] = std::max(b.lo[d], C.dom_lo[d]);
      b.hi[d] = std::min(b.hi[d], C.dom_hi[d]);
    }
  return b;
}

struct RegionBuilder {
  SmallVec<Region, 8> v;  // the lists of one patch fit inline
  int nblk = 0;
  long items = 0;
  void add(int patch, const HBox& b) {
    long cells = b.size();
    if (cells <= 0) return;
    items += cells;
    Region r;
    r.patch = patch;
    for (int d = 0; d < 3; ++d) { r.lo[d] = b.lo[d]; r.n[d] = (int)b.n(d); }
    r.blk0 = nblk;
    nblk += (int)((cells + kBlk - 1) / kBlk);
    v.push_back(r);
  }
  // a region of a tile kernel: one block per tx x ty column tile and
  // z-chunk of tz planes (k_stage: 16 x 16, k_fsum: 64 x 8)
  void add_tiles(int patch, const HBox& b, int tz, int tx = 16, int ty = 16) {
    long cells = b.size();
    if (cells <= 0) return;
    items += cells;
    Region r;
    r.patch = patch;
    for (int d = 0; d < 3; ++d) { r.lo[d] = b.lo[d]; r.n[d] = (int)b.n(d); }
    r.blk0 = nblk;
    nblk += (int)(((b.n(0) + tx - 1) / tx) * ((b.n(1) + ty - 1) / ty) * ((b.n(2) + tz - 1) / tz));
    v.push_back(r);
  }
};

int build_consts(const mfm_params* p, Consts& C, std::string& why) {
  if (!p || p->abi_version != MFM_ABI_VERSION) { why = "abi_version mismatch"; return MFM_ERR_ARG; }
  if (p->spacedim != 2 && p->spacedim != 3) { why = "spacedim must be 2 or 3"; return MFM_ERR_ARG; }
  if (p->nfluids < 1 || p->nfluids > 4) { why = "nfluids must be 1..4"; return MFM_ERR_UNSUPPORTED; }
  // GLM cleaning with neutral fluids (EqSysMHDMF(iDedner, nF), main/MHDAM.cpp:
  // 776-826): with caller-filled ghosts; the problems' boundary states here
  // define no psi for neutral runs
  if (p->dedner && p->nfluids != 1 && p->problem != MFM_PROBLEM_NONE) {
    why = "GLM cleaning with neutral fluids needs caller-filled ghosts (problem NONE)";
    return MFM_ERR_UNSUPPORTED;
  }
  // the factory's cases (src/RiemannSolver.cpp:39-104) but VFRoe (5);
  // 101/103 are what GLM cleaning makes of 1/3; 21 = HLLC, this library's
  if (!((p->flux >= 0 && p->flux <= 4) || p->flux == 11 || p->flux == 13 ||
        p->flux == MFM_FLUX_HLLC)) {
    why = p->flux == 5 ? "flux = 5 (VFRoe) is not supported"
                       : "flux must be 0..4, 11, 13 or 21 (GLM cleaning: 1, 3 or 21)";
    return p->flux == 5 ? MFM_ERR_UNSUPPORTED : MFM_ERR_ARG;
  }
  if (p->dedner && !(p->flux == 1 || p->flux == 3 || p->flux == MFM_FLUX_HLLC)) {
    // main/MHDAM.cpp:1042-1046 (1 or 3), and this library's HLLC
    why = "divB_cleaning = 5 must be used with flux = 1, 3 or 21";
    return MFM_ERR_ARG;
  }
  if (p->dedner && p->eight_wave) { why = "divB_cleaning = 5 can not be used with 8wave_use = 1"; return MFM_ERR_ARG; }
  if (p->time_approx != MFM_TIME_RK2 && p->time_approx != MFM_TIME_FIRST_ORDER &&
      p->time_approx != MFM_TIME_HANCOCK) {
    why = "time_approximation must be 0 (first order), 1 (Hancock) or 2 (RK2)";
    return MFM_ERR_UNSUPPORTED;
  }
  if (p->time_approx == MFM_TIME_HANCOCK && p->dedner) {
    // FLUXESHANCOCK has no flux for psi (srcIdeal/PatchIdealMHDF.ChF:235-297)
    why = "GLM cleaning is not run with the Hancock scheme";
    return MFM_ERR_UNSUPPORTED;
  }
  if (p->coord_sys == MFM_COORD_AXISYMMETRIC) {
    // CS_Axisymmetric: Cartesian fluxes and volumes plus the geometric
    // sources; with the SW/LISM problem or caller-filled ghosts, in 2-D
    if (!(p->problem == MFM_PROBLEM_SWLISM || p->problem == MFM_PROBLEM_NONE) ||
        p->spacedim != 2 || p->dedner || p->time_approx == MFM_TIME_HANCOCK) {
      why = "coord_sys = 1 runs in 2-D (SW/LISM or caller-filled ghosts), RK2 or first order, without GLM";
      return MFM_ERR_UNSUPPORTED;
    }
  } else if (p->coord_sys != MFM_COORD_CARTESIAN) {
    why = "coord_sys must be 0 (Cartesian) or 1 (axisymmetric)";
    return MFM_ERR_UNSUPPORTED;
  }
  if (p->check_tvd < 0 || p->check_tvd > 2) { why = "check_tvd must be 0, 1 or 2"; return MFM_ERR_ARG; }
  if (p->char_reconstruction && p->precision != MFM_PREC_FP64) {
    why = "char_reconstruction runs in the FP64 build only";
    return MFM_ERR_UNSUPPORTED;
  }
  if (p->char_reconstruction)
    for (int v = 0; v < 8; ++v)
      if (p->limiter[v] == MFM_LIM_NONE) {
        why = "char_reconstruction needs a slope limiter on every plasma variable";
        return MFM_ERR_UNSUPPORTED;
      }
  if (p->precision != MFM_PREC_FP64 && p->precision != MFM_PREC_FP32_FLUX) { why = "bad precision"; return MFM_ERR_ARG; }
  if (p->divb_method < 0 || p->divb_method > 2) { why = "bad divB_method"; return MFM_ERR_ARG; }
  if (p->divb_method == 0) {
    // PatchMHDMF::computeBn returns at once for method 0
    // (srcIdeal/PatchMHDMF.cpp:517), leaving the face-field FluxBox of
    // EulerStep (src/PatchMHDAM.cpp:547) unset; the 8-wave source, fluxBC and
    // the divB output then read unset memory, so no result can match it
    why = "divB_method = 0 leaves the face field unset in the reference";
    return MFM_ERR_UNSUPPORTED;
  }
  for (int v = 0; v < 8; ++v)
    if (p->limiter[v] < 0 || p->limiter[v] > MFM_LIM_NONE) { why = "bad limiter"; return MFM_ERR_ARG; }
  if (!(p->gamma > 1.0)) { why = "gamma must exceed 1"; return MFM_ERR_ARG; }
  if (p->problem == MFM_PROBLEM_RIEMANN) {
    if (p->nfluids != 1) { why = "Riemann problem is single-fluid here"; return MFM_ERR_UNSUPPORTED; }
    // RIEMANNGS fills x only, and the update converts U to W on the
    // domain-clipped box (src/PatchMHDAM.cpp:514-538): the W ghosts of a
    // non-periodic transverse direction stay unset in the reference, so no
    // result can match it (riemann14.inputs: is_periodic = 0 1 0 in 3-D)
    for (int d = 1; d < p->spacedim; ++d)
      if (!p->periodic[d]) {
        why = "Riemann problem: a non-periodic transverse direction reads unset ghost primitives in the reference";
        return MFM_ERR_UNSUPPORTED;
      }
  } else if (p->problem == MFM_PROBLEM_HELIO) {
    if (p->spacedim != 3) { why = "heliospheric problem is 3-D"; return MFM_ERR_UNSUPPORTED; }
  } else if (p->problem == MFM_PROBLEM_SWLISM) {
    if (p->spacedim != 2 || p->coord_sys != MFM_COORD_AXISYMMETRIC) {
      why = "SW/LISM problem is 2-D axisymmetric (coord_sys = 1)";
      return MFM_ERR_UNSUPPORTED;
    }
    if (p->dedner || p->time_approx == MFM_TIME_HANCOCK) {
      why = "SW/LISM problem runs RK2 or first order without GLM cleaning here";
      return MFM_ERR_UNSUPPORTED;
    }
    if (p->periodic[0] || p->periodic[1]) { why = "SW/LISM boundaries are not periodic"; return MFM_ERR_ARG; }
  } else if (p->problem != MFM_PROBLEM_NONE) {
    why = "unknown problem"; return MFM_ERR_ARG;
  }
  std::memset(&C, 0, sizeof(C));
  C.dims = p->spacedim;
  C.nf = p->nfluids;
  C.dedner = p->dedner ? 1 : 0;
  C.nc = 8 + C.dedner + 5 * (C.nf - 1);
  C.iPsi = C.dedner ? 8 : -1;
  C.iRhoN = 8 + C.dedner;
  C.flux_kind = flux_kind_of(*p);
  C.iAveraging = (p->flux == 3 || p->flux == 4 || p->flux == 13) ? 1 : 0;
  C.eight_wave = p->eight_wave ? 1 : 0;
  C.divb_method = p->divb_method;
  C.problem = p->problem;
  for (int c = 0; c < kMaxComp; ++c) C.lim[c] = (int8_t)(c < 8 ? p->limiter[c] : MFM_LIM_MINMOD);
  C.char_rec = p->char_reconstruction ? 1 : 0;
  C.tvd_mask = !C.char_rec ? 0 : (p->check_tvd == 1 ? 0xFF
                                  : (p->check_tvd == 2 ? ((1 << 0) | (1 << 4)) : 0));
  C.plasma_lim = C.lim[0];
  for (int c = 1; c < 8 + C.dedner; ++c)
    if (C.lim[c] != C.lim[0]) C.plasma_lim = -1;
  C.gamma = p->gamma;
  C.hgamma = p->gamma - 1.0;
  C.invhgamma = 1.0 / C.hgamma;
  C.smallr = p->smallr;
  C.smallp = p->smallp;
  C.del2 = p->entropy_fix;
  C.smallB = p->small_b;
  C.ch = p->dedner_ch;
  C.cp = p->dedner_cp;
  for (int d = 0; d < 3; ++d) {
    C.dom_lo[d] = p->domain_lo[d];
    C.dom_hi[d] = p->domain_hi[d];
    C.phys[d] = (d < C.dims && p->problem != MFM_PROBLEM_NONE && !p->periodic[d]) ? 1 : 0;
  }
  for (int c = 0; c < 8; ++c) { C.rpL[c] = p->rp_left[c]; C.rpR[c] = p->rp_right[c]; }
  C.rp_start_x = p->rp_start_x;
  C.sun_xc = p->sun_xc; C.sun_yc = p->sun_yc; C.sun_zc = p->sun_zc;
  C.sun_r0 = p->sun_r0; C.sun_rho = p->sun_rho; C.sun_p = p->sun_p;
  C.sun_v = p->sun_v; C.sun_b = p->sun_b; C.sun_tilt = p->sun_tilt;
  C.sun_int_bc_radius = p->sun_int_bc_radius; C.dir_br_n = p->dir_br_n;
  C.omega = p->omega; C.lism_v_dim = p->lism_v_dim; C.init_r = p->init_r;
  C.sun_b_monopolar = p->sun_b_monopolar;
  C.lism_ux = p->lism_ux; C.lism_uy = p->lism_uy; C.lism_uz = p->lism_uz;
  C.lism_p = p->lism_p; C.lism_m = p->lism_m; C.lism_t = p->lism_t;
  C.lism_bx = p->lism_bx; C.lism_by = p->lism_by; C.lism_bz = p->lism_bz;
  C.net_rho = p->net_rho; C.net_p = p->net_p; C.tmlim = p->tmlim;
  C.reg_safe_z_top = p->reg_safe_z_top; C.reg_safe_z_bot = p->reg_safe_z_bot;
  C.ce_scale_vel = p->ce_scale_vel;
  C.ce_sigma = p->ce_sigma_lindsay;
  // Fortran default-real literals of the cross section (see include/mfm.h)
  C.c13493 = p->ce_single_literals ? (double)13.493f : 13.493;
  C.c0531 = p->ce_single_literals ? (double)0.531f : 0.531;
  C.ce_on = ((p->problem == MFM_PROBLEM_HELIO || p->problem == MFM_PROBLEM_SWLISM) &&
             p->nfluids >= 2) ? 1 : 0;
  // CHARGE_EXCHANGE_MF (the SW/LISM problem) rather than _2F/_3F/_4F
  // SWLISMProblem::explicitSource (swlism/SWLISMProblem.cpp:945-1021) calls
  // CHARGE_EXCHANGE_MF for 2, 3 and 5 fluids and CHARGE_EXCHANGE_4F for 4
  C.ce_mf = (p->problem == MFM_PROBLEM_SWLISM && p->nfluids != 4) ? 1 : 0;
  C.axisym = p->coord_sys == MFM_COORD_AXISYMMETRIC ? 1 : 0;
  C.r0_dt_ignore = p->r0_dt_ignore;
  C.sin_tilt = std::sin(p->sun_tilt);
  C.cos_tilt = std::cos(p->sun_tilt);
  C.sun_beta = p->problem == MFM_PROBLEM_HELIO ? p->omega * kAU / (p->sun_v * p->lism_v_dim) : 0.0;
  return MFM_OK;
}

// Geometry of one patch's scratch.
struct PatchGeo {
  HBox ub, vb1, vb2, fb[3];
  size_t off_up, off_w, off_w2, off_d, off_f[3], off_bn[3];
};

// The host checks of one patch of a batch (dimension, ghosts, sizes, the
// domain and the optional outputs), before any device work; *why is a
// string literal, so a refusal allocates nothing.
int check_patch(const Consts& C0, const Consts& Cq, int ng, const PatchIn& P, const char** why) {
  const int dims = C0.dims;
  if (!(P.dt > 0.0) || !(P.dx > 0.0)) { *why = "dt and dx must be positive"; return MFM_ERR_ARG; }
  if (!P.U.data) { *why = "U.data is NULL"; return MFM_ERR_ARG; }
  if (P.U.ncomp != C0.nc) { *why = "U.ncomp does not match the equation system"; return MFM_ERR_SHAPE; }
  HBox ub = hb(P.U.box), v = hb(P.valid);
  if (v.size() <= 0) { *why = "empty valid box"; return MFM_ERR_SHAPE; }
  for (int d = dims; d < 3; ++d)
    if (ub.lo[d] != ub.hi[d] || v.lo[d] != v.hi[d] || ub.lo[d] != v.lo[d]) {
      *why = "unused dimensions must be a single cell";
      return MFM_ERR_SHAPE;
    }
  if (!contains(ub, grow(v, ng, dims))) {
    *why = "U box must contain the valid box grown by numGhostCells";
    return MFM_ERR_SHAPE;
  }
  if (ub.size() >= kMaxFabCells) { *why = "a FAB component must hold fewer than 2^30 cells"; return MFM_ERR_SHAPE; }
  for (int d = 0; d < dims; ++d)
    if (Cq.phys[d] && (v.lo[d] < Cq.dom_lo[d] || v.hi[d] > Cq.dom_hi[d])) {
      *why = "valid box extends outside the problem domain";
      return MFM_ERR_SHAPE;
    }
  if (P.F) {
    for (int d = 0; d < dims; ++d) {
      HBox fo = hb(P.F[d].box), want = v;
      want.hi[d] += 1;
      if (!P.F[d].data || P.F[d].ncomp != C0.nc || !contains(fo, want)) {
        *why = "F[d] must cover the faces of the valid box with ncomp components";
        return MFM_ERR_SHAPE;
      }
      if (fo.size() >= kMaxFabCells) { *why = "a FAB component must hold fewer than 2^30 cells"; return MFM_ERR_SHAPE; }
    }
  }
  if (P.divB && (!P.divB->data || P.divB->ncomp < 1)) { *why = "divB must have one component"; return MFM_ERR_SHAPE; }
  return MFM_OK;
}

// One batched update of n patches.  Each patch names a row of the
// constants table Ct (nct rows) and carries its own dt, dx and time; all
// rows share the launch class of c (dimension, fluids, GLM, time scheme,
// precision).  c provides the scratch arena, staging, timer and errors.
int run_batch(mfm_ctx* c, const Consts* Ct, int nct, const PatchIn* in, int n,
              double* new_dt, int* min_cell, double* level_dt, int* min_patch,
              hipStream_t s, const FixedScratch* fixed) {
  const Consts& C0 = Ct[0];
  const mfm_params& p = c->p;
  const int dims = C0.dims;
  const bool rk2 = p.time_approx == MFM_TIME_RK2;
  const bool han = p.time_approx == MFM_TIME_HANCOCK;
  const int ng = rk2 ? 4 : (han ? 3 : 2);  // PatchMHDAM::numGhostCells
  // tile kernels for the interior faces (k_stage, or k_fsum before
  // k_update); the Hancock scheme and characteristic slopes keep the
  // separate kernels, and so does the FP64 build for the summed stages: it
  // applies the directions one at a time, in the reference's order
  int mode = (han || C0.char_rec) ? MFM_STAGE_SEPARATE : c->stage_mode;
  if (mode == MFM_STAGE_AUTO) {
    // the summed stages where their 64 x 8 column tiles are well filled
    // (FP32-flux build; the allocation-free entry point keeps the separate
    // kernels, whose scratch its bound is stated for)
    mode = MFM_STAGE_SEPARATE;
    if (p.precision == MFM_PREC_FP32_FLUX && !fixed) {
      double cells = 0.0, slots = 0.0;
      for (int q = 0; q < n; ++q) {
        if (in[q].cidx < 0 || in[q].cidx >= nct) continue;
        const HBox v = clip(Ct[in[q].cidx], hb(in[q].valid));
        if (v.size() <= 0) continue;
        cells += (double)v.size();
        slots += (double)((v.n(0) + kSumTX - 1) / kSumTX * kSumTX) *
                 (double)((v.n(1) + kSumTY - 1) / kSumTY * kSumTY) * (double)v.n(2);
      }
      if (slots > 0.0 && cells >= kAutoTileFill * slots) mode = MFM_STAGE_SUMMED;
    }
  }
  if (mode == MFM_STAGE_SUMMED && p.precision != MFM_PREC_FP32_FLUX) mode = MFM_STAGE_SEPARATE;
  const bool fused = mode == MFM_STAGE_FUSED;
  const bool summed = mode == MFM_STAGE_SUMMED;
  const int tz = std::max(1, c->stage_tz);
  if (n <= 0) return fail(c, MFM_ERR_ARG, "empty patch table");
  for (int r = 1; r < nct; ++r)
    if (Ct[r].dims != C0.dims || Ct[r].nf != C0.nf || Ct[r].dedner != C0.dedner)
      return fail(c, MFM_ERR_ARG, "constants of one batch must share dimension, fluids and GLM");
  SmallVec<PatchGeo, 1> G;
  G.resize(n);
  size_t off = 0;
  const size_t fs = c->fsize();
  for (int q = 0; q < n; ++q) {
    const PatchIn& P = in[q];
    if (P.cidx < 0 || P.cidx >= nct) return fail(c, MFM_ERR_ARG, "bad constants row");
    const Consts& Cq = Ct[P.cidx];
    const char* why = nullptr;
    const int rc = check_patch(C0, Cq, ng, P, &why);
    if (rc != MFM_OK) return fail(c, rc, why);
    HBox ub = hb(P.U.box), v = hb(P.valid);
    PatchGeo& g = G[q];
    g.ub = ub;
    g.vb2 = clip(Cq, v);
    // first-stage value box: RK2 box + 2, Hancock predictor box + 1
    g.vb1 = rk2 ? clip(Cq, grow(v, 2, dims)) : (han ? clip(Cq, grow(v, 1, dims)) : g.vb2);
    const long ucells = ub.size();
    g.off_up = off; off += align256((size_t)ucells * C0.nc * 8);
    g.off_w = off; off += align256((size_t)ucells * C0.nc * fs);
    g.off_w2 = off;
    // Hancock: W of the predicted state; fused RK2: W of U*, written by
    // stage 1 while other blocks still read W of U^n
    if (han || (fused && rk2)) off += align256((size_t)ucells * C0.nc * fs);
    // summed stages: the per-cell face sums (NC components + normal B)
    g.off_d = off;
    if (summed) off += align256((size_t)ucells * (C0.nc + 1) * fs);
    for (int d = 0; d < dims; ++d) {
      g.fb[d] = g.vb1;
      g.fb[d].hi[d] += 1;
      g.off_f[d] = off; off += align256((size_t)g.fb[d].size() * C0.nc * fs);
      g.off_bn[d] = off; off += align256((size_t)g.fb[d].size() * fs);
    }
  }
  // region lists; faces on a physical boundary go to their own list
  const PatchIn* pin = in;
  RegionBuilder st1_prim, st1_flux[3], st1_fbc[3], st1_upd, st2_prim, st2_flux[3],
      st2_fbc[3], st2_upd, st1_sum, st2_sum;
  auto add_faces = [&](RegionBuilder& in, RegionBuilder& bc, int q, HBox f, int d) {
    const Consts& Cq = Ct[pin[q].cidx];
    if (Cq.phys[d]) {
      if (f.lo[d] == Cq.dom_lo[d]) {
        HBox b = f;
        b.hi[d] = b.lo[d];
        bc.add(q, b);
        f.lo[d] += 1;
      }
      if (f.hi[d] == Cq.dom_hi[d] + 1 && f.hi[d] >= f.lo[d]) {
        HBox b = f;
        b.lo[d] = b.hi[d];
        bc.add(q, b);
        f.hi[d] -= 1;
      }
    }
    if (f.hi[d] >= f.lo[d]) in.add(q, f);
  };
  for (int q = 0; q < n; ++q) {
    const PatchGeo& g = G[q];
    if (rk2) {
      st1_prim.add(q, grow(g.vb1, 2, dims));
      for (int d = 0; d < dims; ++d) add_faces(st1_flux[d], st1_fbc[d], q, g.fb[d], d);
      if (fused) st1_upd.add_tiles(q, g.vb1, tz);
      else st1_upd.add(q, g.vb1);
      if (summed) st1_sum.add_tiles(q, g.vb1, tz, kSumTX, kSumTY);
    } else if (han) {
      // predictor: W on its stencil (value box + 2), one kernel on the cells
      st1_prim.add(q, grow(g.vb1, 2, dims));
      st1_upd.add(q, g.vb1);
    }
    if (han) {
      // corrector: the predicted W of the physical ghost cells next to the
      // box (the other cells of grow(vb2, 1) lie in vb1)
      HBox rest = grow(g.vb2, 1, dims);
      for (int d = 0; d < dims; ++d) {
        if (rest.lo[d] < g.vb1.lo[d]) {
          HBox b = rest;
          b.hi[d] = g.vb1.lo[d] - 1;
          st2_prim.add(q, b);
          rest.lo[d] = g.vb1.lo[d];
        }
        if (rest.hi[d] > g.vb1.hi[d]) {
          HBox b = rest;
          b.lo[d] = g.vb1.hi[d] + 1;
          st2_prim.add(q, b);
          rest.hi[d] = g.vb1.hi[d];
        }
      }
    } else if (rk2) {
      // the first stage leaves W of U* on vb1; the second stage only needs
      // the physical ghost cells of grow(vb2, 2) outside it
      HBox rest = grow(g.vb2, 2, dims);
      for (int d = 0; d < dims; ++d) {
        if (rest.lo[d] < g.vb1.lo[d]) {
          HBox b = rest;
          b.hi[d] = g.vb1.lo[d] - 1;
          st2_prim.add(q, b);
          rest.lo[d] = g.vb1.lo[d];
        }
        if (rest.hi[d] > g.vb1.hi[d]) {
          HBox b = rest;
          b.lo[d] = g.vb1.hi[d] + 1;
          st2_prim.add(q, b);
          rest.hi[d] = g.vb1.hi[d];
        }
      }
    } else {
      st2_prim.add(q, grow(g.vb2, 2, dims));
    }
    for (int d = 0; d < dims; ++d) {
      HBox f = g.vb2;
      f.hi[d] += 1;
      add_faces(st2_flux[d], st2_fbc[d], q, f, d);
    }
    if (fused) st2_upd.add_tiles(q, g.vb2, tz);
    else st2_upd.add(q, g.vb2);
    if (summed) st2_sum.add_tiles(q, g.vb2, tz, kSumTX, kSumTY);
  }
  constexpr int kLists = 18;
  RegionBuilder* lists[kLists] = {&st1_prim, &st1_flux[0], &st1_flux[1], &st1_flux[2],
                                  &st1_fbc[0], &st1_fbc[1], &st1_fbc[2], &st1_upd,
                                  &st2_prim, &st2_flux[0], &st2_flux[1], &st2_flux[2],
                                  &st2_fbc[0], &st2_fbc[1], &st2_fbc[2], &st2_upd,
                                  &st1_sum, &st2_sum};
  // table layout inside the arena (after the scratch)
  // fused RK2: a second patch table for stage 2, whose W is the W of U*
  // that stage 1 wrote (the kernels always read W through the table)
  const int ntab = (fused && rk2) ? 2 : 1;
  size_t off_patch = off; off += ntab * align256(sizeof(DevPatch) * n);
  size_t off_reg[kLists];
  size_t table_bytes = ntab * align256(sizeof(DevPatch) * n);
  // constants table after the region lists

  for (int l = 0; l < kLists; ++l) {
    off_reg[l] = off;
    size_t b = align256(sizeof(Region) * std::max<size_t>(1, lists[l]->v.size()));
    off += b;
    table_bytes += b;
  }
  size_t off_ct = off;
  off += align256(sizeof(Consts) * nct);
  table_bytes += align256(sizeof(Consts) * nct);
  const int nbu = std::max(1, st2_upd.nblk);
  size_t off_dtp = off; off += align256(sizeof(double) * nbu);
  size_t off_idp = off; off += align256(sizeof(long) * nbu);
  size_t off_dto = off; off += align256(sizeof(double) * n);
  size_t off_ido = off; off += align256(sizeof(long) * n);
  const size_t host_need = table_bytes + align256(sizeof(double) * n) + align256(sizeof(long) * n) + 512;
  char* A;
  char* h;
  if (fixed) {
    // caller-provided memory: refuse what does not fit, allocate nothing
    if (off > fixed->dev_bytes) {
      char msg[160];
      std::snprintf(msg, sizeof(msg),
                    "the update needs %zu bytes of device scratch; the fixed scratch holds %zu",
                    off, fixed->dev_bytes);
      return fail(c, MFM_ERR_NOMEM, msg);
    }
    if (host_need > fixed->host_bytes)
      return fail(c, MFM_ERR_NOMEM, "the update's tables exceed the fixed host staging");
    A = fixed->dev;
    h = fixed->host;
  } else {
    int rc = ensure_arena(c, off);
    if (rc) return rc;
    rc = ensure_hbuf(c, host_need);
    if (rc) return rc;
    A = c->arena;
    h = c->hbuf;
  }

  // host staging of the tables
  DevPatch* hp = reinterpret_cast<DevPatch*>(h);
  for (int q = 0; q < n; ++q) {
    const PatchIn& P = in[q];
    const Consts& Cq = Ct[P.cidx];
    const PatchGeo& g = G[q];
    DevPatch& D = hp[q];
    std::memset(&D, 0, sizeof(D));
    D.cidx = P.cidx;
    D.dx = P.dx;
    D.dt = P.dt;
    D.time = P.time;
    // PSIPARABOLIC factor exp(-dt_stage Ch^2/Cp^2) of each stage
    D.psi_decay[0] = Cq.dedner ? std::exp(-P.dt * Cq.ch * Cq.ch / (Cq.cp * Cq.cp)) : 1.0;
    D.psi_decay[1] = Cq.dedner ? std::exp(-(0.5 * P.dt) * Cq.ch * Cq.ch / (Cq.cp * Cq.cp)) : 1.0;
    D.U = P.U.data;
    D.Up = reinterpret_cast<double*>(A + g.off_up);
    D.W = A + g.off_w;
    D.W2 = han ? A + g.off_w2 : nullptr;
    D.Wn = (fused && rk2) ? A + g.off_w2 : D.W;
    D.D = summed ? A + g.off_d : nullptr;
    for (int d = 0; d < 3; ++d) {
      D.ulo[d] = g.ub.lo[d];
      D.un[d] = (int)g.ub.n(d);
    }
    D.ucs = g.ub.size();
    for (int d = 0; d < dims; ++d) {
      D.F[d] = A + g.off_f[d];
      D.Bn[d] = A + g.off_bn[d];
      for (int e = 0; e < 3; ++e) { D.flo[d][e] = g.fb[d].lo[e]; D.fn[d][e] = (int)g.fb[d].n(e); }
      D.fcs[d] = g.fb[d].size();
      if (P.F) {
        HBox fo = hb(P.F[d].box);
        D.Fo[d] = P.F[d].data;
        for (int e = 0; e < 3; ++e) { D.folo[d][e] = fo.lo[e]; D.fon[d][e] = (int)fo.n(e); }
        D.focs[d] = fo.size();
      }
    }
    if (P.divB) {
      HBox db = hb(P.divB->box);
      D.divB = P.divB->data;
      for (int e = 0; e < 3; ++e) { D.dlo[e] = db.lo[e]; D.dn[e] = (int)db.n(e); }
    }
    // HELIOREINIT_DEFAULT's / SWLISMREINIT's early exit on the FAB bounds
    D.reinit = 0;
    if (Cq.problem == MFM_PROBLEM_HELIO || Cq.problem == MFM_PROBLEM_SWLISM) {
      const double dx = P.dx, sunR = Cq.sun_r0;
      double xl = g.ub.lo[0] * dx - Cq.sun_xc, yl = g.ub.lo[1] * dx - Cq.sun_yc, zl = g.ub.lo[2] * dx - Cq.sun_zc;
      double xh = (g.ub.hi[0] + 1) * dx - Cq.sun_xc, yh = (g.ub.hi[1] + 1) * dx - Cq.sun_yc,
             zh = (g.ub.hi[2] + 1) * dx - Cq.sun_zc;
      bool out = (xl > sunR) || (yl > sunR) || (dims == 3 && zl > sunR) || (xh < -sunR) ||
                 (yh < -sunR) || (dims == 3 && zh < -sunR);
      D.reinit = out ? 0 : 1;
    }
  }
  if (ntab == 2) {
    // stage 2 reads (and its ghost slabs write) the W of U*
    DevPatch* hp2 = reinterpret_cast<DevPatch*>(h + align256(sizeof(DevPatch) * n));
    for (int q = 0; q < n; ++q) {
      hp2[q] = hp[q];
      hp2[q].W = hp[q].Wn;
    }
  }
  size_t hoff = ntab * align256(sizeof(DevPatch) * n);
  for (int l = 0; l < kLists; ++l) {
    std::memcpy(h + hoff, lists[l]->v.data(), sizeof(Region) * lists[l]->v.size());
    hoff += align256(sizeof(Region) * std::max<size_t>(1, lists[l]->v.size()));
  }
  std::memcpy(h + hoff, Ct, sizeof(Consts) * nct);
  hipError_t e = hipMemcpyAsync(A + off_patch, h, table_bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return hipfail(c, e, "hipMemcpyAsync(tables)");

  auto rl = [&](int l) {
    RegionList r;
    r.r = reinterpret_cast<const Region*>(A + off_reg[l]);
    r.n = (int)lists[l]->v.size();
    r.nblk = lists[l]->nblk;
    r.items = lists[l]->items;
    return r;
  };
  const DevPatch* dp = reinterpret_cast<const DevPatch*>(A + off_patch);
  const DevPatch* dp2 = reinterpret_cast<const DevPatch*>(A + off_patch + (ntab - 1) * align256(sizeof(DevPatch) * n));
  double* dtp = reinterpret_cast<double*>(A + off_dtp);
  long* idp = reinterpret_cast<long*>(A + off_idp);
  const Consts* ct = reinterpret_cast<const Consts*>(A + off_ct);
  auto stage = [&](int l0) {
    StageLists L;
    L.prim = rl(l0);
    for (int d = 0; d < 3; ++d) {
      L.flux[d] = rl(l0 + 1 + d);
      L.fbc[d] = rl(l0 + 4 + d);
      if (d >= dims) { L.flux[d].nblk = 0; L.fbc[d].nblk = 0; }
    }
    L.upd = rl(l0 + 7);
    L.sum = rl(l0 == 0 ? 16 : 17);
    L.fused = fused ? 1 : 0;
    L.summed = summed ? 1 : 0;
    return L;
  };
  if (c->timer) c->timer->n = 0;
  int n_stage1 = 0;
  // the whole launch sequence of one update (both stages, dt reduction)
  auto enqueue = [&](hipStream_t st) -> int {
    auto launch = [&](const StageLists& L, const StepArgs& S) -> hipError_t {
      const DevPatch* tp = (S.stage == 1 && rk2) ? dp2 : dp;
      const bool f32 = p.precision == MFM_PREC_FP32_FLUX;
      hipError_t e3 = f32 ? launch_stage<float>(C0, ct, tp, L, S, dtp, idp, st, c->timer)
                          : launch_stage<double>(C0, ct, tp, L, S, dtp, idp, st, c->timer);
      if (e3 != hipSuccess) return e3;
      if (L.fused)
        return f32 ? launch_stage_fused<float>(C0, ct, tp, L, S, dtp, idp, st, c->timer)
                   : launch_stage_fused<double>(C0, ct, tp, L, S, dtp, idp, st, c->timer);
      if (!L.summed) return e3;
      e3 = f32 ? launch_face_sums<float>(C0, ct, tp, L, S, st, c->timer)
               : launch_face_sums<double>(C0, ct, tp, L, S, st, c->timer);
      if (e3 != hipSuccess) return e3;
      return f32 ? launch_update_summed<float>(C0, ct, tp, L, S, dtp, idp, st, c->timer)
                 : launch_update_summed<double>(C0, ct, tp, L, S, dtp, idp, st, c->timer);
    };
    hipError_t e2;
    if (han) {
      // predictorHancock with dt/2, then correctorHancock with dt
      StepArgs S1{0.5, 0, 0, 0, 0, 0, 0, 0, c->walk_tile, c->walk_xcd, 1, tz};
      e2 = launch(stage(0), S1);
      if (e2 != hipSuccess) return hipfail(c, e2, "predictor launch");
      if (c->timer) n_stage1 = c->timer->n;
      StepArgs S2{1.0, 1, 0, 0, 0, 2, 1, 0, c->walk_tile, c->walk_xcd, 2, tz};
      e2 = launch(stage(8), S2);
      if (e2 != hipSuccess) return hipfail(c, e2, "corrector launch");
    } else if (rk2) {
      StepArgs S1{1.0, 0, 0, 1, 0, 0, 0, 1, c->walk_tile, c->walk_xcd, 0, tz};
      e2 = launch(stage(0), S1);
      if (e2 != hipSuccess) return hipfail(c, e2, "stage 1 launch");
      if (c->timer) n_stage1 = c->timer->n;
      StepArgs S2{0.5, 1, 1, 0, 1, 1, 1, 0, c->walk_tile, c->walk_xcd, 0, tz};
      e2 = launch(stage(8), S2);
      if (e2 != hipSuccess) return hipfail(c, e2, "stage 2 launch");
    } else {
      StepArgs S{1.0, 0, 0, 0, 0, 2, 1, 0, c->walk_tile, c->walk_xcd, 0, tz};
      e2 = launch(stage(8), S);
      if (e2 != hipSuccess) return hipfail(c, e2, "first-order launch");
    }
    RegionList U2 = rl(15);
    e2 = launch_dt_final(U2, dtp, idp, reinterpret_cast<double*>(A + off_dto),
                         reinterpret_cast<long*>(A + off_ido), st);
    if (e2 != hipSuccess) return hipfail(c, e2, "dt reduction launch");
    return MFM_OK;
  };
  if (c->graphs && !c->timer && !fixed) {
    // The kernels read every per-call value (dt, time, constants, tables)
    // from the arena, refreshed by the copy above; the sequence itself only
    // depends on the arena layout and the list sizes.
    std::vector<long long> key = {(long long)(uintptr_t)A, (long long)off_patch,
                                  (long long)off_ct, (long long)off_dtp, (long long)off_idp,
                                  (long long)off_dto, (long long)off_ido, dims, p.time_approx,
                                  p.precision, C0.nf, C0.dedner, c->walk_tile,
                                  c->walk_xcd, mode, tz};
    for (int l = 0; l < kLists; ++l) {
      key.push_back((long long)off_reg[l]);
      key.push_back((long long)lists[l]->v.size());
      key.push_back((long long)lists[l]->nblk);
    }
    if (!c->gexec || key != c->gkey) {
      if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
      if (!c->cap_stream) {
        e = hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking);
        if (e != hipSuccess) { c->cap_stream = nullptr; return hipfail(c, e, "hipStreamCreate"); }
      }
      e = hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeThreadLocal);
      if (e != hipSuccess) return hipfail(c, e, "hipStreamBeginCapture");
      int rc2 = enqueue(c->cap_stream);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(c->cap_stream, &g);
      if (rc2 != MFM_OK) { if (g) (void)hipGraphDestroy(g); return rc2; }
      if (e != hipSuccess) return hipfail(c, e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { c->gexec = nullptr; return hipfail(c, e, "hipGraphInstantiate"); }
      c->gkey = key;
    }
    e = hipGraphLaunch(c->gexec, s);
    if (e != hipSuccess) return hipfail(c, e, "hipGraphLaunch");
  } else {
    int rc2 = enqueue(s);
    if (rc2 != MFM_OK) return rc2;
  }
  double* hdt = reinterpret_cast<double*>(h + table_bytes);
  long* hid = reinterpret_cast<long*>(h + table_bytes + align256(sizeof(double) * n));
  e = hipMemcpyAsync(hdt, A + off_dto, sizeof(double) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hid, A + off_ido, sizeof(long) * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hipfail(c, e, "patch update");
  if (c->timer) {
    mfm::KernelTimer& T = *c->timer;
    for (int q = 0; q < T.n; ++q) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, T.ev[2 * q], T.ev[2 * q + 1]) != hipSuccess) continue;
      const int st = ((rk2 || han) && q < n_stage1) ? 0 : 1;
      c->kt_ms[st][T.kind[q]] += ms;
      c->kt_n[st][T.kind[q]] += 1;
      c->kt_items[st][T.kind[q]] += (double)T.items[q];
    }
    T.n = 0;
  }
  // per-patch results; regions of the last update list are one per patch
  double best = std::numeric_limits<double>::max();
  int bestp = -1;
  bool bad = false;
  for (int q = 0; q < n; ++q) {
    const Region& r = st2_upd.v[q];
    double d = hdt[q];
    long li = hid[q];
    int cell[3] = {r.lo[0], r.lo[1], r.lo[2]};
    if (!(d < std::numeric_limits<double>::infinity())) {
      d = std::numeric_limits<double>::max();  // every cell excluded
    } else {
      long n01 = (long)r.n[0] * r.n[1];
      cell[2] = r.lo[2] + (int)(li / n01);
      long rem = li % n01;
      cell[1] = r.lo[1] + (int)(rem / r.n[0]);
      cell[0] = r.lo[0] + (int)(rem % r.n[0]);
    }
    if (!(d > 0.0) || std::isnan(d)) bad = true;
    if (new_dt) new_dt[q] = d;
    if (min_cell) { min_cell[3 * q] = cell[0]; min_cell[3 * q + 1] = cell[1]; min_cell[3 * q + 2] = cell[2]; }
    if (d < best) { best = d; bestp = q; }
  }
  if (level_dt) *level_dt = best;
  if (min_patch) *min_patch = bestp;
  if (bad) return fail(c, MFM_ERR_NUMERIC, "non-positive or NaN time step (reference: CH_assert(dt > 0))");
  return MFM_OK;
}

int run_dt(mfm_ctx* c, const Consts& C0, const mfm_fab& U, const mfm_box& box,
           double dx, double* dt, int* min_cell, hipStream_t s) {
  if (!c || !U.data || !dt) return MFM_ERR_ARG;
  if (U.ncomp != C0.nc) return fail(c, MFM_ERR_SHAPE, "U.ncomp mismatch");
  HBox ub = hb(U.box), b = hb(box);
  // as the update: an empty box has no cell to reduce over (and would leave
  // the region list below without its one region)
  if (b.size() == 0) return fail(c, MFM_ERR_SHAPE, "empty box");
  if (!contains(ub, b)) return fail(c, MFM_ERR_SHAPE, "box outside U");
  if (ub.size() >= kMaxFabCells) return fail(c, MFM_ERR_SHAPE, "a FAB component must hold fewer than 2^30 cells");
  if (!(dx > 0.0)) return fail(c, MFM_ERR_ARG, "dx must be positive");
  RegionBuilder rb;
  rb.add(0, b);
  size_t off = 0;
  size_t off_p = off; off += align256(sizeof(DevPatch));
  size_t off_r = off; off += align256(sizeof(Region));
  size_t off_c = off; off += align256(sizeof(Consts));
  size_t off_dtp = off; off += align256(sizeof(double) * rb.nblk);
  size_t off_idp = off; off += align256(sizeof(long) * rb.nblk);
  size_t off_dto = off; off += align256(sizeof(double));
  size_t off_ido = off; off += align256(sizeof(long));
  int rc = ensure_arena(c, off);
  if (!rc) rc = ensure_hbuf(c, 4096 + align256(sizeof(DevPatch)) + align256(sizeof(Region)) +
                                 align256(sizeof(Consts)));
  if (rc) return rc;
  DevPatch* hp = reinterpret_cast<DevPatch*>(c->hbuf);
  std::memset(hp, 0, sizeof(DevPatch));
  hp->U = U.data;
  for (int d = 0; d < 3; ++d) { hp->ulo[d] = ub.lo[d]; hp->un[d] = (int)ub.n(d); }
  hp->ucs = ub.size();
  hp->cidx = 0;
  hp->dx = dx;
  Region* hr = reinterpret_cast<Region*>(c->hbuf + off_r);
  *hr = rb.v[0];
  std::memcpy(c->hbuf + off_c, &C0, sizeof(Consts));
  char* A = c->arena;
  hipError_t e = hipMemcpyAsync(A + off_p, c->hbuf, off_c + sizeof(Consts), hipMemcpyHostToDevice, s);
  RegionList L{reinterpret_cast<const Region*>(A + off_r), 1, rb.nblk, rb.items};
  if (e == hipSuccess)
    e = launch_dt_only(C0, reinterpret_cast<const Consts*>(A + off_c),
                       reinterpret_cast<const DevPatch*>(A + off_p), L,
                       reinterpret_cast<double*>(A + off_dtp), reinterpret_cast<long*>(A + off_idp), s);
  if (e == hipSuccess)
    e = launch_dt_final(L, reinterpret_cast<double*>(A + off_dtp), reinterpret_cast<long*>(A + off_idp),
                        reinterpret_cast<double*>(A + off_dto), reinterpret_cast<long*>(A + off_ido), s);
  double* hdt = reinterpret_cast<double*>(c->hbuf + off_c + align256(sizeof(Consts)));
  long* hid = reinterpret_cast<long*>(c->hbuf + off_c + align256(sizeof(Consts)) + 256);
  if (e == hipSuccess) e = hipMemcpyAsync(hdt, A + off_dto, sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(hid, A + off_ido, sizeof(long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hipfail(c, e, "compute_dt");
  double d = *hdt;
  const Region& r = rb.v[0];
  int cell[3] = {r.lo[0], r.lo[1], r.lo[2]};
  if (!(d < std::numeric_limits<double>::infinity())) {
    d = std::numeric_limits<double>::max();
  } else {
    long n01 = (long)r.n[0] * r.n[1];
    long li = *hid;
    cell[2] = r.lo[2] + (int)(li / n01);
    cell[1] = r.lo[1] + (int)((li % n01) / r.n[0]);
    cell[0] = r.lo[0] + (int)((li % n01) % r.n[0]);
  }
  *dt = d;
  if (min_cell) { min_cell[0] = cell[0]; min_cell[1] = cell[1]; min_cell[2] = cell[2]; }
  if (!(d > 0.0)) return fail(c, MFM_ERR_NUMERIC, "non-positive time step");
  return MFM_OK;
}

}  // namespace host
}  // namespace mfm

extern "C" {

size_t mfm_params_sizeof(void) { return sizeof(mfm_params); }

int mfm_params_default(mfm_params* p) {
  if (!p) return MFM_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->abi_version = MFM_ABI_VERSION;
  p->spacedim = 3;
  p->nfluids = 1;
  p->dedner = 0;
  p->flux = MFM_FLUX_ROE8;                       // main/MHDAM.cpp:911
  for (int v = 0; v < 8; ++v) p->limiter[v] = MFM_LIM_MINMOD;  // :656 iLimiter = 1
  p->check_positivity = 0;
  p->time_approx = MFM_TIME_FIRST_ORDER;         // :1016-1022
  p->divb_method = 1;                            // :1030
  p->eight_wave = 1;                             // :1035-1040
  p->coord_sys = MFM_COORD_CARTESIAN;
  p->precision = MFM_PREC_FP64;
  p->entropy_fix = 0.3;                          // :916
  p->small_b = 1.0e-9;                           // src/RiemannSolver.cpp:42
  p->gamma = 5.0 / 3.0;
  p->smallr = 1.0e-8;
  p->smallp = 1.0e-8;
  p->dedner_ch = 1.0;                            // srcIdeal/EqSysMHDMF.cpp:17-18
  p->dedner_cp = 1.0;
  for (int d = 0; d < 3; ++d) { p->domain_lo[d] = 0; p->domain_hi[d] = 0; p->periodic[d] = 1; }
  p->problem = MFM_PROBLEM_NONE;
  p->sun_int_bc_radius = 1.0;
  p->dir_br_n = 1.0;
  p->tmlim = 50000.0;
  p->init_r = 80.0;
  return MFM_OK;
}

// SETRIEMANN (problems/RiemannF.ChF:4-63)
int mfm_params_set_riemann(mfm_params* p, double gamma, const double left[8],
                           const double right[8], double start_x) {
  if (!p || !left || !right) return MFM_ERR_ARG;
  p->problem = MFM_PROBLEM_RIEMANN;
  p->gamma = gamma;
  for (int c = 0; c < 8; ++c) { p->rp_left[c] = left[c]; p->rp_right[c] = right[c]; }
  p->rp_start_x = start_x;
  const double r0 = std::min(left[0], right[0]);
  const double p0 = std::min(left[4], right[4]);
  const double umag = std::sqrt(gamma * p0 / r0);
  p->smallr = 1.0e-8 * r0;
  p->smallp = 1.0e-8 * r0 * 0.5 * umag * umag;
  return MFM_OK;
}

// SETORSZAG (problems/OrszagF.ChF:4-25)
int mfm_params_set_orszag(mfm_params* p, double gamma) {
  if (!p) return MFM_ERR_ARG;
  p->problem = MFM_PROBLEM_NONE;
  p->gamma = gamma;
  const double r0 = gamma * gamma, p0 = gamma;
  const double umag = std::sqrt(gamma * p0 / r0);
  p->smallr = 1.0e-8 * r0;
  p->smallp = 1.0e-8 * r0 * 0.5 * umag * umag;
  return MFM_OK;
}

// HeliosphericProblem::input/define (swlism/HeliosphericProblem.cpp:81-330,
// 463-500), SETHELIOS (swlism/HeliosphericF.ChF:5-131) and
// SETCHARGEEX_PARS (swlism/ChargeExchange2F.ChF:4-21).
int mfm_params_set_helio(mfm_params* p, double gamma, double lismN, double lismV,
                         double lismT, double lismB, const double lismBdir[3],
                         const double lismUdir[3], const double sunXYZ[3],
                         double R0, double sunN, double sunV, double sunT,
                         double sunB, double sunTILT, double initR, double netN,
                         double TMLIM, double regSafeDZ, double dx_level0,
                         const int num_cells[3]) {
  if (!p || !lismBdir || !lismUdir || !sunXYZ || !num_cells) return MFM_ERR_ARG;
  if (!(dx_level0 > 0.0)) return MFM_ERR_ARG;
  p->problem = MFM_PROBLEM_HELIO;
  p->gamma = gamma;
  p->spacedim = 3;
  for (int d = 0; d < 3; ++d) {
    p->domain_lo[d] = 0;
    p->domain_hi[d] = num_cells[d] - 1;
    p->periodic[d] = 0;
  }
  double sun[3];
  for (int d = 0; d < 3; ++d) sun[d] = ((int)std::floor(sunXYZ[d] / dx_level0 + 0.5)) * dx_level0;
  p->sun_xc = sun[0]; p->sun_yc = sun[1]; p->sun_zc = sun[2];
  const double minZ = -sun[2], maxZ = dx_level0 * num_cells[2] - sun[2];
  p->reg_safe_z_top = maxZ - regSafeDZ;
  p->reg_safe_z_bot = minZ + regSafeDZ;
  p->tmlim = TMLIM;
  p->sun_r0 = R0;
  p->init_r = initR;
  p->sun_int_bc_radius = 1.0;
  p->dir_br_n = 1.0;
  p->sun_b_monopolar = 0;
  p->lism_t = lismT;
  p->lism_v_dim = lismV;
  // SETHELIOS
  const double rho = lismN * eos_mp;
  const double pref = rho * lismV * lismV;
  const double pl = 2.0 * eos_k * lismN * lismT;
  p->lism_p = pl / pref;
  p->lism_m = 1.0 / std::sqrt(gamma * p->lism_p);
  const double b = lismB / std::sqrt(pref);
  p->lism_bx = b * lismBdir[0];
  p->lism_by = b * lismBdir[1];
  p->lism_bz = b * lismBdir[2];
  p->lism_ux = lismUdir[0];
  p->lism_uy = lismUdir[1];
  p->lism_uz = lismUdir[2];
  p->sun_tilt = sunTILT * 0.0174532925199432957692369076848861271;
  p->omega = 6.28318530717958647692528676655900576 / (25.38 * 24.0 * 3600.0);
  p->sun_rho = sunN / lismN;
  p->sun_p = (2.0 * eos_k * sunN * sunT) / pref;
  p->sun_v = sunV / lismV;
  p->sun_b = sunB / std::sqrt(pref);
  if (netN > 0.0) {
    p->net_rho = netN / lismN;
    p->net_p = (eos_k * netN * lismT) / pref;
  }
  const double r0 = 1.0;
  const double p0 = 1.0 / (gamma * p->lism_m * p->lism_m);
  const double umag = std::sqrt(gamma * p0 / r0);
  p->smallr = 1.0e-8 * r0;
  p->smallp = 1.0e-8 * r0 * 0.5 * umag * umag;
  if (p->smallr < 1.0e-12) p->smallr = 1.0e-12;
  if (p->smallp < 1.0e-12) p->smallp = 1.0e-12;
  // charge exchange scales, scaleLen = 1 AU
  p->ce_scale_vel = lismV;
  p->ce_sigma_lindsay = 1.0e-16 * 1.5e13 * lismN;
  return MFM_OK;
}

int mfm_params_set_swlism(mfm_params* p, double gamma, double lismN, double lismV,
                          double lismT, double lismB, const double sunXY[2], double R0,
                          double sunN, double sunV, double sunT, double initR, double netN,
                          double TMLIM, double dx_level0, const int num_cells[2]) {
  if (!p || !sunXY || !num_cells) return MFM_ERR_ARG;
  if (!(dx_level0 > 0.0) || num_cells[0] < 1 || num_cells[1] < 1) return MFM_ERR_ARG;
  p->problem = MFM_PROBLEM_SWLISM;
  p->coord_sys = MFM_COORD_AXISYMMETRIC;
  p->gamma = gamma;
  p->spacedim = 2;
  for (int d = 0; d < 3; ++d) {
    p->domain_lo[d] = 0;
    p->domain_hi[d] = d < 2 ? num_cells[d] - 1 : 0;
    p->periodic[d] = 0;
  }
  // the sun on a grid node (SWLISMProblem::input, :146-147); z unused in 2-D
  p->sun_xc = ((int)std::floor(sunXY[0] / dx_level0 + 0.5)) * dx_level0;
  p->sun_yc = ((int)std::floor(sunXY[1] / dx_level0 + 0.5)) * dx_level0;
  p->sun_zc = 0.0;
  p->sun_r0 = R0;
  p->r0_dt_ignore = R0 - dx_level0;  // m_R0dtIgnore (:149)
  p->init_r = initR;
  p->tmlim = TMLIM;
  p->lism_t = lismT;
  p->lism_v_dim = lismV;
  // SETSWLISM
  const double rho = lismN * eos_mp;
  const double pref = rho * lismV * lismV;
  p->lism_p = (2.0 * eos_k * lismN * lismT) / pref;
  p->lism_m = 1.0 / std::sqrt(gamma * p->lism_p);
  p->lism_bx = lismB / std::sqrt(pref);  // lismBX: the field along the flow
  p->lism_by = 0.0;
  p->lism_bz = 0.0;
  p->lism_ux = 1.0;
  p->lism_uy = 0.0;
  p->lism_uz = 0.0;
  p->sun_rho = sunN / lismN;
  p->sun_p = (2.0 * eos_k * sunN * sunT) / pref;
  p->sun_v = sunV / lismV;
  if (netN > 0.0) {
    p->net_rho = netN / lismN;
    p->net_p = (eos_k * netN * lismT) / pref;
  }
  const double r0 = 1.0;
  const double p0 = 1.0 / (gamma * p->lism_m * p->lism_m);
  const double umag = std::sqrt(gamma * p0 / r0);
  p->smallr = 1.0e-8 * r0;
  p->smallp = 1.0e-8 * r0 * 0.5 * umag * umag;
  if (p->smallr < 1.0e-12) p->smallr = 1.0e-12;
  if (p->smallp < 1.0e-12) p->smallp = 1.0e-12;
  // SETCHARGEEX_PARS(lismV, scaleLen = 1.5e13, lismN) (:240-247)
  p->ce_scale_vel = lismV;
  p->ce_sigma_lindsay = 1.0e-16 * 1.5e13 * lismN;
  return MFM_OK;
}

int mfm_num_components(const mfm_params* p) {
  if (!p) return MFM_ERR_ARG;
  return 8 + (p->dedner ? 1 : 0) + 5 * (p->nfluids - 1);
}

int mfm_num_ghost_cells(const mfm_params* p) {
  if (!p) return MFM_ERR_ARG;
  if (p->time_approx == MFM_TIME_RK2) return 4;
  if (p->time_approx == MFM_TIME_HANCOCK) return 3;
  return 2;
}

int mfm_ctx_create(const mfm_params* p, mfm_ctx** out) {
  if (!out) return MFM_ERR_ARG;
  *out = nullptr;
  Consts C;
  std::string why;
  int rc = build_consts(p, C, why);
  if (rc) {
    std::fprintf(stderr, "mfm_ctx_create: %s\n", why.c_str());
    return rc;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    std::fprintf(stderr, "mfm_ctx_create: no HIP device available\n");
    return MFM_ERR_DEVICE;
  }
  mfm_ctx* c = new mfm_ctx();
  c->p = *p;
  c->C = C;
  c->err.reserve(1024);  // error messages then need no allocation
  (void)hipGetDevice(&c->device);
  rc = mfm::host::fixed_scratch_setup(c);
  if (rc != MFM_OK) {
    std::fprintf(stderr, "mfm_ctx_create: %s\n", c->err.c_str());
    delete c;
    return rc;
  }
  *out = c;
  return MFM_OK;
}

int mfm_ctx_set_timing(mfm_ctx* c, int on) {
  if (!c) return MFM_ERR_ARG;
  if (on && !c->timer) {
    mfm::KernelTimer* T = new mfm::KernelTimer();
    for (int q = 0; q < 2 * mfm::KernelTimer::kMax; ++q) {
      hipError_t e = hipEventCreate(&T->ev[q]);
      if (e != hipSuccess) {
        for (int r = 0; r < q; ++r) (void)hipEventDestroy(T->ev[r]);
        delete T;
        return hipfail(c, e, "hipEventCreate");
      }
    }
    c->timer = T;
  } else if (!on && c->timer) {
    (void)hipDeviceSynchronize();
    for (int q = 0; q < 2 * mfm::KernelTimer::kMax; ++q) (void)hipEventDestroy(c->timer->ev[q]);
    delete c->timer;
    c->timer = nullptr;
  }
  for (int st = 0; st < 2; ++st)
    for (int k = 0; k < mfm::KK_COUNT; ++k) {
      c->kt_ms[st][k] = 0.0;
      c->kt_n[st][k] = 0;
      c->kt_items[st][k] = 0.0;
    }
  return MFM_OK;
}

int mfm_ctx_set_walk(mfm_ctx* c, int row_tile, int xcd_order) {
  if (!c || row_tile < 1 || row_tile > 64) return MFM_ERR_ARG;
  c->walk_tile = row_tile;
  c->walk_xcd = xcd_order ? 1 : 0;
  return MFM_OK;
}

int mfm_ctx_set_stage_kernel(mfm_ctx* c, int mode, int tz) {
  if (!c || tz < 1 || tz > 4096) return MFM_ERR_ARG;
  if (mode < MFM_STAGE_SEPARATE || mode > MFM_STAGE_AUTO) return MFM_ERR_ARG;
  c->stage_mode = mode;
  c->stage_tz = tz;
  return MFM_OK;
}

int mfm_ctx_get_walk(const mfm_ctx* c, int* row_tile, int* xcd_order) {
  if (!c) return MFM_ERR_ARG;
  if (row_tile) *row_tile = c->walk_tile;
  if (xcd_order) *xcd_order = c->walk_xcd;
  return MFM_OK;
}

int mfm_ctx_autotune(mfm_ctx* c, mfm_fab U, mfm_box valid, double dt, double dx,
                     double time, int reps, int* row_tile, int* xcd_order,
                     double* best_ms) {
  if (!c || !U.data || reps < 1) return MFM_ERR_ARG;
  const HBox ub = hb(U.box);
  const size_t bytes = (size_t)ub.size() * U.ncomp * sizeof(double);
  double* tmp = nullptr;
  hipError_t e = hipMalloc((void**)&tmp, bytes);
  if (e != hipSuccess) return hipfail(c, e, "hipMalloc (autotune copy)");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MFM_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = MFM_ERR_DEVICE;
  mfm_fab T = U;
  T.data = tmp;
  const int tiles[] = {2, 4, 8, 16, 32};
  int best_t = c->walk_tile, best_x = c->walk_xcd;
  double best = std::numeric_limits<double>::infinity();
  for (int x = 0; x < 2 && rc == MFM_OK; ++x)
    for (int t : tiles) {
      c->walk_tile = t;
      c->walk_xcd = x;
      // the same starting state for every candidate
      e = hipMemcpy(tmp, U.data, bytes, hipMemcpyDeviceToDevice);
      if (e != hipSuccess) { rc = hipfail(c, e, "autotune copy"); break; }
      double nd;
      int cell[3];
      rc = mfm_patch_update_dev(c, T, valid, dt, dx, time, nullptr, nullptr, &nd, cell, nullptr);
      if (rc != MFM_OK && rc != MFM_ERR_NUMERIC) break;
      rc = MFM_OK;
      (void)hipEventRecord(e0, nullptr);
      for (int r = 0; r < reps && rc == MFM_OK; ++r) {
        rc = mfm_patch_update_dev(c, T, valid, dt, dx, time, nullptr, nullptr, &nd, cell, nullptr);
        if (rc == MFM_ERR_NUMERIC) rc = MFM_OK;
      }
      (void)hipEventRecord(e1, nullptr);
      (void)hipEventSynchronize(e1);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (rc == MFM_OK && ms / reps < best) { best = ms / reps; best_t = t; best_x = x; }
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(tmp);
  c->walk_tile = best_t;
  c->walk_xcd = best_x;
  if (row_tile) *row_tile = best_t;
  if (xcd_order) *xcd_order = best_x;
  if (best_ms) *best_ms = best;
  return rc;
}

int mfm_ctx_set_graphs(mfm_ctx* c, int on) {
  if (!c) return MFM_ERR_ARG;
  c->graphs = on != 0;
  if (!c->graphs && c->gexec) {
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
  }
  return MFM_OK;
}

int mfm_ctx_get_timing(mfm_ctx* c, int stage, int kind, double* ms, long* launches,
                       double* items) {
  if (!c || stage < 0 || stage > 1 || kind < 0 || kind >= mfm::KK_COUNT) return MFM_ERR_ARG;
  if (ms) *ms = c->kt_ms[stage][kind];
  if (launches) *launches = c->kt_n[stage][kind];
  if (items) *items = c->kt_items[stage][kind];
  return MFM_OK;
}

int mfm_ctx_destroy(mfm_ctx* c) {
  if (!c) return MFM_OK;
  (void)hipDeviceSynchronize();
  if (c->timer) (void)mfm_ctx_set_timing(c, 0);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->cap_stream) (void)hipStreamDestroy(c->cap_stream);
  if (c->arena) (void)hipFree(c->arena);
  if (c->hbuf) (void)hipHostFree(c->hbuf);
  delete c;
  return MFM_OK;
}

int mfm_ctx_set_dedner(mfm_ctx* c, double ch, double cp) {
  if (!c) return MFM_ERR_ARG;
  if (!(ch > 0.0) || !(cp > 0.0)) return fail(c, MFM_ERR_ARG, "Ch and Cp must be positive");
  c->p.dedner_ch = ch;
  c->p.dedner_cp = cp;
  c->C.ch = ch;
  c->C.cp = cp;
  return MFM_OK;
}

const char* mfm_ctx_last_error(const mfm_ctx* c) {
  if (!c) return "null context";
  return c->err.c_str();
}

int mfm_patch_update_dev(mfm_ctx* c, mfm_fab U, mfm_box valid, double dt,
                         double dx, double time, mfm_fab* F, mfm_fab* divB,
                         double* new_dt, int min_cell[3], void* stream) {
  if (!c) return MFM_ERR_ARG;
  PatchIn in{U, valid, F, divB, 0, dt, dx, time};
  double d;
  int cell[3];
  int rc = run_batch(c, &c->C, 1, &in, 1, &d, cell, nullptr, nullptr, (hipStream_t)stream);
  if (new_dt) *new_dt = d;
  if (min_cell) { min_cell[0] = cell[0]; min_cell[1] = cell[1]; min_cell[2] = cell[2]; }
  return rc;
}

int mfm_level_update_dev(mfm_ctx* c, const mfm_patch* table, int n, double dt,
                         double dx, double time, double* new_dt, int* min_cell,
                         double* level_dt, int* min_patch, void* stream) {
  if (!c || !table) return MFM_ERR_ARG;
  std::vector<PatchIn> in(n);
  for (int q = 0; q < n; ++q)
    in[q] = PatchIn{table[q].U, table[q].valid, table[q].F, table[q].divB, 0, dt, dx, time};
  return run_batch(c, &c->C, 1, in.data(), n, new_dt, min_cell, level_dt, min_patch,
                   (hipStream_t)stream);
}

int mfm_level_check(const mfm_params* p, const mfm_patch* table, int n, double dt,
                    double dx, int* bad_patch, char* why, size_t why_len) {
  if (bad_patch) *bad_patch = -1;
  if (why && why_len) why[0] = 0;
  if (!p || !table || n <= 0) return MFM_ERR_ARG;
  Consts C;
  std::string w;
  int rc = build_consts(p, C, w);
  if (rc == MFM_OK) {
    const int ng = mfm_num_ghost_cells(p);
    for (int q = 0; q < n; ++q) {
      const PatchIn P{table[q].U, table[q].valid, table[q].F, table[q].divB, 0, dt, dx, 0.0};
      const char* m = nullptr;
      rc = check_patch(C, C, ng, P, &m);
      if (rc != MFM_OK) {
        if (bad_patch) *bad_patch = q;
        w = m;
        break;
      }
    }
  }
  if (rc != MFM_OK && why && why_len) std::snprintf(why, why_len, "%s", w.c_str());
  return rc;
}

int mfm_patch_update_host(mfm_ctx* c, mfm_fab U, mfm_box valid, double dt,
                          double dx, double time, mfm_fab* F, mfm_fab* divB,
                          double* new_dt, int min_cell[3]) {
  if (!c || !U.data) return MFM_ERR_ARG;
  const int dims = c->C.dims;
  HBox ub = hb(U.box);
  // sizes of the staged copies come from the host FABs: refuse empty or
  // negative ones before any allocation (the device path checks the rest)
  if (U.ncomp <= 0 || ub.size() == 0) return fail(c, MFM_ERR_SHAPE, "U: empty FAB");
  for (int d = 0; F && d < dims; ++d)
    if (!F[d].data || F[d].ncomp <= 0 || hb(F[d].box).size() == 0)
      return fail(c, MFM_ERR_SHAPE, "F[d]: empty FAB");
  if (divB && (!divB->data || divB->ncomp <= 0 || hb(divB->box).size() == 0))
    return fail(c, MFM_ERR_SHAPE, "divB: empty FAB");
  size_t ub_bytes = (size_t)ub.size() * U.ncomp * sizeof(double);
  double* dU = nullptr;
  std::vector<double*> dF(3, nullptr);
  double* dD = nullptr;
  mfm_fab Fd[3];
  mfm_fab Dd;
  int rc = MFM_OK;
  hipError_t e = hipMalloc((void**)&dU, ub_bytes);
  if (e != hipSuccess) return fail(c, MFM_ERR_NOMEM, "hipMalloc(U)");
  e = hipMemcpy(dU, U.data, ub_bytes, hipMemcpyHostToDevice);
  if (F) {
    for (int d = 0; d < dims && e == hipSuccess; ++d) {
      size_t b = (size_t)hb(F[d].box).size() * F[d].ncomp * sizeof(double);
      e = hipMalloc((void**)&dF[d], b);
      // the caller's values travel both ways, so cells of F outside the
      // faces of `valid` come back unchanged, as with device FABs
      if (e == hipSuccess) e = hipMemcpy(dF[d], F[d].data, b, hipMemcpyHostToDevice);
      Fd[d] = F[d];
      Fd[d].data = dF[d];
    }
  }
  if (divB && e == hipSuccess) {
    size_t b = (size_t)hb(divB->box).size() * divB->ncomp * sizeof(double);
    e = hipMalloc((void**)&dD, b);
    if (e == hipSuccess) e = hipMemcpy(dD, divB->data, b, hipMemcpyHostToDevice);
    Dd = *divB;
    Dd.data = dD;
  }
  if (e != hipSuccess) rc = hipfail(c, e, "staging");
  if (!rc) {
    mfm_fab Ud = U;
    Ud.data = dU;
    rc = mfm_patch_update_dev(c, Ud, valid, dt, dx, time, F ? Fd : nullptr,
                              divB ? &Dd : nullptr, new_dt, min_cell, nullptr);
  }
  if (!rc || rc == MFM_ERR_NUMERIC) {
    e = hipMemcpy(U.data, dU, ub_bytes, hipMemcpyDeviceToHost);
    for (int d = 0; F && d < dims && e == hipSuccess; ++d)
      e = hipMemcpy(F[d].data, dF[d], (size_t)hb(F[d].box).size() * F[d].ncomp * sizeof(double),
                    hipMemcpyDeviceToHost);
    if (divB && e == hipSuccess)
      e = hipMemcpy(divB->data, dD, (size_t)hb(divB->box).size() * divB->ncomp * sizeof(double),
                    hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hipfail(c, e, "copy back");
  }
  (void)hipFree(dU);
  for (int d = 0; d < 3; ++d)
    if (dF[d]) (void)hipFree(dF[d]);
  if (dD) (void)hipFree(dD);
  return rc;
}

int mfm_patch_compute_dt_dev(mfm_ctx* c, mfm_fab U, mfm_box box, double dx,
                             double* dt, int min_cell[3], void* stream) {
  if (!c) return MFM_ERR_ARG;
  return run_dt(c, c->C, U, box, dx, dt, min_cell, (hipStream_t)stream);
}

int mfm_init_problem_dev(mfm_ctx* c, mfm_fab U, mfm_box box, double dx,
                         void* stream) {
  if (!c || !U.data) return MFM_ERR_ARG;
  if (U.ncomp != c->C.nc) return fail(c, MFM_ERR_SHAPE, "U.ncomp mismatch");
  HBox ub = hb(U.box), b = hb(box);
  if (!contains(ub, b)) return fail(c, MFM_ERR_SHAPE, "box outside U");
  if (ub.size() >= (1L << 31)) return fail(c, MFM_ERR_SHAPE, "FAB of 2^31 cells or more");
  if (b.size() == 0) return MFM_OK;  // an empty box: nothing to set
  int ulo[3], un[3], lo[3], n[3];
  for (int d = 0; d < 3; ++d) { ulo[d] = ub.lo[d]; un[d] = (int)ub.n(d); lo[d] = b.lo[d]; n[d] = (int)b.n(d); }
  hipError_t e = launch_init(c->C, U.data, ulo, un, ub.size(), lo, n, dx, (hipStream_t)stream);
  if (e != hipSuccess) return hipfail(c, e, "init launch");
  return MFM_OK;
}

int mfm_fill_periodic_dev(mfm_ctx* c, mfm_fab U, mfm_box valid, void* stream) {
  if (!c || !U.data) return MFM_ERR_ARG;
  HBox ub = hb(U.box), v = hb(valid);
  // the kernel takes every index modulo the valid extent: it must be >= 1
  if (v.size() == 0 || U.ncomp <= 0) return fail(c, MFM_ERR_SHAPE, "empty valid box or FAB");
  if (!contains(ub, v)) return fail(c, MFM_ERR_SHAPE, "valid outside U");
  if (ub.size() >= (1L << 31)) return fail(c, MFM_ERR_SHAPE, "FAB of 2^31 cells or more");
  int ulo[3], un[3], vlo[3], vn[3];
  for (int d = 0; d < 3; ++d) { ulo[d] = ub.lo[d]; un[d] = (int)ub.n(d); vlo[d] = v.lo[d]; vn[d] = (int)v.n(d); }
  hipError_t e = launch_fill_periodic(U.data, ulo, un, ub.size(), U.ncomp, vlo, vn,
                                      c->C.dims, (hipStream_t)stream);
  if (e != hipSuccess) return hipfail(c, e, "periodic fill launch");
  return MFM_OK;
}

}  // extern "C"

// ---- device utilities ------------------------------------------------------
extern "C" {

int mfm_dev_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int mfm_dev_alloc(size_t bytes, void** ptr) {
  if (!ptr) return MFM_ERR_ARG;
  *ptr = nullptr;
  if (hipMalloc(ptr, bytes) != hipSuccess) return MFM_ERR_NOMEM;
  return MFM_OK;
}

int mfm_dev_free(void* ptr) {
  if (!ptr) return MFM_OK;
  return hipFree(ptr) == hipSuccess ? MFM_OK : MFM_ERR_DEVICE;
}

int mfm_dev_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  if (kind < 0 || kind > 2) return MFM_ERR_ARG;
  return hipMemcpy(dst, src, bytes, k) == hipSuccess ? MFM_OK : MFM_ERR_DEVICE;
}

int mfm_dev_memset(void* ptr, int value, size_t bytes) {
  return hipMemset(ptr, value, bytes) == hipSuccess ? MFM_OK : MFM_ERR_DEVICE;
}

int mfm_dev_synchronize(void) {
  return hipDeviceSynchronize() == hipSuccess ? MFM_OK : MFM_ERR_DEVICE;
}

int mfm_event_create(void** ev) {
  if (!ev) return MFM_ERR_ARG;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return MFM_ERR_DEVICE;
  *ev = (void*)e;
  return MFM_OK;
}

int mfm_event_destroy(void* ev) {
  return hipEventDestroy((hipEvent_t)ev) == hipSuccess ? MFM_OK : MFM_ERR_DEVICE;
}

int mfm_event_record(void* ev, void* stream) {
  return hipEventRecord((hipEvent_t)ev, (hipStream_t)stream) == hipSuccess ? MFM_OK
                                                                          : MFM_ERR_DEVICE;
}

int mfm_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!ms) return MFM_ERR_ARG;
  if (hipEventSynchronize((hipEvent_t)stop) != hipSuccess) return MFM_ERR_DEVICE;
  return hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop) == hipSuccess
             ? MFM_OK : MFM_ERR_DEVICE;
}

}  // extern "C"
