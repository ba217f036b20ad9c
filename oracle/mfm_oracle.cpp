// mfm_oracle.cpp -- CPU FP64 restatement of MHDAM's per-patch Godunov update.
//
// TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline).  The
// product library never links or calls this file.
//
// It follows the reference's structure on purpose: full-box passes over
// FArrayBox-shaped temporaries (W, W-/W+ per direction, F per direction, S),
// in the same operation order as the ChomboFortran kernels, so that it can be
// compared against both the reference's semantics and the fused HIP kernels.
// Paths are relative to MHDAM-3.2.4/ of the reference.
//
// Pinned: every function on the path has a check of its own in tests/ --
// closed forms, invariants, mutation-tested boundary and ghost-fill pins,
// and the reference's own fixtures where it holds them.
#include "mfm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// srcIdeal/EosCommon.fh:13-37
const double d_1_4PI = 0.079577471545947667884441881686257181;
const double d_1_8PI = 0.039788735772973833942220940843128591;
const double d_SQRT_2 = 0.707106781186547524400844362104849039;
const double d_PI = 3.14159265358979323846264338327950288;
const double d_2PI = 6.28318530717958647692528676655900576;
const double d_4PI = 12.566370614359172953850573533118;
const double d_PI_2 = 1.57079632679489661923132169163975144;
const double eos_AU = 1.5e13;
const double netFluid2Dens = 1.0e-6, netFluid2Pres = 1.0e-8;
const double netFluid3Dens = 1.0e-7, netFluid3Pres = 1.0e-9;
const double half = 0.5, one = 1.0, two = 2.0, four = 4.0, zero = 0.0;

// srcIdeal/LGintegrator.H:10-46
enum { WRHO = 0, WVELX = 1, WVELY = 2, WVELZ = 3, WPRES = 4, WBX = 5, WBY = 6,
       WBZ = 7 };
enum { URHO = 0, UMOMX = 1, UMOMY = 2, UMOMZ = 3, UENG = 4, UBX = 5, UBY = 6,
       UBZ = 7 };
const int WNUM_E = 5;

// ------------------------------------------------------------------ boxes --
struct Box {
  int lo[3], hi[3];
  bool empty() const {
    for (int d = 0; d < 3; ++d)
      if (hi[d] < lo[d]) return true;
    return false;
  }
  long n(int d) const { return hi[d] - lo[d] + 1; }
  long size() const { return empty() ? 0 : n(0) * n(1) * n(2); }
};
Box mkbox(const mfm_box& b) {
  Box r;
  for (int d = 0; d < 3; ++d) { r.lo[d] = b.lo[d]; r.hi[d] = b.hi[d]; }
  return r;
}
Box grow(Box b, int n, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] -= n; b.hi[d] += n; }
  return b;
}
Box grow_dir(Box b, int d, int n) { b.lo[d] -= n; b.hi[d] += n; return b; }
bool contains(const Box& a, const Box& b) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < a.lo[d] || b.hi[d] > a.hi[d]) return false;
  return true;
}

// FArrayBox: component-major, x fastest.
struct Fab {
  std::vector<double> own;
  double* p = nullptr;
  Box b{};
  int nc = 0;
  long s1 = 0, s2 = 0, sc = 0;
  void setup() {
    s1 = b.n(0);
    s2 = s1 * b.n(1);
    sc = s2 * b.n(2);
  }
  void define(const Box& bx, int ncomp, double fill) {
    b = bx; nc = ncomp; setup();
    own.assign((size_t)sc * nc, fill);
    p = own.data();
  }
  void wrap(const mfm_fab& f) {
    b = mkbox(f.box); nc = f.ncomp; setup(); p = f.data;
  }
  double& operator()(int i, int j, int k, int c) {
    return p[(i - b.lo[0]) + s1 * (j - b.lo[1]) + s2 * (k - b.lo[2]) + sc * c];
  }
  double operator()(int i, int j, int k, int c) const {
    return p[(i - b.lo[0]) + s1 * (j - b.lo[1]) + s2 * (k - b.lo[2]) + sc * c];
  }
};
struct IFab {
  std::vector<int> v;
  Box b{};
  long s1 = 0, s2 = 0;
  void define(const Box& bx) {
    b = bx; s1 = b.n(0); s2 = s1 * b.n(1);
    v.assign((size_t)b.size(), 0);
  }
  int& operator()(int i, int j, int k) {
    return v[(i - b.lo[0]) + s1 * (j - b.lo[1]) + s2 * (k - b.lo[2])];
  }
};

#define FORBOX(bx, i, j, k)                          \
  for (int k = (bx).lo[2]; k <= (bx).hi[2]; ++k)     \
    for (int j = (bx).lo[1]; j <= (bx).hi[1]; ++j)   \
      for (int i = (bx).lo[0]; i <= (bx).hi[0]; ++i)

// OpenMP over z planes (the CPU-baseline parallelisation); results do not
// depend on the thread count because every loop body writes only its own cell.
#define PFORBOX(bx, i, j, k)                                            \
  _Pragma("omp parallel for schedule(static)")                        \
  for (int k = (bx).lo[2]; k <= (bx).hi[2]; ++k)                        \
    for (int j = (bx).lo[1]; j <= (bx).hi[1]; ++j)                      \
      for (int i = (bx).lo[0]; i <= (bx).hi[0]; ++i)

// --------------------------------------------------- equation-system state --
// Replaces the COMMON blocks /eos/, /RiemSol/, /Dedner/ and the EqSysMHDMF
// index maps (srcIdeal/EqSysMHDMF.cpp:11-48, 257-306).
struct Eq {
  const mfm_params* P;
  int dims, nf, nc, dedner, iPsi, iRhoN;
  double gamma, hgamma, invhgamma, smallr, smallp, del2, smallB;
  double ch, cp;
  int iAveraging, iLaxFriedrix;
  explicit Eq(const mfm_params* p, double ch_, double cp_) : P(p) {
    dims = p->spacedim;
    nf = p->nfluids;
    dedner = p->dedner ? 1 : 0;
    nc = 8 + dedner + 5 * (nf - 1);
    iPsi = dedner ? 8 : -1;
    iRhoN = 8 + dedner;
    gamma = p->gamma;
    hgamma = gamma - 1.0;
    invhgamma = one / hgamma;
    smallr = p->smallr;
    smallp = p->smallp;
    del2 = p->entropy_fix;
    smallB = p->small_b;
    ch = ch_;
    cp = cp_;
    // src/RiemannSolver.cpp:36-105 (SETRSCONST arguments)
    iAveraging = (p->flux == 3 || p->flux == 4 || p->flux == 13) ? 1 : 0;
    iLaxFriedrix = 0;
  }
  int limiter(int var) const {
    // main/MHDAM.cpp:927-1000: minmod for all slopes, then the plasma
    // primitive limiters replace RHO..BZ.
    if (var < 8) return P->limiter[var];
    return MFM_LIM_MINMOD;
  }
  bool physical(int d) const {
    return P->problem != MFM_PROBLEM_NONE && !P->periodic[d];
  }
  Box clip(Box b) const {  // Box &= ProblemDomain (periodic dirs unclipped)
    for (int d = 0; d < dims; ++d)
      if (physical(d)) {
        b.lo[d] = std::max(b.lo[d], P->domain_lo[d]);
        b.hi[d] = std::min(b.hi[d], P->domain_hi[d]);
      }
    return b;
  }
};

// =============================================================== limiters ==
// src/reconstructionF.ChF
inline double fsign1(double a) { return std::copysign(1.0, a); }
inline double minmod(double A, double B) {  // :4-14
  double signA = fsign1(A);
  return signA * std::max(zero, std::min(std::fabs(A), signA * B));
}

// Face values of a cell from its two neighbours; left = W- (lo face),
// right = W+ (hi face).
void limit(int lim, double WL_, double WP, double WR_, double* left,
           double* right) {
  double dWR = WR_ - WP, dWL = WP - WL_;
  switch (lim) {
    case MFM_LIM_NONE: *left = WP; *right = WP; return;  // :28-44
    case MFM_LIM_CD:  // CDFACEVALUES :1056-1079 (W- takes the +1 neighbour)
      *left = half * (WP + WR_);
      *right = half * (WP + WL_);
      return;
    case MFM_LIM_MINMOD: {  // :126-155
      double SLOPE = half * minmod(dWR, dWL);
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_SUPERBEE: {  // :258-300
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double AR = std::fabs(dWR), AL = std::fabs(dWL), SLOPE;
      if (AL <= half * AR) SLOPE = dWL;
      else if (AL <= AR) SLOPE = half * dWR;
      else if (AL <= AR + AR) SLOPE = half * dWL;
      else SLOPE = dWR;
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_HYPERBEE: {  // :430-464
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double A = one + std::sqrt(dWR / dWL);
      double SLOPE = dWR / (A * A);
      SLOPE = SLOPE + SLOPE;
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_HARMONIC: {  // :492-524
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double SLOPE = dWR / (one + dWR / dWL);
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_VANALBADA: {  // :550-589
      const double dSave = 1.0e-6, dSaveS = dSave + dSave;
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double SLOPE = half * (dWR * dWL + dSave) * (dWR + dWL) /
                     (dWR * dWR + dWL * dWL + dSaveS);
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_MC: {  // :622-662
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double AR = std::fabs(dWR), AL = std::fabs(dWL), SLOPE;
      if (3.0 * AL <= AR) SLOPE = dWL;
      else if (AL <= 3.0 * AR) SLOPE = 0.25 * (dWR + dWL);
      else SLOPE = dWR;
      *left = WP - SLOPE; *right = WP + SLOPE;
      return;
    }
    case MFM_LIM_VANLEER: {  // :794-841, kappa = 1/3
      const double dSLim = one / 3.0;
      const double BT = (3.0 - dSLim) / (one - dSLim);
      const double SL1 = 0.25 * (one - dSLim), SL2 = 0.25 * (one + dSLim);
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double SL = std::min(std::fabs(dWR), std::fabs(BT * dWL));
      double SR = std::min(std::fabs(dWL), std::fabs(BT * dWR));
      double SLOPL = SL1 * SL + SL2 * SR;
      double SLOPR = SL1 * SR + SL2 * SL;
      if (dWR < zero) { SLOPL = -SLOPL; SLOPR = -SLOPR; }
      *left = WP - SLOPL; *right = WP + SLOPR;
      return;
    }
    case MFM_LIM_KOREN: {  // :883-918
      const double dSave = 1.0e-10;
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double H = dWL * dWR / (two * (dWL * dWL + dWR * dWR + dSave) - dWL * dWR);
      *left = WP - (dWL + half * dWR) * H;
      *right = WP + (dWR + half * dWL) * H;
      return;
    }
    case MFM_LIM_VENKAT: {  // :950-988
      const double dSave = 1.0e-10, dSave2 = dSave * dSave;
      if (dWL * dWR <= zero) { *left = WP; *right = WP; return; }
      double H = (dWR + dWL) / (6.0 * (dWR * dWR + dWL * dWL + dSave2));
      *left = WP - (dWR + two * dWL) * (dWL + dSave) * H;
      *right = WP + (dWL + two * dWR) * (dWR + dSave) * H;
      return;
    }
    case MFM_LIM_WENO3: {  // :1096-1136
      const double dSave = 1.0e-6;
      double BR = dWR * dWR + dSave, AR = one / (BR * BR);
      double BL = dWL * dWL + dSave, AL = one / (BL * BL);
      double H = AR / (AR + two * AL);
      *left = WP - half * (dWL + H * (dWR - dWL));
      H = AL / (AL + two * AR);
      *right = WP + half * (dWR - H * (dWR - dWL));
      return;
    }
    case MFM_LIM_WENO3YC: {  // :1169-1212
      const double dSave = 1.0e-6;
      double BR = dWR * dWR + dSave, BL = dWL * dWL + dSave;
      double D2P = dWR - dWL;
      D2P = D2P * D2P;
      double AL = one + D2P / BL, AR = one + D2P / BR;
      double H = AR / (AR + two * AL);
      *left = WP - half * (dWL + H * (dWR - dWL));
      H = AL / (AL + two * AR);
      *right = WP + half * (dWR - H * (dWR - dWL));
      return;
    }
    default: *left = WP; *right = WP; return;
  }
}

// ======================================================= cons <-> prim =====
// CONSTOPRIM (srcIdeal/PatchIdealMHDF.ChF:143-189) for one fluid;
// CONSTOPRIM_MF (srcIdeal/PatchMHDMFF.ChF:5-80) floors rho for every fluid.
void cons_to_prim_cell(const Eq& E, const double* u, double* w) {
  double rho = u[URHO];
  if (E.nf > 1) rho = std::max(u[URHO], E.smallr);
  double ru = u[UMOMX], rv = u[UMOMY], rw = u[UMOMZ];
  double invrho = one / rho;
  double uu = ru * invrho, vv = rv * invrho, ww = rw * invrho;
  double ke = ru * uu + rv * vv + rw * ww;
  double bx = u[UBX], by = u[UBY], bz = u[UBZ];
  double b = d_1_4PI * (bx * bx + by * by + bz * bz);
  double p = E.hgamma * (u[UENG] - half * (ke + b));
  p = std::max(p, E.smallp);
  w[WVELX] = uu; w[WVELY] = vv; w[WVELZ] = ww;
  w[WRHO] = rho; w[WPRES] = p;
  w[WBX] = bx; w[WBY] = by; w[WBZ] = bz;
  if (E.dedner) w[E.iPsi] = u[E.iPsi];  // EqSysMHDMF::stateToPrim :194-197
  for (int f = 1; f < E.nf; ++f) {
    int s = E.iRhoN + (f - 1) * WNUM_E;
    double r = std::max(u[s + URHO], E.smallr);
    double nru = u[s + UMOMX], nrv = u[s + UMOMY], nrw = u[s + UMOMZ];
    double ir = one / r;
    double nu = nru * ir, nv = nrv * ir, nw = nrw * ir;
    double nke = nru * nu + nrv * nv + nrw * nw;
    double np = E.hgamma * (u[s + UENG] - half * nke);
    np = std::max(np, E.smallp);
    w[s + WVELX] = nu; w[s + WVELY] = nv; w[s + WVELZ] = nw;
    w[s + WRHO] = r; w[s + WPRES] = np;
  }
}

void cons_to_prim(const Eq& E, Fab& W, const Fab& U, const Box& b) {
  PFORBOX(b, i, j, k) {
    double u[32], w[32];
    for (int c = 0; c < E.nc; ++c) u[c] = U(i, j, k, c);
    cons_to_prim_cell(E, u, w);
    for (int c = 0; c < E.nc; ++c) W(i, j, k, c) = w[c];
  }
}

// ====================================================== Riemann solvers ====
// Plasma primitive state in face-rotated order.
struct PState { double p, rho, u, v, w, bx, by, bz; };

// Body of RIEMANNF / ROE8WAVEF after the input clamps
// (src/Roe8WaveF.ChF:99-588 and :666-1141).  out: rho, mom_n, mom_t1,
// mom_t2, E, B_n, B_t1, B_t2.
void roe8_core(const Eq& E, int iAveraging, int iLaxFriedrix, double PGasL,
               double RhoL, double UL, double VL, double WL, double BXL,
               double BYL, double BZL, double PGasR, double RhoR, double UR,
               double VR, double WR, double BXR, double BYR, double BZR,
               double* out) {
  const double gamma = E.gamma, hgamma = E.hgamma, del2 = E.del2,
               smallB = E.smallB;
  double RUL = RhoL * UL, RVL = RhoL * VL, RWL = RhoL * WL;
  double BxBL = d_1_8PI * (BXL * BXL + BYL * BYL + BZL * BZL);
  double RUxUL = half * (RUL * UL + RVL * VL + RWL * WL);
  double PL = PGasL + BxBL;
  double EL = PGasL / hgamma + RUxUL + BxBL;
  double AL2 = gamma * PGasL / RhoL;
  double RUR = RhoR * UR, RVR = RhoR * VR, RWR = RhoR * WR;
  double BxBR = d_1_8PI * (BXR * BXR + BYR * BYR + BZR * BZR);
  double RUxUR = half * (RUR * UR + RVR * VR + RWR * WR);
  double PR = PGasR + BxBR;
  double ER = PGasR / hgamma + RUxUR + BxBR;
  double AR2 = gamma * PGasR / RhoR;
  double ccmi, ccma;
  if (AL2 > AR2) { ccmi = AR2; ccma = AL2; } else { ccmi = AL2; ccma = AR2; }

  double Rho, U, V, W, BX, BY, BZ, BX2, BY2, BZ2, BYZ2, BxBH, kk, cc;
  if (iAveraging == 1) {
    double HL = (EL + PL) / RhoL, HR = (ER + PR) / RhoR;
    double coefRL = one / (one + std::sqrt(RhoR / RhoL));
    double coefRR = one - coefRL;
    Rho = RhoR * coefRL + RhoL * coefRR;
    U = UL * coefRL + UR * coefRR;
    V = VL * coefRL + VR * coefRR;
    W = WL * coefRL + WR * coefRR;
    double HAV = HL * coefRL + HR * coefRR;
    BX = BXR * coefRL + BXL * coefRR;
    BY = BYR * coefRL + BYL * coefRR;
    BZ = BZR * coefRL + BZL * coefRR;
    double BXAV = half * (BXL + BXR);
    BX2 = BXAV * BXAV; BY2 = BY * BY; BZ2 = BZ * BZ;
    BYZ2 = BY2 + BZ2;
    BxBH = d_1_8PI * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    double DBX = BXR - BXL, DBY = BYR - BYL, DBZ = BZR - BZL;
    double X = d_1_8PI * (DBX * DBX + DBY * DBY + DBZ * DBZ) * coefRL * coefRL /
               RhoL;
    cc = hgamma * (HAV - kk - d_1_4PI * (BX * BXAV + BYZ2) / Rho) +
         (two - gamma) * X;
  } else {
    Rho = half * (RhoL + RhoR);
    U = half * (UL + UR); V = half * (VL + VR); W = half * (WL + WR);
    BX = half * (BXL + BXR); BY = half * (BYL + BYR); BZ = half * (BZL + BZR);
    double P = half * (PL + PR);
    BX2 = BX * BX; BY2 = BY * BY; BZ2 = BZ * BZ;
    BYZ2 = BY2 + BZ2;
    BxBH = d_1_8PI * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    double PGas = P - BxBH;
    cc = gamma * PGas / Rho;
  }
  double bb = d_1_4PI * BX2 / Rho;
  double b = std::sqrt(bb);
  if (cc < ccmi) cc = ccmi;
  if (cc > ccma) cc = ccma;
  double c = std::sqrt(cc);

  double hyy, hzz, aaf, aas, af, as;
  if (BYZ2 < smallB * smallB) {
    hyy = d_SQRT_2; hzz = d_SQRT_2;
    if (bb > cc) { aaf = bb; aas = cc; af = b; as = c; }
    else { aaf = cc; aas = bb; af = c; as = b; }
  } else {
    double hyz = std::sqrt(BYZ2);
    hyy = BY / hyz; hzz = BZ / hyz;
    double szb = cc + (BxBH + BxBH) / Rho;
    double h = std::sqrt(szb * szb - 4.0 * cc * bb);
    aas = half * (szb - h);
    if (aas <= zero) { aas = zero; as = zero; }
    else if (aas >= cc) { aas = cc; as = c; }
    else as = std::sqrt(aas);
    aaf = half * (szb + h);
    if (aaf <= cc) { aaf = cc; af = c; }
    else af = std::sqrt(aaf);
  }

  double A[8];
  if (iLaxFriedrix == 1) {
    double eigmax = std::fabs(U) + af;
    for (int q = 0; q < 8; ++q) A[q] = eigmax;
  } else {
    double eig[8] = {U, U, U + b, U - b, U + af, U - af, U + as, U - as};
    double eigmax = std::fabs(U) + std::fabs(V) + std::fabs(W) + af;
    double del = del2 * eigmax;
    double dell2 = del * del;
    for (int q = 0; q < 8; ++q) {
      double z = std::fabs(eig[q]);
      if (z >= del) A[q] = z;
      else A[q] = half * (z * z + dell2) / del;
    }
  }

  double dfs = aaf - aas, alf, als;
  if (dfs < 1.0e-8) { alf = one; als = zero; }
  else if ((cc - aas) <= zero) { alf = zero; als = one; }
  else if ((aaf - cc) <= zero) { alf = one; als = zero; }
  else { alf = std::sqrt((cc - aas) / dfs); als = std::sqrt((aaf - cc) / dfs); }

  double sih = -one;
  if (BX >= zero) sih = one;

  double fl[8], fr[8], du[8];
  fl[0] = RUL;
  fl[1] = UL * RUL - BXL * BXL * d_1_4PI + PL;
  fl[2] = UL * RVL - BXL * BYL * d_1_4PI;
  fl[3] = UL * RWL - BXL * BZL * d_1_4PI;
  fl[4] = UL * (EL + PL) - BXL * (UL * BXL + VL * BYL + WL * BZL) * d_1_4PI;
  fl[5] = zero;
  fl[6] = UL * BYL - BXL * VL;
  fl[7] = UL * BZL - BXL * WL;
  fr[0] = RUR;
  fr[1] = UR * RUR - BXR * BXR * d_1_4PI + PR;
  fr[2] = UR * RVR - BXR * BYR * d_1_4PI;
  fr[3] = UR * RWR - BXR * BZR * d_1_4PI;
  fr[4] = UR * (ER + PR) - BXR * (UR * BXR + VR * BYR + WR * BZR) * d_1_4PI;
  fr[5] = zero;
  fr[6] = UR * BYR - BXR * VR;
  fr[7] = UR * BZR - BXR * WR;
  du[0] = RhoL - RhoR; du[1] = RUL - RUR; du[2] = RVL - RVR; du[3] = RWL - RWR;
  du[4] = EL - ER; du[5] = BXL - BXR; du[6] = BYL - BYR; du[7] = BZL - BZR;
  du[4] = -hgamma * (-kk * du[0] + U * du[1] + V * du[2] + W * du[3] - du[4] +
                     (BX * du[5] + BY * du[6] + BZ * du[7]) * d_1_4PI);
  du[1] = (-U * du[0] + du[1]) / Rho;
  du[2] = (-V * du[0] + du[2]) / Rho;
  du[3] = (-W * du[0] + du[3]) / Rho;
  double sqPiRho = std::sqrt(d_PI * Rho), sq2PiRho = std::sqrt(d_2PI * Rho);

  double sp[8][8], so[8][8];
  // right eigenvectors, src/Roe8WaveF.ChF:393-463
  sp[0][0] = one; sp[0][1] = zero; sp[0][2] = zero; sp[0][3] = zero;
  sp[0][4] = Rho * alf; sp[0][5] = Rho * alf; sp[0][6] = Rho * als; sp[0][7] = Rho * als;
  sp[1][0] = zero; sp[1][1] = zero; sp[1][2] = zero; sp[1][3] = zero;
  sp[1][4] = alf * af; sp[1][5] = -alf * af; sp[1][6] = als * as; sp[1][7] = -als * as;
  sp[2][0] = zero; sp[2][1] = zero;
  sp[2][2] = -hzz * d_SQRT_2; sp[2][3] = -hzz * d_SQRT_2;
  sp[2][4] = -als * as * hyy * sih; sp[2][5] = als * as * hyy * sih;
  sp[2][6] = alf * af * hyy * sih; sp[2][7] = -alf * af * hyy * sih;
  sp[3][0] = zero; sp[3][1] = zero;
  sp[3][2] = hyy * d_SQRT_2; sp[3][3] = hyy * d_SQRT_2;
  sp[3][4] = -als * as * hzz * sih; sp[3][5] = als * as * hzz * sih;
  sp[3][6] = alf * af * hzz * sih; sp[3][7] = -alf * af * hzz * sih;
  sp[4][0] = zero; sp[4][1] = zero; sp[4][2] = zero; sp[4][3] = zero;
  sp[4][4] = alf * Rho * cc; sp[4][5] = alf * Rho * cc;
  sp[4][6] = als * Rho * cc; sp[4][7] = als * Rho * cc;
  sp[5][0] = zero; sp[5][1] = one;
  for (int q = 2; q < 8; ++q) sp[5][q] = zero;
  sp[6][0] = zero; sp[6][1] = zero;
  sp[6][2] = hzz * sq2PiRho * sih; sp[6][3] = -hzz * sq2PiRho * sih;
  sp[6][4] = two * als * sqPiRho * c * hyy; sp[6][5] = two * als * sqPiRho * c * hyy;
  sp[6][6] = -two * alf * sqPiRho * c * hyy; sp[6][7] = -two * alf * sqPiRho * c * hyy;
  sp[7][0] = zero; sp[7][1] = zero;
  sp[7][2] = -hyy * sq2PiRho * sih; sp[7][3] = hyy * sq2PiRho * sih;
  sp[7][4] = two * als * sqPiRho * c * hzz; sp[7][5] = two * als * sqPiRho * c * hzz;
  sp[7][6] = -two * alf * sqPiRho * c * hzz; sp[7][7] = -two * alf * sqPiRho * c * hzz;
  // left eigenvectors, src/Roe8WaveF.ChF:465-527
  for (int r = 0; r < 8; ++r)
    for (int q = 0; q < 8; ++q) so[r][q] = zero;
  so[0][0] = one; so[0][4] = -one / cc;
  so[1][5] = one;
  so[2][2] = -hzz * d_SQRT_2; so[2][3] = hyy * d_SQRT_2;
  so[2][6] = half * hzz * sih / sq2PiRho; so[2][7] = -half * hyy * sih / sq2PiRho;
  so[3][2] = -hzz * d_SQRT_2; so[3][3] = hyy * d_SQRT_2;
  so[3][6] = -half * hzz * sih / sq2PiRho; so[3][7] = half * hyy * sih / sq2PiRho;
  so[4][1] = half * alf * af / cc; so[4][2] = -half * als * as * hyy * sih / cc;
  so[4][3] = -half * als * as * hzz * sih / cc; so[4][4] = half * alf / (cc * Rho);
  so[4][6] = 0.25 * als * hyy / (c * sqPiRho); so[4][7] = 0.25 * als * hzz / (c * sqPiRho);
  so[5][1] = -half * alf * af / cc; so[5][2] = half * als * as * hyy * sih / cc;
  so[5][3] = half * als * as * hzz * sih / cc; so[5][4] = half * alf / (cc * Rho);
  so[5][6] = 0.25 * als * hyy / (c * sqPiRho); so[5][7] = 0.25 * als * hzz / (c * sqPiRho);
  so[6][1] = half * als * as / cc; so[6][2] = half * alf * af * hyy * sih / cc;
  so[6][3] = half * alf * af * hzz * sih / cc; so[6][4] = half * als / (Rho * cc);
  so[6][6] = -0.25 * alf * hyy / (c * sqPiRho); so[6][7] = -0.25 * alf * hzz / (c * sqPiRho);
  so[7][1] = -half * als * as / cc; so[7][2] = -half * alf * af * hyy * sih / cc;
  so[7][3] = -half * alf * af * hzz * sih / cc; so[7][4] = half * als / (Rho * cc);
  so[7][6] = -0.25 * alf * hyy / (c * sqPiRho); so[7][7] = -0.25 * alf * hzz / (c * sqPiRho);

  double lu[8], ra[8];
  for (int r = 0; r < 8; ++r) {
    double s = so[r][0] * du[0];
    for (int q = 1; q < 8; ++q) s = s + so[r][q] * du[q];
    lu[r] = A[r] * s;
  }
  for (int r = 0; r < 8; ++r) {
    double s = sp[r][0] * lu[0];
    for (int q = 1; q < 8; ++q) s = s + sp[r][q] * lu[q];
    ra[r] = s;
  }
  ra[4] = ra[0] * kk + Rho * (U * ra[1] + V * ra[2] + W * ra[3]) + ra[4] / hgamma +
          d_1_4PI * (BX * ra[5] + BY * ra[6] + BZ * ra[7]);
  ra[1] = U * ra[0] + Rho * ra[1];
  ra[2] = V * ra[0] + Rho * ra[2];
  ra[3] = W * ra[0] + Rho * ra[3];
  for (int q = 0; q < 8; ++q) out[q] = half * (fl[q] + fr[q] + ra[q]);
}

// RIEMANN7F (src/Roe7WaveF.ChF:5-558): 7-wave Roe without the B_n wave;
// out[5] (B_n flux) = 0.  The (B_n, psi) block is DEDNER2X2F's.
void roe7_core(const Eq& E, int iAveraging, int iLaxFriedrix, double PGasL,
               double RhoL, double UL, double VL, double WL, double BXL,
               double BYL, double BZL, double PGasR, double RhoR, double UR,
               double VR, double WR, double BXR, double BYR, double BZR,
               double* out) {
  const double gamma = E.gamma, hgamma = E.hgamma, del2 = E.del2,
               smallB = E.smallB;
  double RUL = RhoL * UL, RVL = RhoL * VL, RWL = RhoL * WL;
  double BxBL = d_1_8PI * (BXL * BXL + BYL * BYL + BZL * BZL);
  double RUxUL = half * (RUL * UL + RVL * VL + RWL * WL);
  double PL = PGasL + BxBL;
  double EL = PGasL / hgamma + RUxUL + BxBL;
  double AL2 = gamma * PGasL / RhoL;
  double RUR = RhoR * UR, RVR = RhoR * VR, RWR = RhoR * WR;
  double BxBR = d_1_8PI * (BXR * BXR + BYR * BYR + BZR * BZR);
  double RUxUR = half * (RUR * UR + RVR * VR + RWR * WR);
  double PR = PGasR + BxBR;
  double ER = PGasR / hgamma + RUxUR + BxBR;
  double AR2 = gamma * PGasR / RhoR;
  double ccmi, ccma;
  if (AL2 > AR2) { ccmi = AR2; ccma = AL2; } else { ccmi = AL2; ccma = AR2; }
  double Rho, U, V, W, BX, BY, BZ, BX2, BY2, BZ2, BYZ2, BxBH, kk, cc;
  if (iAveraging == 1) {
    double HL = (EL + PL) / RhoL, HR = (ER + PR) / RhoR;
    double coefRL = one / (one + std::sqrt(RhoR / RhoL));
    double coefRR = one - coefRL;
    Rho = RhoR * coefRL + RhoL * coefRR;
    U = UL * coefRL + UR * coefRR;
    V = VL * coefRL + VR * coefRR;
    W = WL * coefRL + WR * coefRR;
    double HAV = HL * coefRL + HR * coefRR;
    BX = BXR * coefRL + BXL * coefRR;
    BY = BYR * coefRL + BYL * coefRR;
    BZ = BZR * coefRL + BZL * coefRR;
    double BXAV = half * (BXL + BXR);
    BX2 = BXAV * BXAV; BY2 = BY * BY; BZ2 = BZ * BZ;
    BYZ2 = BY2 + BZ2;
    BxBH = d_1_8PI * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    double DBX = BXR - BXL, DBY = BYR - BYL, DBZ = BZR - BZL;
    double X = d_1_8PI * (DBX * DBX + DBY * DBY + DBZ * DBZ) * coefRL * coefRL /
               RhoL;
    cc = hgamma * (HAV - kk - d_1_4PI * (BX * BXAV + BYZ2) / Rho) +
         (two - gamma) * X;
  } else {
    Rho = half * (RhoL + RhoR);
    U = half * (UL + UR); V = half * (VL + VR); W = half * (WL + WR);
    BX = half * (BXL + BXR); BY = half * (BYL + BYR); BZ = half * (BZL + BZR);
    double P = half * (PL + PR);
    BX2 = BX * BX; BY2 = BY * BY; BZ2 = BZ * BZ;
    BYZ2 = BY2 + BZ2;
    BxBH = d_1_8PI * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    double PGas = P - BxBH;
    cc = gamma * PGas / Rho;
  }
  double bb = d_1_4PI * BX2 / Rho;
  double b = std::sqrt(bb);
  if (cc < ccmi) cc = ccmi;
  if (cc > ccma) cc = ccma;
  double c = std::sqrt(cc);
  double hyy, hzz, aaf, aas, af, as;
  if (BYZ2 < smallB * smallB) {
    hyy = d_SQRT_2; hzz = d_SQRT_2;
    if (bb > cc) { aaf = bb; aas = cc; af = b; as = c; }
    else { aaf = cc; aas = bb; af = c; as = b; }
  } else {
    double hyz = std::sqrt(BYZ2);
    hyy = BY / hyz; hzz = BZ / hyz;
    double szb = cc + (BxBH + BxBH) / Rho;
    double h = std::sqrt(szb * szb - 4.0 * cc * bb);
    aas = half * (szb - h);
    if (aas <= zero) { aas = zero; as = zero; }
    else if (aas >= cc) { aas = cc; as = c; }
    else as = std::sqrt(aas);
    aaf = half * (szb + h);
    if (aaf <= cc) { aaf = cc; af = c; }
    else af = std::sqrt(aaf);
  }
  double A[7];
  if (iLaxFriedrix == 1) {
    double eigmax = std::fabs(U) + af;
    for (int q = 0; q < 7; ++q) A[q] = eigmax;
  } else {
    double eig[7] = {U, U + b, U - b, U + af, U - af, U + as, U - as};
    double eigmax = std::fabs(U) + std::fabs(V) + std::fabs(W) + af;
    double del = del2 * eigmax;
    double dell2 = del * del;
    for (int q = 0; q < 7; ++q) {
      double z = std::fabs(eig[q]);
      if (z >= del) A[q] = z;
      else A[q] = half * (z * z + dell2) / del;
    }
  }
  double dfs = aaf - aas, alf, als;
  if (dfs < 1.0e-8) { alf = one; als = zero; }
  else if ((cc - aas) <= zero) { alf = zero; als = one; }
  else if ((aaf - cc) <= zero) { alf = one; als = zero; }
  else { alf = std::sqrt((cc - aas) / dfs); als = std::sqrt((aaf - cc) / dfs); }
  double sih = -one;
  if (BX >= zero) sih = one;
  double fl[7], fr[7], du[7];
  fl[0] = RUL;
  fl[1] = UL * RUL - BXL * BXL * d_1_4PI + PL;
  fl[2] = UL * RVL - BXL * BYL * d_1_4PI;
  fl[3] = UL * RWL - BXL * BZL * d_1_4PI;
  fl[4] = UL * (EL + PL) - BXL * (UL * BXL + VL * BYL + WL * BZL) * d_1_4PI;
  fl[5] = UL * BYL - BXL * VL;
  fl[6] = UL * BZL - BXL * WL;
  fr[0] = RUR;
  fr[1] = UR * RUR - BXR * BXR * d_1_4PI + PR;
  fr[2] = UR * RVR - BXR * BYR * d_1_4PI;
  fr[3] = UR * RWR - BXR * BZR * d_1_4PI;
  fr[4] = UR * (ER + PR) - BXR * (UR * BXR + VR * BYR + WR * BZR) * d_1_4PI;
  fr[5] = UR * BYR - BXR * VR;
  fr[6] = UR * BZR - BXR * WR;
  du[0] = RhoL - RhoR; du[1] = RUL - RUR; du[2] = RVL - RVR; du[3] = RWL - RWR;
  du[4] = EL - ER; du[5] = BYL - BYR; du[6] = BZL - BZR;
  du[4] = -hgamma * (-kk * du[0] + U * du[1] + V * du[2] + W * du[3] - du[4] +
                     (BY * du[5] + BZ * du[6]) * d_1_4PI);
  du[1] = (-U * du[0] + du[1]) / Rho;
  du[2] = (-V * du[0] + du[2]) / Rho;
  du[3] = (-W * du[0] + du[3]) / Rho;
  double sqPiRho = std::sqrt(d_PI * Rho), sq2PiRho = std::sqrt(d_2PI * Rho);
  double sp[7][7], so[7][7];
  for (int r = 0; r < 7; ++r)
    for (int q = 0; q < 7; ++q) { sp[r][q] = zero; so[r][q] = zero; }
  // src/Roe7WaveF.ChF (7x7 blocks; identical to DEDNER9X9F,
  // src/DednerF.ChF:566-676)
  sp[0][0] = one; sp[0][3] = Rho * alf; sp[0][4] = Rho * alf;
  sp[0][5] = Rho * als; sp[0][6] = Rho * als;
  sp[1][3] = alf * af; sp[1][4] = -alf * af; sp[1][5] = als * as; sp[1][6] = -als * as;
  sp[2][1] = -hzz * d_SQRT_2; sp[2][2] = -hzz * d_SQRT_2;
  sp[2][3] = -als * as * hyy * sih; sp[2][4] = als * as * hyy * sih;
  sp[2][5] = alf * af * hyy * sih; sp[2][6] = -alf * af * hyy * sih;
  sp[3][1] = hyy * d_SQRT_2; sp[3][2] = hyy * d_SQRT_2;
  sp[3][3] = -als * as * hzz * sih; sp[3][4] = als * as * hzz * sih;
  sp[3][5] = alf * af * hzz * sih; sp[3][6] = -alf * af * hzz * sih;
  sp[4][3] = alf * Rho * cc; sp[4][4] = alf * Rho * cc;
  sp[4][5] = als * Rho * cc; sp[4][6] = als * Rho * cc;
  sp[5][1] = hzz * sq2PiRho * sih; sp[5][2] = -hzz * sq2PiRho * sih;
  sp[5][3] = two * als * sqPiRho * c * hyy; sp[5][4] = two * als * sqPiRho * c * hyy;
  sp[5][5] = -two * alf * sqPiRho * c * hyy; sp[5][6] = -two * alf * sqPiRho * c * hyy;
  sp[6][1] = -hyy * sq2PiRho * sih; sp[6][2] = hyy * sq2PiRho * sih;
  sp[6][3] = two * als * sqPiRho * c * hzz; sp[6][4] = two * als * sqPiRho * c * hzz;
  sp[6][5] = -two * alf * sqPiRho * c * hzz; sp[6][6] = -two * alf * sqPiRho * c * hzz;
  so[0][0] = one; so[0][4] = -one / cc;
  so[1][2] = -hzz * d_SQRT_2; so[1][3] = hyy * d_SQRT_2;
  so[1][5] = half * hzz * sih / sq2PiRho; so[1][6] = -half * hyy * sih / sq2PiRho;
  so[2][2] = -hzz * d_SQRT_2; so[2][3] = hyy * d_SQRT_2;
  so[2][5] = -half * hzz * sih / sq2PiRho; so[2][6] = half * hyy * sih / sq2PiRho;
  so[3][1] = half * alf * af / cc; so[3][2] = -half * als * as * hyy * sih / cc;
  so[3][3] = -half * als * as * hzz * sih / cc; so[3][4] = half * alf / (cc * Rho);
  so[3][5] = 0.25 * als * hyy / (c * sqPiRho); so[3][6] = 0.25 * als * hzz / (c * sqPiRho);
  so[4][1] = -half * alf * af / cc; so[4][2] = half * als * as * hyy * sih / cc;
  so[4][3] = half * als * as * hzz * sih / cc; so[4][4] = half * alf / (cc * Rho);
  so[4][5] = 0.25 * als * hyy / (c * sqPiRho); so[4][6] = 0.25 * als * hzz / (c * sqPiRho);
  so[5][1] = half * als * as / cc; so[5][2] = half * alf * af * hyy * sih / cc;
  so[5][3] = half * alf * af * hzz * sih / cc; so[5][4] = half * als / (Rho * cc);
  so[5][5] = -0.25 * alf * hyy / (c * sqPiRho); so[5][6] = -0.25 * alf * hzz / (c * sqPiRho);
  so[6][1] = -half * als * as / cc; so[6][2] = -half * alf * af * hyy * sih / cc;
  so[6][3] = -half * alf * af * hzz * sih / cc; so[6][4] = half * als / (Rho * cc);
  so[6][5] = -0.25 * alf * hyy / (c * sqPiRho); so[6][6] = -0.25 * alf * hzz / (c * sqPiRho);
  double lu[7], ra[7];
  for (int r = 0; r < 7; ++r) {
    double s = so[r][0] * du[0];
    for (int q = 1; q < 7; ++q) s = s + so[r][q] * du[q];
    lu[r] = A[r] * s;
  }
  for (int r = 0; r < 7; ++r) {
    double s = sp[r][0] * lu[0];
    for (int q = 1; q < 7; ++q) s = s + sp[r][q] * lu[q];
    ra[r] = s;
  }
  ra[4] = ra[0] * kk + Rho * (U * ra[1] + V * ra[2] + W * ra[3]) + ra[4] / hgamma +
          d_1_4PI * (BY * ra[5] + BZ * ra[6]);
  ra[1] = U * ra[0] + Rho * ra[1];
  ra[2] = V * ra[0] + Rho * ra[2];
  ra[3] = W * ra[0] + Rho * ra[3];
  out[0] = half * (fl[0] + fr[0] + ra[0]);
  out[1] = half * (fl[1] + fr[1] + ra[1]);
  out[2] = half * (fl[2] + fr[2] + ra[2]);
  out[3] = half * (fl[3] + fr[3] + ra[3]);
  out[4] = half * (fl[4] + fr[4] + ra[4]);
  out[6] = half * (fl[5] + fr[5] + ra[5]);
  out[7] = half * (fl[6] + fr[6] + ra[6]);
  out[5] = zero;
}

// RUSANOVF (src/RusanovF.ChF:4-197), after the input clamps.
// HLLC flux of ideal MHD after S. Li, "An HLLC Riemann solver for
// magneto-hydrodynamics", J. Comput. Phys. 203 (2005) 344-357, eqs. (14)-
// (31), written in Gaussian units (B^2/8pi, B B/4pi).  The reference has no
// HLLC solver: this function is pinned by its closed forms (consistency,
// exact stationary contacts, upwinding beyond the outer waves, agreement
// with HLL in the density of a jump without a contact), not by reference
// parity.  Outer speeds after Davis: S_L = min(u_L - cf_L, u_R - cf_R),
// S_R = max(u_L + cf_L, u_R + cf_R); star field = the HLL field; contact
// speed S_M = (rho u)_hll / rho_hll.
void hllc_core(const Eq& E, double PGasL, double RhoL, double UL, double VL,
               double WL, double BXL, double BYL, double BZL, double PGasR,
               double RhoR, double UR, double VR, double WR, double BXR,
               double BYR, double BZR, double* out) {
  struct Side {
    double rho, u, v, w, p, bx, by, bz, pt, U[8], F[8];
  } s[2] = {{RhoL, UL, VL, WL, PGasL, BXL, BYL, BZL},
            {RhoR, UR, VR, WR, PGasR, BXR, BYR, BZR}};
  double cf[2];
  for (Side& q : s) {
    const double B2 = q.bx * q.bx + q.by * q.by + q.bz * q.bz;
    q.pt = q.p + d_1_8PI * B2;
    const double En = q.p / E.hgamma + half * q.rho * (q.u * q.u + q.v * q.v + q.w * q.w) +
                      d_1_8PI * B2;
    const double vB = q.u * q.bx + q.v * q.by + q.w * q.bz;
    const double cons[8] = {q.rho, q.rho * q.u, q.rho * q.v, q.rho * q.w, En, q.bx, q.by, q.bz};
    const double flx[8] = {q.rho * q.u,
                           q.rho * q.u * q.u + q.pt - d_1_4PI * q.bx * q.bx,
                           q.rho * q.u * q.v - d_1_4PI * q.bx * q.by,
                           q.rho * q.u * q.w - d_1_4PI * q.bx * q.bz,
                           (En + q.pt) * q.u - d_1_4PI * q.bx * vB,
                           zero,
                           q.u * q.by - q.v * q.bx,
                           q.u * q.bz - q.w * q.bx};
    for (int c = 0; c < 8; ++c) { q.U[c] = cons[c]; q.F[c] = flx[c]; }
  }
  for (int k = 0; k < 2; ++k) {
    const Side& q = s[k];
    const double a2 = E.gamma * q.p / q.rho;
    const double va2 = d_1_4PI * (q.bx * q.bx + q.by * q.by + q.bz * q.bz) / q.rho;
    const double vax2 = d_1_4PI * q.bx * q.bx / q.rho;
    const double disc = std::max(zero, (a2 + va2) * (a2 + va2) - 4.0 * a2 * vax2);
    cf[k] = std::sqrt(half * (a2 + va2 + std::sqrt(disc)));
  }
  const double SL = std::min(s[0].u - cf[0], s[1].u - cf[1]);
  const double SR = std::max(s[0].u + cf[0], s[1].u + cf[1]);
  if (SL >= zero) { for (int c = 0; c < 8; ++c) out[c] = s[0].F[c]; return; }
  if (SR <= zero) { for (int c = 0; c < 8; ++c) out[c] = s[1].F[c]; return; }
  double hll[8];
  for (int c = 0; c < 8; ++c)
    hll[c] = (SR * s[1].U[c] - SL * s[0].U[c] - (s[1].F[c] - s[0].F[c])) / (SR - SL);
  const double SM = hll[1] / hll[0];
  const double bxs = hll[5], bys = hll[6], bzs = hll[7];
  const double vs = hll[2] / hll[0], ws = hll[3] / hll[0];
  const double pts = s[0].rho * (SL - s[0].u) * (SM - s[0].u) + s[0].pt -
                     d_1_4PI * (s[0].bx * s[0].bx - bxs * bxs);
  const int k = SM >= zero ? 0 : 1;
  const Side& q = s[k];
  const double S = k == 0 ? SL : SR;
  const double den = S - SM;
  double st[8];
  st[0] = q.rho * (S - q.u) / den;
  st[1] = st[0] * SM;
  st[2] = (q.rho * q.v * (S - q.u) - d_1_4PI * (bxs * bys - q.bx * q.by)) / den;
  st[3] = (q.rho * q.w * (S - q.u) - d_1_4PI * (bxs * bzs - q.bx * q.bz)) / den;
  st[4] = (q.U[4] * (S - q.u) - q.pt * q.u + pts * SM +
           d_1_4PI * (q.bx * (q.u * q.bx + q.v * q.by + q.w * q.bz) -
                      bxs * (SM * bxs + vs * bys + ws * bzs))) / den;
  st[5] = bxs; st[6] = bys; st[7] = bzs;
  for (int c = 0; c < 8; ++c) out[c] = q.F[c] + S * (st[c] - q.U[c]);
}

void rusanov_core(const Eq& E, int iAveraging, double PGasL, double RhoL,
                  double UL, double VL, double WL, double BXL, double BYL,
                  double BZL, double PGasR, double RhoR, double UR, double VR,
                  double WR, double BXR, double BYR, double BZR, double* out) {
  const double gamma = E.gamma, hgamma = E.hgamma;
  double RUL = RhoL * UL, RVL = RhoL * VL, RWL = RhoL * WL;
  double BxBL = d_1_8PI * (BXL * BXL + BYL * BYL + BZL * BZL);
  double RUxUL = half * (RUL * UL + RVL * VL + RWL * WL);
  double PL = PGasL + BxBL;
  double EL = PGasL / hgamma + RUxUL + BxBL;
  double BXL_4PI = BXL * d_1_4PI;
  double RUR = RhoR * UR, RVR = RhoR * VR, RWR = RhoR * WR;
  double BxBR = d_1_8PI * (BXR * BXR + BYR * BYR + BZR * BZR);
  double RUxUR = half * (RUR * UR + RVR * VR + RWR * WR);
  double PR = PGasR + BxBR;
  double ER = PGasR / hgamma + RUxUR + BxBR;
  double BXR_4PI = BXR * d_1_4PI;
  double Rho, U, BX2, BxBH, A2;
  if (iAveraging == 1) {
    double HL = (EL + PL) / RhoL, HR = (ER + PR) / RhoR;
    double coefRL = one / (one + std::sqrt(RhoR / RhoL));
    double coefRR = one - coefRL;
    Rho = RhoR * coefRL + RhoL * coefRR;
    U = UL * coefRL + UR * coefRR;
    double V = VL * coefRL + VR * coefRR;
    double W = WL * coefRL + WR * coefRR;
    double HAV = HL * coefRL + HR * coefRR;
    double BX = BXR * coefRL + BXL * coefRR;
    double BY = BYR * coefRL + BYL * coefRR;
    double BZ = BZR * coefRL + BZL * coefRR;
    double BXAV = half * (BXL + BXR);
    BX2 = BXAV * BXAV;
    double BY2 = BY * BY, BZ2 = BZ * BZ;
    double BYZ2 = BY2 + BZ2;
    BxBH = d_1_8PI * (BX2 + BYZ2);
    double kk = half * (U * U + V * V + W * W);
    double DBX = BXR - BXL, DBY = BYR - BYL, DBZ = BZR - BZL;
    double X = d_1_8PI * (DBX * DBX + DBY * DBY + DBZ * DBZ) * coefRL * coefRL /
               RhoL;
    A2 = hgamma * (HAV - kk - d_1_4PI * (BX * BXAV + BYZ2) / Rho) +
         (two - gamma) * X;
  } else {
    Rho = half * (RhoL + RhoR);
    U = half * (UL + UR);
    double BX = half * (BXL + BXR), BY = half * (BYL + BYR),
           BZ = half * (BZL + BZR);
    double PGas = half * (PGasL + PGasR);
    BX2 = BX * BX;
    double BY2 = BY * BY, BZ2 = BZ * BZ;
    BxBH = d_1_8PI * (BX2 + BY2 + BZ2);
    A2 = gamma * PGas / Rho;
  }
  double bb = d_1_4PI * BX2 / Rho;
  double szb = A2 + (BxBH + BxBH) / Rho;
  double h = std::sqrt(szb * szb - 4.0 * A2 * bb);
  double aaf = half * (szb + h);
  double af = std::sqrt(aaf);
  double eigmax = std::fabs(U) + af;
  double fl[8], fr[8], du[8];
  fl[0] = RUL;
  fl[1] = UL * RUL - BXL_4PI * BXL + PL;
  fl[2] = UL * RVL - BXL_4PI * BYL;
  fl[3] = UL * RWL - BXL_4PI * BZL;
  fl[4] = UL * (EL + PL) - BXL_4PI * (UL * BXL + VL * BYL + WL * BZL);
  fl[5] = zero;
  fl[6] = UL * BYL - BXL * VL;
  fl[7] = UL * BZL - BXL * WL;
  fr[0] = RUR;
  fr[1] = UR * RUR - BXR_4PI * BXR + PR;
  fr[2] = UR * RVR - BXR_4PI * BYR;
  fr[3] = UR * RWR - BXR_4PI * BZR;
  fr[4] = UR * (ER + PR) - BXR_4PI * (UR * BXR + VR * BYR + WR * BZR);
  fr[5] = zero;
  fr[6] = UR * BYR - BXR * VR;
  fr[7] = UR * BZR - BXR * WR;
  du[0] = RhoL - RhoR; du[1] = RUL - RUR; du[2] = RVL - RVR; du[3] = RWL - RWR;
  du[4] = EL - ER; du[5] = BXL - BXR; du[6] = BYL - BYR; du[7] = BZL - BZR;
  for (int q = 0; q < 8; ++q) out[q] = half * (fl[q] + fr[q] + eigmax * du[q]);
}

// VANLEERF / VANLEERRS (src/VanLeerRSF.ChF:4-130, :132-240) after clamps.
// Inputs in face-rotated order; out = rho, mom_n, mom_t1, mom_t2, E.
void vanleer_core(const Eq& E, double PGasL, double RhoL, double UL,
                  double VL, double WL, double PGasR, double RhoR, double UR,
                  double VR, double WR, double* out) {
  const double gamma = E.gamma, hgamma = E.hgamma;
  double RUL = RhoL * UL, RVL = RhoL * VL, RWL = RhoL * WL;
  double RUxUL = half * (RUL * UL + RVL * VL + RWL * WL);
  double EL = PGasL / hgamma + RUxUL;
  double AL2 = gamma * PGasL / RhoL;
  double AL = std::sqrt(AL2);
  double fr1, fr2, fr3, fr4, fr5;
  if (UL <= -AL) { fr1 = fr2 = fr3 = fr4 = fr5 = zero; }
  else if (UL >= AL) {
    fr1 = RhoL * UL; fr2 = fr1 * UL + PGasL; fr3 = fr1 * VL; fr4 = fr1 * WL;
    fr5 = UL * (EL + PGasL);
  } else {
    double UA = (-UL + (AL + AL)) / gamma;
    double H = (EL + PGasL) / RhoL;
    double HU = UL + AL;
    fr1 = 0.25 * RhoL * HU * HU / AL;
    fr2 = fr1 * (UL + UA); fr3 = fr1 * (VL); fr4 = fr1 * (WL); fr5 = fr1 * (H);
  }
  double RUR = RhoR * UR, RVR = RhoR * VR, RWR = RhoR * WR;
  double RUxUR = half * (RUR * UR + RVR * VR + RWR * WR);
  double ER = PGasR / hgamma + RUxUR;
  double AR2 = gamma * PGasR / RhoR;
  double AR = std::sqrt(AR2);
  double fl1, fl2, fl3, fl4, fl5;
  if (UR >= AR) { fl1 = fl2 = fl3 = fl4 = fl5 = zero; }
  else if (UR <= -AR) {
    fl1 = RhoR * UR; fl2 = fl1 * UR + PGasR; fl3 = fl1 * VR; fl4 = fl1 * WR;
    fl5 = UR * (ER + PGasR);
  } else {
    double UA = (-UR - (AR + AR)) / gamma;
    double H = (ER + PGasR) / RhoR;
    double HU = UR - AR;
    fl1 = -0.25 * RhoR * HU * HU / AR;
    fl2 = fl1 * (UR + UA); fl3 = fl1 * (VR); fl4 = fl1 * (WR); fl5 = fl1 * (H);
  }
  out[0] = fl1 + fr1; out[1] = fl2 + fr2; out[2] = fl3 + fr3;
  out[3] = fl4 + fr4; out[4] = fl5 + fr5;
}

// Rotated component indices of fluid starting at iRho (RIEMANNF :54-68).
struct Rot { int rho, n, t1, t2, p, bn, bt1, bt2; };
Rot rot(int iRho, int idir) {
  Rot r;
  int iVelX = iRho + 1, iBX = iRho + 5;
  r.rho = iRho; r.p = iRho + 4;
  r.n = iVelX + idir; r.t1 = iVelX + (idir + 1) % 3; r.t2 = iVelX + (idir + 2) % 3;
  r.bn = iBX + idir; r.bt1 = iBX + (idir + 1) % 3; r.bt2 = iBX + (idir + 2) % 3;
  return r;
}

// Plasma flux at one face: m_RS->fluxes (src/RiemannSolver.cpp:36-105).
// wl/wr: full primitive vectors; out: full flux vector (plasma part, and
// psi when Dedner is on).
void plasma_flux_face(const Eq& E, const double* wl, const double* wr,
                      int idir, double* out) {
  Rot r = rot(0, idir);
  double PGasL = std::max(E.smallp, wl[r.p]), RhoL = std::max(E.smallr, wl[r.rho]);
  double PGasR = std::max(E.smallp, wr[r.p]), RhoR = std::max(E.smallr, wr[r.rho]);
  double f[8];
  int flux = E.P->flux;
  if (flux == MFM_FLUX_HLLC) {
    // this library's HLLC (no counterpart in the reference); with GLM the
    // (B_n, psi) block below replaces its normal-field flux
    hllc_core(E, PGasL, RhoL, wl[r.n], wl[r.t1], wl[r.t2], wl[r.bn], wl[r.bt1], wl[r.bt2],
              PGasR, RhoR, wr[r.n], wr[r.t1], wr[r.t2], wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  } else if (E.dedner || flux == 11 || flux == 13) {
    // Roe7WavesRS::fluxes (src/Roe7WavesRS.cpp:38-78), flux 11 / 13 of the
    // factory (src/RiemannSolver.cpp:76-89); DednerRS::fluxes: RIEMANN7F
    // then DEDNER2X2F (src/DednerRS.cpp:38-95)
    roe7_core(E, E.iAveraging, E.iLaxFriedrix, PGasL, RhoL, wl[r.n], wl[r.t1],
              wl[r.t2], wl[r.bn], wl[r.bt1], wl[r.bt2], PGasR, RhoR, wr[r.n],
              wr[r.t1], wr[r.t2], wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  } else if (flux == MFM_FLUX_RUSANOV || flux == MFM_FLUX_RUSANOV_ROEAVG) {
    rusanov_core(E, E.iAveraging, PGasL, RhoL, wl[r.n], wl[r.t1], wl[r.t2],
                 wl[r.bn], wl[r.bt1], wl[r.bt2], PGasR, RhoR, wr[r.n], wr[r.t1],
                 wr[r.t2], wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  } else {
    roe8_core(E, E.iAveraging, E.iLaxFriedrix, PGasL, RhoL, wl[r.n], wl[r.t1],
              wl[r.t2], wl[r.bn], wl[r.bt1], wl[r.bt2], PGasR, RhoR, wr[r.n],
              wr[r.t1], wr[r.t2], wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  }
  out[r.rho] = f[0]; out[r.n] = f[1]; out[r.t1] = f[2]; out[r.t2] = f[3];
  out[r.p] = f[4]; out[r.bn] = f[5]; out[r.bt1] = f[6]; out[r.bt2] = f[7];
  if (E.dedner) {
    // DEDNER2X2F (src/DednerF.ChF:89-172)
    double dCh = E.ch, dCh2 = dCh * dCh;
    int inormB = WBX + idir, iPsi = E.iPsi;
    double BXL = wl[inormB], PsiL = wl[iPsi], BXR = wr[inormB], PsiR = wr[iPsi];
    double sp11 = one, sp12 = one, sp21 = -dCh, sp22 = dCh;
    double so11 = half, so12 = -half / dCh, so21 = half, so22 = half / dCh;
    double fl1 = PsiL, fl2 = dCh2 * BXL, fr1 = PsiR, fr2 = dCh2 * BXR;
    double du1 = BXL - BXR, du2 = PsiL - PsiR;
    double lu1 = dCh * (so11 * du1 + so12 * du2);
    double lu2 = dCh * (so21 * du1 + so22 * du2);
    double ra1 = sp11 * lu1 + sp12 * lu2;
    double ra2 = sp21 * lu1 + sp22 * lu2;
    out[inormB] = half * (fl1 + fr1 + ra1);
    out[iPsi] = half * (fl2 + fr2 + ra2);
  }
}

// Neutral flux of fluid starting at iRho: m_RSGD->fluxes = VanLeerRS
void neutral_flux_face(const Eq& E, const double* wl, const double* wr,
                       int idir, int iRho, double* out) {
  Rot r = rot(iRho, idir);
  double PGasL = std::max(E.smallp, wl[r.p]), RhoL = std::max(E.smallr, wl[r.rho]);
  double PGasR = std::max(E.smallp, wr[r.p]), RhoR = std::max(E.smallr, wr[r.rho]);
  double f[5];
  vanleer_core(E, PGasL, RhoL, wl[r.n], wl[r.t1], wl[r.t2], PGasR, RhoR,
               wr[r.n], wr[r.t1], wr[r.t2], f);
  out[r.rho] = f[0]; out[r.n] = f[1]; out[r.t1] = f[2]; out[r.t2] = f[3];
  out[r.p] = f[4];
}

// =========================================================== problems ======
// ---- Riemann problem: RIEMANNGS (problems/RiemannF.ChF:276-367), dir 0 only
void riemann_fill_ghost(const Eq& E, Fab& W, int dir) {
  if (dir != 0 || !E.physical(0)) return;
  Box WBox = W.b;
  Box tmp = E.clip(WBox);
  const double* L = E.P->rp_left;
  const double* R = E.P->rp_right;
  for (int side = 0; side < 2; ++side) {
    Box bb = WBox;
    if (side == 1) {
      if (WBox.hi[0] == tmp.hi[0]) continue;
      bb.lo[0] = tmp.hi[0] + 1;
    } else {
      if (WBox.lo[0] == tmp.lo[0]) continue;
      bb.hi[0] = tmp.lo[0] - 1;
    }
    const double* S = side ? R : L;
    FORBOX(bb, i, j, k) {
      for (int c = 0; c < 8; ++c) W(i, j, k, c) = S[c];
      if (E.dedner) W(i, j, k, E.iPsi) = zero;
    }
  }
}
// FLUXBC (problems/RiemannF.ChF:129-275), single-fluid part
void riemann_flux_bc_face(const Eq& E, int side, double* F, double* Bn) {
  const double* S = side ? E.P->rp_right : E.P->rp_left;
  double rho = S[0], vx = S[1], vy = S[2], vz = S[3], pr = S[4];
  double Bx = S[5], By = S[6], Bz = S[7];
  double b = d_1_8PI * (Bx * Bx + By * By + Bz * Bz);
  double ke = rho * (vx * vx + vy * vy + vz * vz);
  double bu = Bx * vx + By * vy + Bz * vz;
  double P = pr + b;
  double e = pr / E.hgamma + half * ke + b;
  F[URHO] = rho * vx;
  F[UMOMX] = rho * vx * vx - d_1_4PI * Bx * Bx + P;
  F[UMOMY] = rho * vx * vy - d_1_4PI * Bx * By;
  F[UMOMZ] = rho * vx * vz - d_1_4PI * Bx * Bz;
  F[UENG] = vx * (e + P) - d_1_4PI * Bx * bu;
  F[UBX] = zero;
  F[UBY] = vx * By - vy * Bx;
  F[UBZ] = vx * Bz - vz * Bx;
  *Bn = Bx;
  if (E.dedner) F[E.iPsi] = E.ch * E.ch * Bx;
}

// ---- Heliospheric problem
// HELIOGS (swlism/HeliosphericF.ChF:978-1149) on one ghost strip
void helio_gs(const Eq& E, Fab& W, int lohisign, int dir, const Box& box) {
  const mfm_params* P = E.P;
  FORBOX(box, i, j, k) {
    W(i, j, k, WRHO) = one;
    W(i, j, k, WVELX) = P->lism_ux; W(i, j, k, WVELY) = P->lism_uy;
    W(i, j, k, WVELZ) = P->lism_uz;
    W(i, j, k, WPRES) = P->lism_p;
    W(i, j, k, WBX) = P->lism_bx; W(i, j, k, WBY) = P->lism_by;
    W(i, j, k, WBZ) = P->lism_bz;
  }
  if (E.nf > 1) {
    int s = E.iRhoN;
    FORBOX(box, i, j, k) {
      W(i, j, k, s + WRHO) = P->net_rho;
      W(i, j, k, s + WVELX) = P->lism_ux; W(i, j, k, s + WVELY) = P->lism_uy;
      W(i, j, k, s + WVELZ) = P->lism_uz;
      W(i, j, k, s + WPRES) = P->net_p;
    }
    for (int f = 2; f <= E.nf - 1; ++f) {
      int s2 = E.iRhoN + (f - 1) * WNUM_E;
      int src = (lohisign == -1) ? box.hi[dir] + 1 : box.lo[dir] - 1;
      FORBOX(box, i, j, k) {
        int ii = i, jj = j, kk = k;
        if (dir == 0) ii = src; else if (dir == 1) jj = src; else kk = src;
        for (int c = 0; c < 5; ++c) W(i, j, k, s2 + c) = W(ii, jj, kk, s2 + c);
      }
    }
  }
}
void helio_fill_ghost(const Eq& E, Fab& W, int dir) {
  if (!E.physical(dir)) return;
  Box WBox = W.b;
  Box tmp = E.clip(WBox);
  if (WBox.hi[dir] != tmp.hi[dir]) {
    Box bb = WBox; bb.lo[dir] = tmp.hi[dir] + 1;
    helio_gs(E, W, 1, dir, bb);
  }
  if (WBox.lo[dir] != tmp.lo[dir]) {
    Box bb = WBox; bb.hi[dir] = tmp.lo[dir] - 1;
    helio_gs(E, W, -1, dir, bb);
  }
}

// BFLUX / BFLUXR (swlism/HeliosphericF.ChF:1363-1560, :1562-1774).
// radial = true selects BFLUXR with the outward direction (DX,DY,DZ).
void helio_bflux(const Eq& E, bool radial, double rho_in, double vn_in,
                 double vt1_in, double vt2_in, double p_in, double Bn_in,
                 double Bt1_in, double Bt2_in, double rho_out, double vn_out,
                 double vt1_out, double vt2_out, double p_out, double Bn_out,
                 double Bt1_out, double Bt2_out, double DX, double DY,
                 double DZ, double* FL, double* BnValue,
                 double* wchar = nullptr) {
  // wchar (tests only): the characteristic boundary state of the sub-fast
  // outflow branch, (rho, vn, vt1, vt2, p, Bn, Bt1, Bt2), when it is taken
  const double gamma = E.gamma, hgamma = E.hgamma;
  double U = vn_in, V = vt1_in, W = vt2_in, R = rho_in;
  double BX = Bn_in, BY = Bt1_in, BZ = Bt2_in, PG = p_in;
  double BB = BX * BX + BY * BY + BZ * BZ;
  double P = PG + d_1_8PI * BB;
  double KE = half * (vn_in * vn_in + vt1_in * vt1_in + vt2_in * vt2_in);
  double Ee = PG / hgamma + rho_in * KE + d_1_8PI * BB;
  double SZ = std::sqrt(gamma * PG / R);
  double B2 = BB;
  double SP = half * std::sqrt(d_4PI * R);
  double EM = SZ * SZ + B2 / (d_4PI * R) + SZ * std::fabs(BX) / SP;
  double EMM = SZ * SZ + B2 / (d_4PI * R) - SZ * std::fabs(BX) / SP;
  double CF = half * (std::sqrt(EM) + std::sqrt(EMM));
  double CS = half * (std::sqrt(EM) - std::sqrt(EMM));
  if (U >= CF) {
    *BnValue = BX;
    FL[0] = R * U;
    FL[1] = R * U * U + P - d_1_4PI * BX * BX;
    FL[2] = R * U * V - d_1_4PI * BX * BY;
    FL[3] = R * U * W - d_1_4PI * BX * BZ;
    FL[4] = U * (Ee + P) - d_1_4PI * BX * (U * BX + V * BY + W * BZ);
    FL[5] = zero;
    FL[6] = U * BY - V * BX;
    FL[7] = U * BZ - W * BX;
    return;
  }
  double U1 = vn_out, V1 = vt1_out, W1 = vt2_out, R1 = rho_out;
  double BX1 = Bn_out, BY1 = Bt1_out, BZ1 = Bt2_out, PG1 = p_out;
  double BB1 = BX1 * BX1 + BY1 * BY1 + BZ1 * BZ1;
  double P1 = PG1 + d_1_8PI * BB1;
  double KE1 = half * (vn_out * vn_out + vt1_out * vt1_out + vt2_out * vt2_out);
  double E1 = PG1 / hgamma + rho_out * KE1 + d_1_8PI * BB1;
  if (U < -CF) {
    *BnValue = BX1;
    FL[0] = R1 * U1;
    FL[1] = R1 * U1 * U1 + P1 - d_1_4PI * BX1 * BX1;
    FL[2] = R1 * U1 * V1 - d_1_4PI * BX1 * BY1;
    FL[3] = R1 * U1 * W1 - d_1_4PI * BX1 * BZ1;
    FL[4] = U1 * (E1 + P1) - d_1_4PI * BX1 * (U1 * BX1 + V1 * BY1 + W1 * BZ1);
    FL[5] = zero;
    FL[6] = U1 * BY1 - V1 * BX1;
    FL[7] = U1 * BZ1 - W1 * BX1;
    return;
  }
  double SZ1 = std::sqrt(gamma * PG1 / R1);
  double B21 = BB1;
  double SP1 = half * std::sqrt(d_4PI * R1);
  double EM1 = SZ1 * SZ1 + B21 / (d_4PI * R1) + SZ1 * std::fabs(BX1) / SP1;
  double EMM1 = SZ1 * SZ1 + B21 / (d_4PI * R1) - SZ1 * std::fabs(BX1) / SP1;
  double CF1 = half * (std::sqrt(EM1) + std::sqrt(EMM1));
  bool branch;
  if (radial) {
    double RR = std::sqrt(DX * DX + DY * DY + DZ * DZ);
    (void)RR;  // UR1 only feeds a disabled condition in BFLUXR
    branch = (U < CF) && (U >= zero);
  } else {
    branch = (U < CF) && (U >= zero) && (U1 > CF1);
  }
  if (branch) {
    double SCM = SZ * SZ - CS * CS, SCP = SZ * SZ + CS * CS;
    double DEM = std::sqrt(EM * EMM);
    double THE = -half * CF * (BX * BX) / (R * d_4PI * DEM);
    THE = THE + (BZ * BZ + BY * BY) * CF * SCP * half / (R * DEM * d_4PI * SCM);
    THE = THE + half * CF * hgamma * (SZ * SZ) / DEM;
    THE = THE - SZ * std::fabs(BX) * CS * (gamma - two) * 0.25 / (SP * DEM);
    double hyz = BY * BY + BZ * BZ, hy2, hz2;
    double R16, DRho, HRho, CF16, SZ16, BE;
    if (!radial) {
      CF16 = (U * THE + CF * CF) / (THE + CF);
      HRho = (U - CF) / (THE + CF);
      DRho = R * HRho;
      R16 = R + DRho;
      SZ16 = SZ * (one + half * hgamma * HRho);
    }
    if (hyz < 1.0e-10) { hy2 = half; hz2 = half; }
    else { hy2 = BY * BY / hyz; hz2 = BZ * BZ / hyz; }
    double BZ16, BY16;
    if (!radial) {
      BE = two * d_4PI * (CF * CF - SZ * SZ) * DRho;
      BZ16 = BZ * BZ + BE * hz2;
    } else {
      BE = two * d_4PI * (CF * CF - SZ * SZ);
      R16 = std::max(R + R * (U - CF) / (THE + CF), 1.0e-14);
      DRho = R16 - R;
      HRho = DRho / R;
      CF16 = U - HRho * CF;
      SZ16 = SZ * (one + half * hgamma * HRho);
      BZ16 = BZ * BZ + BE * hz2 * DRho;
    }
    double U16 = CF16;
    if (BZ16 <= zero) BZ16 = zero;
    else { BZ16 = std::sqrt(BZ16); if (BZ < zero) BZ16 = -BZ16; }
    BY16 = radial ? BY * BY + BE * hy2 * DRho : BY * BY + BE * hy2;
    if (BY16 <= zero) BY16 = zero;
    else { BY16 = std::sqrt(BY16); if (BY < zero) BY16 = -BY16; }
    double BX16 = BX;
    double ALF = d_1_4PI * BX / (R * CF);
    double W16 = W + ALF * (BZ16 - BZ);
    double V16 = V + ALF * (BY16 - BY);
    double PG16 = (SZ16 * SZ16) * R16 / gamma;
    double BB2 = d_1_4PI * (BX16 * BX16 + BY16 * BY16 + BZ16 * BZ16);
    double P16 = PG16 + half * BB2;
    double E16 = PG16 / hgamma +
                 half * (R16 * ((U16 * U16) + (V16 * V16) + (W16 * W16)) + BB2);
    double BXs = d_1_4PI * BX16;
    if (wchar) {
      const double ws[8] = {R16, U16, V16, W16, PG16, BX16, BY16, BZ16};
      for (int c = 0; c < 8; ++c) wchar[c] = ws[c];
    }
    *BnValue = BX16;
    FL[0] = R16 * U16;
    FL[1] = R16 * U16 * U16 + P16 - BXs * BX16;
    FL[2] = R16 * U16 * V16 - BXs * BY16;
    FL[3] = R16 * U16 * W16 - BXs * BZ16;
    FL[4] = U16 * (E16 + P16) - BXs * (U16 * BX16 + V16 * BY16 + W16 * BZ16);
    FL[5] = zero;
    FL[6] = U16 * BY16 - V16 * BX16;
    FL[7] = U16 * BZ16 - W16 * BX16;
  } else if (radial && U >= zero) {
    *BnValue = BX;
    roe8_core(E, E.iAveraging, E.iLaxFriedrix, PG, R, U, V, W, BX, BY, BZ, PG1,
              R, U, V, W, BX1, BY1, BZ1, FL);
  } else {
    *BnValue = half * (BX + BX1);
    roe8_core(E, E.iAveraging, E.iLaxFriedrix, PG, R, U, V, W, BX, BY, BZ, PG1,
              R1, U1, V1, W1, BX1, BY1, BZ1, FL);
  }
}

// VANLEERRS with already-clamped inputs (pointwise).
void vanleer_rs(const Eq& E, double pL, double rhoL, double vnL, double vt1L,
                double vt2L, double pR, double rhoR, double vnR, double vt1R,
                double vt2R, double* f) {
  vanleer_core(E, pL, rhoL, vnL, vt1L, vt2L, pR, rhoR, vnR, vt1R, vt2R, f);
}

// HELIOBC (swlism/HeliosphericF.ChF:221-925) at one boundary face.
// wl = W+ of the cell below the face, wr = W- of the cell above.
void helio_bc_face(const Eq& E, int lohisign, int dir, double dx, int i,
                   int j, int k, const double* WLeft, const double* WRight,
                   double* F, double* Bn) {
  const mfm_params* Pp = E.P;
  Rot r = rot(0, dir);
  double FL[8], BnValue;
  if (dir == 0 && lohisign == 1) {
    // inflow: fixed LISM flux
    double density = one, velx = Pp->lism_ux, vely = Pp->lism_uy,
           velz = Pp->lism_uz, pressure = Pp->lism_p;
    double Bx = Pp->lism_bx, By = Pp->lism_by, Bz = Pp->lism_bz;
    double b = d_1_8PI * (Bx * Bx + By * By + Bz * Bz);
    double ke = density * (velx * velx + vely * vely + velz * velz);
    double bu = Bx * velx + By * vely + Bz * velz;
    double P = pressure + b;
    double e = pressure / E.hgamma + half * ke + b;
    F[URHO] = density * velx;
    F[UMOMX] = density * velx * velx - d_1_4PI * Bx * Bx + P;
    F[UMOMY] = density * velx * vely - d_1_4PI * Bx * By;
    F[UMOMZ] = density * velx * velz - d_1_4PI * Bx * Bz;
    F[UENG] = velx * (e + P) - d_1_4PI * Bx * bu;
    F[UBX] = zero;
    F[UBY] = velx * By - vely * Bx;
    F[UBZ] = velx * Bz - velz * Bx;
    *Bn = Bx;
  } else if (lohisign == -1) {
    double rhoR = WLeft[WRHO], vnR = -WLeft[r.n], vt1R = -WLeft[r.t1],
           vt2R = -WLeft[r.t2], pR = WLeft[WPRES], BnR = -WLeft[r.bn],
           Bt1R = -WLeft[r.bt1], Bt2R = -WLeft[r.bt2];
    rhoR = std::max(E.smallr, rhoR); pR = std::max(E.smallp, pR);
    double rhoL = WRight[WRHO], vnL = -WRight[r.n], vt1L = -WRight[r.t1],
           vt2L = -WRight[r.t2], pL = WRight[WPRES], BnL = -WRight[r.bn],
           Bt1L = -WRight[r.bt1], Bt2L = -WRight[r.bt2];
    rhoL = std::max(E.smallr, rhoL); pL = std::max(E.smallp, pL);
    double x = 0, y = 0, z = 0;
    bool radial = (dir == 0);
    if (radial) {
      x = -((i + half) * dx - Pp->sun_xc);
      y = -((j + half) * dx - Pp->sun_yc);
      z = -((k + half) * dx - Pp->sun_zc);
    }
    helio_bflux(E, radial, rhoL, vnL, vt1L, vt2L, pL, BnL, Bt1L, Bt2L, rhoR,
                vnR, vt1R, vt2R, pR, BnR, Bt1R, Bt2R, x, y, z, FL, &BnValue);
    F[URHO] = -FL[0];
    F[r.n] = FL[1]; F[r.t1] = FL[2]; F[r.t2] = FL[3];
    F[UENG] = -FL[4];
    F[r.bn] = FL[5]; F[r.bt1] = FL[6]; F[r.bt2] = FL[7];
    *Bn = -BnValue;
  } else {
    double rhoL = WLeft[WRHO], vnL = WLeft[r.n], vt1L = WLeft[r.t1],
           vt2L = WLeft[r.t2], pL = WLeft[WPRES], BnL = WLeft[r.bn],
           Bt1L = WLeft[r.bt1], Bt2L = WLeft[r.bt2];
    rhoL = std::max(E.smallr, rhoL); pL = std::max(E.smallp, pL);
    double rhoR = WRight[WRHO], vnR = WRight[r.n], vt1R = WRight[r.t1],
           vt2R = WRight[r.t2], pR = WRight[WPRES], BnR = WRight[r.bn],
           Bt1R = WRight[r.bt1], Bt2R = WRight[r.bt2];
    rhoR = std::max(E.smallr, rhoR); pR = std::max(E.smallp, pR);
    helio_bflux(E, false, rhoL, vnL, vt1L, vt2L, pL, BnL, Bt1L, Bt2L, rhoR,
                vnR, vt1R, vt2R, pR, BnR, Bt1R, Bt2R, 0, 0, 0, FL, &BnValue);
    F[URHO] = FL[0];
    F[r.n] = FL[1]; F[r.t1] = FL[2]; F[r.t2] = FL[3];
    F[UENG] = FL[4];
    F[r.bn] = FL[5]; F[r.bt1] = FL[6]; F[r.bt2] = FL[7];
    *Bn = BnValue;
  }
  // neutrals (:530-923)
  for (int f = 1; f < E.nf; ++f) {
    int iW = E.iRhoN + (f - 1) * WNUM_E;
    Rot q = rot(iW, dir);
    int iUENG = iW + UENG, iWPRES = iW + WPRES;
    double fr[5];
    bool floor2 = (f == 2), floor3 = (f == 3);
    if (f == 1 || f == 4) {
      if (lohisign == -1 && dir == 0) {
        double rhoL = std::max(E.smallr, WRight[iW]);
        double vnL = WRight[q.n], vt1L = WRight[q.t1], vt2L = WRight[q.t2];
        double pL = std::max(E.smallp, WRight[iWPRES]);
        vnL = std::min(-1.0e-6, vnL);
        double ke = rhoL * (vnL * vnL + vt1L * vt1L + vt2L * vt2L);
        double e = pL / E.hgamma + half * ke;
        F[iW] = rhoL * vnL;
        F[q.n] = rhoL * vnL * vnL + pL;
        F[q.t1] = rhoL * vnL * vt1L;
        F[q.t2] = rhoL * vnL * vt2L;
        F[iUENG] = vnL * (e + pL);
        continue;
      }
      double rhoL, vnL, vt1L, vt2L, pL, rhoR, vnR, vt1R, vt2R, pR;
      if (lohisign == -1) {
        rhoR = std::max(E.smallr, WLeft[iW]); vnR = -WLeft[q.n];
        vt1R = -WLeft[q.t1]; vt2R = -WLeft[q.t2];
        pR = std::max(E.smallp, WLeft[iWPRES]);
        rhoL = std::max(E.smallr, WRight[iW]); vnL = -WRight[q.n];
        vt1L = -WRight[q.t1]; vt2L = -WRight[q.t2];
        pL = std::max(E.smallp, WRight[iWPRES]);
      } else {
        rhoL = std::max(E.smallr, WLeft[iW]); vnL = WLeft[q.n];
        vt1L = WLeft[q.t1]; vt2L = WLeft[q.t2];
        pL = std::max(E.smallp, WLeft[iWPRES]);
        rhoR = std::max(E.smallr, WRight[iW]); vnR = WRight[q.n];
        vt1R = WRight[q.t1]; vt2R = WRight[q.t2];
        pR = std::max(E.smallp, WRight[iWPRES]);
      }
      if (vnL > zero) {
        vnR = vnL; vt1R = vt1L; vt2R = vt2L;
        if (f == 1) rhoR = rhoL;  // neutral 4 keeps rhoR (:885-889)
      }
      vanleer_rs(E, pL, rhoL, vnL, vt1L, vt2L, pR, rhoR, vnR, vt1R, vt2R, fr);
      double sgn = (lohisign == -1) ? -one : one;
      F[iW] = sgn == -one ? -fr[0] : fr[0];
      F[q.n] = fr[1]; F[q.t1] = fr[2]; F[q.t2] = fr[3];
      F[iUENG] = sgn == -one ? -fr[4] : fr[4];
    } else if (floor2 || floor3) {
      double dmin = floor2 ? netFluid2Dens : netFluid3Dens;
      double pmin = floor2 ? netFluid2Pres : netFluid3Pres;
      const double* Ws = (lohisign == -1) ? WRight : WLeft;
      double rhoL = std::max(dmin, Ws[iW]);
      double vnL = Ws[q.n], vt1L = Ws[q.t1], vt2L = Ws[q.t2];
      double pL = std::max(pmin, Ws[iWPRES]);
      vnL = (lohisign == -1) ? std::min(zero, vnL) : std::max(zero, vnL);
      double ke = rhoL * (vnL * vnL + vt1L * vt1L + vt2L * vt2L);
      double e = pL / E.hgamma + half * ke;
      F[iW] = rhoL * vnL;
      F[q.n] = rhoL * vnL * vnL + pL;
      F[q.t1] = rhoL * vnL * vt1L;
      F[q.t2] = rhoL * vnL * vt2L;
      F[iUENG] = vnL * (e + pL);
    }
  }
}

// DEFINE_REGIONS_2F (swlism/SWLISMF.ChF:654-742), Cartesian branch (dx > 0)
int region_cell(const Eq& E, const double* w, int i, int j, int k, double dx) {
  const mfm_params* P = E.P;
  double LIMSUN = 0.6 * P->sun_v;
  double coeff = E.gamma * P->lism_m * P->lism_m * P->lism_t;
  double sunR2 = P->sun_r0 * P->sun_r0;
  double x = (i + half) * dx - P->sun_xc;
  double y = (j + half) * dx - P->sun_yc;
  double z = (E.dims == 3) ? (k + half) * dx - P->sun_zc : 0.0;
  double r2 = (E.dims == 3) ? x * x + y * y + z * z : x * x + y * y;
  double invr = one / std::sqrt(r2);
  double cosx = x * invr, cosy = y * invr, cosz = z * invr;
  double ur = w[WVELX] * cosx + w[WVELY] * cosy;
  if (E.dims == 3) ur = ur + w[WVELZ] * cosz;
  double rho = w[WRHO], p = w[WPRES];
  double TEMP = coeff * p / rho;
  int reg;
  if (std::fabs(ur) >= LIMSUN) {
    if ((TEMP >= 1.0e6) && (r2 > sunR2)) reg = 2; else reg = 3;
    if (E.dims == 3 && ((z < P->reg_safe_z_bot) || (z > P->reg_safe_z_top)))
      reg = 1;
  } else {
    if (TEMP >= P->tmlim) {
      reg = 2;
      if (E.dims == 3 && ((z < P->reg_safe_z_bot) || (z > P->reg_safe_z_top)))
        reg = 1;
    } else {
      reg = 1;
    }
  }
  return reg;
}

// Charge exchange, CHARGE_EXCHANGE_2F/_3F/_4F (swlism/ChargeExchange2F.ChF:
// 294-432, 434-671, 673-993): plasma sink summed over neutrals in order,
// each neutral born in its region or lost.
void charge_exchange_cell(const Eq& E, const double* w, int REG, double dt,
                          double* s) {
  const mfm_params* P = E.P;
  const bool sl = P->ce_single_literals != 0;
  // Fortran default-real literals (REAL*4 unless promoted by the compiler)
  const double c13493 = sl ? (double)13.493f : 13.493;
  const double c0531 = sl ? (double)0.531f : 0.531;
  int nn = E.nf - 1;
  double rho = w[WRHO], ux = w[WVELX], uy = w[WVELY], uz = w[WVELZ],
         p = w[WPRES];
  double U2 = ux * ux + uy * uy + uz * uz;
  double VVTP = p / rho;
  double UH2[3], VPHH[3], VPHP[3], UD[3], UPM[3], UHM[3], UPE[3], UHE[3],
      ARR[3], VVTH[3], VVTHU[3], un[3][3], DENPLU[3];
  for (int f = 0; f < nn; ++f) {
    int b = E.iRhoN + f * WNUM_E;
    double rho1 = w[b], ux1 = w[b + 1], uy1 = w[b + 2], uz1 = w[b + 3],
           p1 = w[b + 4];
    un[f][0] = ux1; un[f][1] = uy1; un[f][2] = uz1;
    UH2[f] = ux1 * ux1 + uy1 * uy1 + uz1 * uz1;
    double DU2 = (ux - ux1) * (ux - ux1) + (uy - uy1) * (uy - uy1) +
                 (uz - uz1) * (uz - uz1);
    VPHH[f] = (ux - ux1) * ux1 + (uy - uy1) * uy1 + (uz - uz1) * uz1;
    VPHP[f] = (ux - ux1) * ux + (uy - uy1) * uy + (uz - uz1) * uz;
    VVTH[f] = 2.0 * p1 / rho1;
    UD[f] = std::sqrt(4.0 * (VVTP + VVTH[f]) / d_PI + DU2);
    UPM[f] = std::sqrt(16.0 * VVTP / d_PI + 2.25 * d_PI * VVTH[f] + 4.0 * DU2);
    UHM[f] = std::sqrt(2.25 * d_PI * VVTP + 16.0 * VVTH[f] / d_PI + 4.0 * DU2);
    UPE[f] = std::sqrt((4.0 * VVTP + 64.0 * VVTH[f] / 9.0) / d_PI + DU2);
    UHE[f] = std::sqrt((64.0 * VVTP / 9.0 + 4.0 * VVTH[f]) / d_PI + DU2);
    double UDS = P->ce_scale_vel * UD[f];
    double t = c13493 - c0531 * std::log(UDS);
    double SIG = (t * t) * std::pow(one - std::exp(-2.946e+9 / UDS), 4.5);
    SIG = P->ce_sigma_lindsay * SIG;
    ARR[f] = SIG * rho1 * rho;
  }
  for (int f = 0; f < nn; ++f) VVTHU[f] = VVTH[f] / UPM[f];
  double h2 = 0, h3 = 0, h4 = 0, h5 = 0;
  for (int f = 0; f < nn; ++f) {
    double t2 = ARR[f] * (ux - un[f][0]) * (UD[f] + VVTHU[f] + VVTP / UHM[f]);
    double t3 = ARR[f] * (uy - un[f][1]) * (UD[f] + VVTHU[f] + VVTP / UHM[f]);
    double t4 = ARR[f] * (uz - un[f][2]) * (UD[f] + VVTHU[f] + VVTP / UHM[f]);
    double t5 = 0.5 * ARR[f] *
                (1.5 * (UHE[f] * VVTP - UPE[f] * VVTH[f]) + UD[f] * (U2 - UH2[f]) +
                 2.0 * (VVTHU[f] * VPHH[f] + VVTP * VPHP[f] / UHM[f]));
    if (f == 0) { h2 = t2; h3 = t3; h4 = t4; h5 = t5; }
    else { h2 = h2 + t2; h3 = h3 + t3; h4 = h4 + t4; h5 = h5 + t5; }
  }
  for (int c = 0; c < E.nc; ++c) s[c] = zero;
  s[URHO] = zero;
  s[UMOMX] = -dt * h2; s[UMOMY] = -dt * h3; s[UMOMZ] = -dt * h4;
  s[UENG] = -dt * h5;
  s[UBX] = zero; s[UBY] = zero; s[UBZ] = zero;
  double PPLU[3] = {0, 0, 0}, ENPLU = 0;
  const double uu[3] = {ux, uy, uz};
  for (int f = 0; f < nn; ++f) {
    DENPLU[f] = -ARR[f] * UD[f];
    for (int d = 0; d < 3; ++d) {
      double t = ARR[f] * (uu[d] * UD[f] - (un[f][d] - uu[d]) * VVTP / UHM[f]);
      PPLU[d] = (f == 0) ? -t : PPLU[d] - t;
    }
    double te = ARR[f] * (0.5 * U2 * UD[f] + VVTP * (VPHP[f] / UHM[f] + 0.75 * UHE[f]));
    ENPLU = (f == 0) ? -te : ENPLU - te;
  }
  for (int f = 0; f < nn; ++f) {
    double PMI[3];
    for (int d = 0; d < 3; ++d)
      PMI[d] = ARR[f] * (un[f][d] * UD[f] + (un[f][d] - uu[d]) * VVTHU[f]);
    double ENMI = ARR[f] * (0.5 * UH2[f] * UD[f] - VPHH[f] * VVTHU[f] +
                            0.75 * VVTH[f] * UPE[f]);
    // neutral 1 is born in region 1 (and 4 in the 2- and 4-fluid kernels,
    // CHARGE_EXCHANGE_3F tests REG == 1 only, :618); neutral k in region k
    // CHARGE_EXCHANGE_MF (:995-1149, the SW/LISM problem's kernel): neutral
    // k is born where REG == k
    // (SWLISMProblem::explicitSource, swlism/SWLISMProblem.cpp:945-1021,
    // calls it for 2 and 3 fluids and CHARGE_EXCHANGE_4F for 4)
    const bool mf = E.P->problem == MFM_PROBLEM_SWLISM && E.nf != 4;
    bool born = mf ? (REG == f + 1)
                   : ((f == 0) ? (REG == 1 || (REG == 4 && E.nf != 3)) : (REG == f + 1));
    double H1, H2, H3, H4, H5;
    if (born && mf) {
      // sum(DENPLU) - DENPLU(iFluid), the sum in index order
      double sum = zero;
      for (int g = 0; g < nn; ++g) sum = sum + DENPLU[g];
      H1 = sum - DENPLU[f];
      H2 = PPLU[0] + PMI[0]; H3 = PPLU[1] + PMI[1]; H4 = PPLU[2] + PMI[2];
      H5 = ENPLU + ENMI;
    } else if (born) {
      // sum of the other fluids' DENPLU in index order (zero if none)
      bool first = true;
      H1 = zero;
      for (int g = 0; g < nn; ++g) {
        if (g == f) continue;
        if (first) { H1 = DENPLU[g]; first = false; }
        else H1 = H1 + DENPLU[g];
      }
      H2 = PPLU[0] + PMI[0]; H3 = PPLU[1] + PMI[1]; H4 = PPLU[2] + PMI[2];
      H5 = ENPLU + ENMI;
    } else {
      H1 = -DENPLU[f];
      H2 = PMI[0]; H3 = PMI[1]; H4 = PMI[2];
      H5 = ENMI;
    }
    int b = E.iRhoN + f * WNUM_E;
    s[b] = -dt * H1;
    s[b + 1] = -dt * H2; s[b + 2] = -dt * H3; s[b + 3] = -dt * H4;
    s[b + 4] = -dt * H5;
  }
}

// ================================================================ SW/LISM ==
// The 2-D axisymmetric solar-wind / LISM problem (swlism/SWLISMProblem.cpp,
// swlism/SWLISMF.ChF) in coord_sys = 1 (CS_Axisymmetric: Cartesian fluxes
// and volumes, src/CSHCartesian.cpp:254-305, plus the geometric sources).

// LISMINIT (swlism/SWLISMF.ChF:512-652): the LISM potential flow past the
// sphere initR on a ghost box; populations 2.. copy the interior cell next
// to the boundary (right / left boundary in x, top in y).
void swlism_lisminit(const Eq& E, Fab& W, int lohisign, int dir, const Box& box, double dx) {
  const mfm_params* P = E.P;
  FORBOX(box, i, j, k) {
    double x = (i + half) * dx - P->sun_xc;
    double y = (j + half) * dx - P->sun_yc;
    double r = std::sqrt(x * x + y * y);
    double sinT = y / r, cosT = x / r;
    double h = P->init_r / r;
    h = h * h * h;
    double ux = one - (one - 1.5 * sinT * sinT) * h;
    double uy = -1.5 * sinT * cosT * h;
    double Bx = P->lism_bx * ux, By = P->lism_bx * uy;
    W(i, j, k, WRHO) = one;
    W(i, j, k, WVELX) = ux; W(i, j, k, WVELY) = uy; W(i, j, k, WVELZ) = zero;
    W(i, j, k, WPRES) = P->lism_p;
    W(i, j, k, WBX) = Bx; W(i, j, k, WBY) = By; W(i, j, k, WBZ) = zero;
  }
  if (E.nf > 1) {
    int s0 = E.iRhoN;
    FORBOX(box, i, j, k) {
      W(i, j, k, s0 + WRHO) = P->net_rho;
      W(i, j, k, s0 + WVELX) = W(i, j, k, WVELX);
      W(i, j, k, s0 + WVELY) = W(i, j, k, WVELY);
      W(i, j, k, s0 + WVELZ) = W(i, j, k, WVELZ);
      W(i, j, k, s0 + WPRES) = P->net_p;
    }
  }
  if (E.nf > 2) {
    int vb = E.iRhoN + WNUM_E, ve = E.iRhoN + (E.nf - 1) * WNUM_E - 1;
    if (dir == 0) {
      int indb = lohisign == 1 ? box.lo[0] - 1 : box.hi[0] + 1;
      FORBOX(box, i, j, k) for (int v = vb; v <= ve; ++v) W(i, j, k, v) = W(indb, j, k, v);
    } else if (lohisign == 1) {
      int indb = box.lo[1] - 1;
      FORBOX(box, i, j, k) for (int v = vb; v <= ve; ++v) W(i, j, k, v) = W(i, indb, k, v);
    }
  }
}

// SWLISMProblem::fillGhostCellsDefault (swlism/SWLISMProblem.cpp:453-635):
// LISM inflow / outflow states on the x boundaries and the top, and the
// mirror image across the symmetry axis (y = 0) at the bottom, with the
// y velocities and B_y reversed.  At most 4 ghost layers.
void swlism_fill_ghost(const Eq& E, Fab& W, int dir, double dx) {
  if (!E.physical(dir) || dir > 1) return;
  Box WBox = W.b;
  Box tmp = E.clip(WBox);
  if (WBox.hi[dir] != tmp.hi[dir]) {
    int nGS = std::min(std::abs(WBox.hi[dir] - tmp.hi[dir]), 4);
    Box bb = WBox; bb.lo[dir] = tmp.hi[dir] + 1; bb.hi[dir] = tmp.hi[dir] + nGS;
    swlism_lisminit(E, W, 1, dir, bb, dx);
  }
  if (WBox.lo[dir] != tmp.lo[dir]) {
    int nGS = std::min(std::abs(WBox.lo[dir] - tmp.lo[dir]), 4);
    Box bb = WBox; bb.lo[dir] = tmp.lo[dir] - nGS; bb.hi[dir] = tmp.lo[dir] - 1;
    if (dir == 0) {
      swlism_lisminit(E, W, -1, dir, bb, dx);
    } else {
      int src = tmp.lo[1], des = tmp.lo[1] - 1;
      for (int q = 0; q < nGS; ++q, ++src, --des) {
        for (int k = WBox.lo[2]; k <= WBox.hi[2]; ++k)
          for (int i = WBox.lo[0]; i <= WBox.hi[0]; ++i)
            for (int c = 0; c < E.nc; ++c) W(i, des, k, c) = W(i, src, k, c);
      }
      FORBOX(bb, i, j, k) {
        W(i, j, k, WVELY) = -W(i, j, k, WVELY);
        W(i, j, k, WBY) = -W(i, j, k, WBY);
        for (int f = 1; f < E.nf; ++f) {
          int s0 = E.iRhoN + (f - 1) * WNUM_E;
          W(i, j, k, s0 + WVELY) = -W(i, j, k, s0 + WVELY);
        }
      }
    }
  }
}

// BFLUXGAS (swlism/SWLISMF.ChF:1434-1553): characteristic outflow flux of
// a neutral population; the subsonic branches that return early keep the
// Riemann flux passed in.
void swlism_bfluxgas(const Eq& E, double rho_in, double vn_in, double vt1_in,
                     double vt2_in, double p_in, double rho_out, double vn_out,
                     double vt1_out, double vt2_out, double p_out, double* F5) {
  const double BET = E.hgamma, GAM = E.gamma, GAL = E.invhgamma;
  double UBO = vn_in, VBO = vt1_in, WBO = vt2_in;
  double RBO = std::max(E.smallr, rho_in), PBO = std::max(E.smallp, p_in);
  double KEBO = half * (vn_in * vn_in + vt1_in * vt1_in + vt2_in * vt2_in);
  double EBO = PBO / E.hgamma + rho_in * KEBO;
  double SZBO = std::sqrt(GAM * PBO / RBO);
  double U1 = vn_out, V1 = vt1_out, W1 = vt2_out;
  double R1 = std::max(E.smallr, rho_out), P1 = std::max(E.smallp, p_out);
  double KE1 = half * (vn_out * vn_out + vt1_out * vt1_out + vt2_out * vt2_out);
  double E1 = P1 / E.hgamma + rho_out * KE1;
  double SZ1 = std::sqrt(GAM * P1 / R1);
  double EN = PBO / std::pow(RBO, GAM), EN1 = P1 / std::pow(R1, GAM);
  double U16 = BET * (UBO + 2.0 * SZBO / BET) / (GAM + 1.0);
  double SZ16 = U16;
  double R16 = std::pow(SZ16 * SZ16 / (GAM * EN), GAL);
  double V16 = VBO, W16 = WBO;
  double P16 = SZ16 * SZ16 * R16 / GAM;
  double E16 = P16 * GAL + .5 * R16 * (U16 * U16 + V16 * V16 + W16 * W16);
  double U56 = BET * (U1 - 2.0 * SZ1 / BET) / (GAM + 1.0);
  double SZ56 = -U56;
  double R56 = std::pow(SZ56 * SZ56 / (GAM * EN1), GAL);
  double V56 = V1, W56 = W1;
  double P56 = SZ56 * SZ56 * R56 / GAM;
  double E56 = P56 * GAL + .5 * R56 * (U56 * U56 + V56 * V56 + W56 * W56);
  if ((UBO - SZBO) >= 0.0) {
    F5[0] = RBO * UBO; F5[1] = RBO * UBO * UBO + PBO; F5[2] = RBO * UBO * VBO;
    F5[3] = RBO * UBO * WBO; F5[4] = (EBO + PBO) * UBO;
    return;
  }
  if (((UBO - SZBO) < 0.0) && (UBO >= 0.0)) {
    if ((U1 - SZ1) <= 0.0) return;
    F5[0] = R16 * U16; F5[1] = R16 * U16 * U16 + P16; F5[2] = R16 * U16 * V16;
    F5[3] = R16 * U16 * W16; F5[4] = (E16 + P16) * U16;
    return;
  }
  if ((UBO < 0.0) && (UBO + SZBO) > 0.0) {
    if ((U1 + SZ1) >= 0.0) return;
    F5[0] = R56 * U56; F5[1] = R56 * U56 * U56 + P56; F5[2] = R56 * U56 * V56;
    F5[3] = R56 * U56 * W56; F5[4] = (E56 + P56) * U56;
    return;
  }
  if ((UBO + SZBO) < 0.0) {
    F5[0] = R1 * U1; F5[1] = R1 * U1 * U1 + P1; F5[2] = R1 * U1 * V1;
    F5[3] = R1 * U1 * W1; F5[4] = (E1 + P1) * U1;
  }
}

// SWLISMProblem::fluxBC (swlism/SWLISMProblem.cpp:727-862) at one boundary
// face, after the Riemann fluxes of the face states are in out:
// SWLISMBC (swlism/SWLISMF.ChF:1025-1205) -- BFLUX for the plasma and
// BFLUXGAS for the neutrals on the right and top boundaries, the mean normal
// field on the left -- and a zero normal field on the axis.
void swlism_bc_face(const Eq& E, int lohisign, int dir, const double* wl,
                    const double* wr, double* out, double* bn) {
  Rot r = rot(0, dir);
  if (lohisign == 1) {
    double FL[8], BnValue;
    // SWLISMBC passes the face states unclamped (BFLUX, HeliosphericF.ChF:1363)
    helio_bflux(E, false, wl[WRHO], wl[r.n], wl[r.t1], wl[r.t2], wl[WPRES], wl[r.bn],
                wl[r.bt1], wl[r.bt2], wr[WRHO], wr[r.n], wr[r.t1], wr[r.t2], wr[WPRES],
                wr[r.bn], wr[r.bt1], wr[r.bt2], zero, zero, zero, FL, &BnValue);
    out[URHO] = FL[0]; out[r.n] = FL[1]; out[r.t1] = FL[2]; out[r.t2] = FL[3];
    out[UENG] = FL[4]; out[r.bn] = FL[5]; out[r.bt1] = FL[6]; out[r.bt2] = FL[7];
    *bn = BnValue;
    for (int f = 1; f < E.nf; ++f) {
      int s0 = E.iRhoN + (f - 1) * WNUM_E;
      Rot q = rot(s0, dir);
      double F5[5] = {out[q.rho], out[q.n], out[q.t1], out[q.t2], out[q.p]};
      swlism_bfluxgas(E, wl[q.rho], wl[q.n], wl[q.t1], wl[q.t2], wl[q.p], wr[q.rho],
                      wr[q.n], wr[q.t1], wr[q.t2], wr[q.p], F5);
      out[q.rho] = F5[0]; out[q.n] = F5[1]; out[q.t1] = F5[2]; out[q.t2] = F5[3];
      out[q.p] = F5[4];
    }
  } else if (dir == 0) {
    *bn = half * (wl[r.bn] + wr[r.bn]);
  } else {
    *bn = zero;  // the symmetry axis (swlism/SWLISMProblem.cpp:854)
  }
}

// SWLISMREINIT (swlism/SWLISMF.ChF:305-403): the supersonic solar wind
// inside R0 (early exit on the FAB bounds, as the reference).
void swlism_reinit(const Eq& E, Fab& U, double dx, const Box& box) {
  const mfm_params* P = E.P;
  double sunR = P->sun_r0;
  double xl = U.b.lo[0] * dx - P->sun_xc, yl = U.b.lo[1] * dx - P->sun_yc;
  double xh = (U.b.hi[0] + 1) * dx - P->sun_xc, yh = (U.b.hi[1] + 1) * dx - P->sun_yc;
  if ((xl > sunR) || (yl > sunR)) return;
  if ((xh < -sunR) || (yh < -sunR)) return;
  double sunR2 = sunR * sunR;
  FORBOX(box, i, j, k) {
    double x = (i + half) * dx - P->sun_xc;
    double y = (j + half) * dx - P->sun_yc;
    double r2 = x * x + y * y;
    if (r2 <= sunR2) {
      double r = std::sqrt(r2);
      double rho = P->sun_rho / (r * r);
      double ux = P->sun_v * x / r, uy = P->sun_v * y / r;
      double p = P->sun_p / std::pow(r, two * E.gamma);
      double e0 = p / E.hgamma + half * rho * P->sun_v * P->sun_v;
      U(i, j, k, URHO) = rho;
      U(i, j, k, UMOMX) = rho * ux; U(i, j, k, UMOMY) = rho * uy; U(i, j, k, UMOMZ) = zero;
      U(i, j, k, UENG) = e0;
      U(i, j, k, UBX) = zero; U(i, j, k, UBY) = zero; U(i, j, k, UBZ) = zero;
      if (E.nf > 2) {
        int s0 = E.iRhoN + WNUM_E;
        U(i, j, k, s0 + URHO) = netFluid2Dens;
        U(i, j, k, s0 + UMOMX) = zero; U(i, j, k, s0 + UMOMY) = zero;
        U(i, j, k, s0 + UMOMZ) = zero;
        U(i, j, k, s0 + UENG) = netFluid2Pres / E.hgamma;
      }
      if (E.nf > 3) {
        int s0 = E.iRhoN + 2 * WNUM_E;
        U(i, j, k, s0 + URHO) = netFluid3Dens;
        U(i, j, k, s0 + UMOMX) = zero; U(i, j, k, s0 + UMOMY) = zero;
        U(i, j, k, s0 + UMOMZ) = zero;
        U(i, j, k, s0 + UENG) = netFluid3Pres / E.hgamma;
      }
    }
  }
}

// Geometric sources of coord_sys = 1: SOURCEAXISYMMETRIC
// (srcIdeal/PatchIdealMHDF.ChF:839-903) and, with neutral fluids,
// SOURCEAXISYMMETRIC_MF (srcIdeal/PatchMHDMFF.ChF:152-244) of one cell;
// s[0 .. nc-1] (zero where the reference writes zero).
void axisym_source_cell(const Eq& E, const double* w, int j, double dt, double dx, double* s) {
  double y = dx * (j + half);
  double h = -dt / y;
  double Rho = w[WRHO], Ux = w[WVELX], Uy = w[WVELY], Uz = w[WVELZ], P = w[WPRES];
  double Bx = w[WBX], By = w[WBY], Bz = w[WBZ];
  double By4Pi = d_1_4PI * By;
  double BxB = d_1_4PI * (Bx * Bx + By * By + Bz * Bz);
  double UxU = half * (Ux * Ux + Uy * Uy + Uz * Uz);
  double e = P / E.hgamma + Rho * UxU;
  for (int c = 0; c < E.nc; ++c) s[c] = zero;
  s[URHO] = h * Rho * Uy;
  s[UMOMX] = h * (Rho * Ux * Uy - By4Pi * Bx);
  s[UMOMY] = h * (Rho * Uy * Uy - By4Pi * By);
  s[UMOMZ] = zero;
  s[UENG] = h * (Uy * (e + P + BxB) - By4Pi * (Bx * Ux + By * Uy));
  s[UBX] = h * (Uy * Bx - Ux * By);
  s[UBY] = zero;
  s[UBZ] = zero;
  for (int f = 1; f < E.nf; ++f) {
    int sh = E.iRhoN + (f - 1) * WNUM_E;
    double RhoN = w[sh + WRHO], UxN = w[sh + WVELX], UyN = w[sh + WVELY],
           UzN = w[sh + WVELZ], PN = w[sh + WPRES];
    double UxUN = half * (UxN * UxN + UyN * UyN + UzN * UzN);
    double eN = PN / E.hgamma + RhoN * UxUN;
    double hu = h * UyN;
    s[sh + URHO] = hu * RhoN;
    s[sh + UMOMX] = hu * RhoN * UxN;
    s[sh + UMOMY] = hu * RhoN * UyN;
    s[sh + UMOMZ] = zero;
    s[sh + UENG] = hu * (eN + PN);
  }
}

// HELIOREINIT_DEFAULT (swlism/HeliosphericPlasmaBCF.ChF:1435-1578)
void helio_reinit(const Eq& E, Fab& U, double dx, double t, const Box& box) {
  const mfm_params* P = E.P;
  double Rinit = P->sun_r0;
  double xl = U.b.lo[0] * dx - P->sun_xc, yl = U.b.lo[1] * dx - P->sun_yc;
  double zl = U.b.lo[2] * dx - P->sun_zc;
  double xh = (U.b.hi[0] + 1) * dx - P->sun_xc, yh = (U.b.hi[1] + 1) * dx - P->sun_yc;
  double zh = (U.b.hi[2] + 1) * dx - P->sun_zc;
  bool d3 = E.dims == 3;
  if ((xl > Rinit) || (yl > Rinit) || (d3 && zl > Rinit)) return;
  if ((xh < -Rinit) || (yh < -Rinit) || (d3 && zh < -Rinit)) return;
  double sunR2 = Rinit * Rinit;
  double sinTILT = std::sin(P->sun_tilt), cosTILT = std::cos(P->sun_tilt);
  double sunBeta = P->omega * eos_AU / (P->sun_v * P->lism_v_dim);
  FORBOX(box, i, j, k) {
    double x = (i + half) * dx - P->sun_xc;
    double y = (j + half) * dx - P->sun_yc;
    double z = d3 ? (k + half) * dx - P->sun_zc : zero;
    double r2 = d3 ? x * x + y * y + z * z : x * x + y * y;
    if (r2 <= sunR2) {
      double r = std::sqrt(r2);
      double rr = r / P->sun_int_bc_radius;
      double xy = std::sqrt(x * x + y * y);
      double cosT = z / r;
      double sinT = std::sqrt(one - cosT * cosT);
      double cosF = x / xy, sinF = y / xy;
      double theta = std::acos(cosT);
      double cosBr = std::cos(sunBeta * (r - P->sun_v * t));
      double sinBr = std::sin(sunBeta * (r - P->sun_v * t));
      double sinB = sinF * cosBr + cosF * sinBr;
      double cosB = cosF * cosBr - sinF * sinBr;
      double thetaS = sinTILT * sinB /
                      std::sqrt((cosTILT * cosB) * (cosTILT * cosB) + sinB * sinB);
      thetaS = std::asin(thetaS) + d_PI_2;
      double rho = P->sun_rho / (rr * rr);
      double ux = P->sun_v * sinT * cosF, uy = P->sun_v * sinT * sinF,
             uz = P->sun_v * cosT;
      double p = P->sun_p / std::pow(rr, two * E.gamma);
      double br = P->dir_br_n * P->sun_b / (rr * rr);
      double bp = -sunBeta * br * r * sinT;
      double bx = br * sinT * cosF - bp * sinF;
      double by = br * sinT * sinF + bp * cosF;
      double bz = br * cosT;
      double b = bx * bx + by * by + bz * bz;
      double e0 = p / E.hgamma + half * (rho * P->sun_v * P->sun_v + d_1_4PI * b);
      U(i, j, k, URHO) = rho;
      U(i, j, k, UMOMX) = rho * ux; U(i, j, k, UMOMY) = rho * uy;
      U(i, j, k, UMOMZ) = rho * uz;
      U(i, j, k, UENG) = e0;
      double sg = (P->sun_b_monopolar == 1 || theta < thetaS) ? one : -one;
      U(i, j, k, UBX) = sg == one ? bx : -bx;
      U(i, j, k, UBY) = sg == one ? by : -by;
      U(i, j, k, UBZ) = sg == one ? bz : -bz;
    }
  }
}

// POSTPROCESSING (srcIdeal/PatchIdealMHDF.ChF:1019-1070) and
// POSTPROCESSING_MF (srcIdeal/PatchMHDMFF.ChF:301-511)
void floors_cell(const Eq& E, double* u) {
  if (E.nf == 1) {
    double Rho = u[URHO];
    int iWas = 0;
    if (Rho < E.smallr) { Rho = E.smallr; iWas = 1; }
    double invRho = one / Rho;
    double Ux = u[UMOMX] * invRho, Uy = u[UMOMY] * invRho, Uz = u[UMOMZ] * invRho;
    double Ee = u[UENG];
    double Bx = u[UBX], By = u[UBY], Bz = u[UBZ];
    double BxB = d_1_8PI * (Bx * Bx + By * By + Bz * Bz);
    double UxU = half * (Ux * Ux + Uy * Uy + Uz * Uz);
    double P = E.hgamma * (Ee - BxB - Rho * UxU);
    if (P < E.smallp) { P = E.smallp; iWas = 1; }
    if (iWas == 1) {
      u[URHO] = Rho;
      u[UENG] = P / E.hgamma + Rho * UxU + BxB;
    }
    return;
  }
  {
    double Rho = u[URHO];
    double invRho = one / Rho;
    double Ux = u[UMOMX] * invRho, Uy = u[UMOMY] * invRho, Uz = u[UMOMZ] * invRho;
    double Ee = u[UENG];
    double Bx = u[UBX], By = u[UBY], Bz = u[UBZ];
    double BxB = d_1_8PI * (Bx * Bx + By * By + Bz * Bz);
    double UxU = half * (Ux * Ux + Uy * Uy + Uz * Uz);
    double P = E.hgamma * (Ee - BxB - Rho * UxU);
    int iWas = 0;
    if (Rho < E.smallr) { Rho = E.smallr; iWas = 1; }
    if (P < E.smallp) { P = E.smallp; iWas = 1; }
    if (iWas == 1) {
      u[URHO] = Rho;
      u[UMOMX] = Rho * Ux; u[UMOMY] = Rho * Uy; u[UMOMZ] = Rho * Uz;
      u[UENG] = P / E.hgamma + Rho * UxU + BxB;
    }
  }
  for (int f = 1; f < E.nf; ++f) {
    int s = E.iRhoN + (f - 1) * WNUM_E;
    double dmin = E.smallr, pmin = E.smallp;
    if (f == 2 || f == 4) { dmin = netFluid2Dens; pmin = netFluid2Pres; }
    if (f == 3) { dmin = netFluid3Dens; pmin = netFluid3Pres; }
    double Rho = u[s + URHO];
    double invRho = one / Rho;
    double Ux = u[s + UMOMX] * invRho, Uy = u[s + UMOMY] * invRho,
           Uz = u[s + UMOMZ] * invRho;
    double Ee = u[s + UENG];
    double UxU = half * (Ux * Ux + Uy * Uy + Uz * Uz);
    double P = E.hgamma * (Ee - Rho * UxU);
    int iWas = 0;
    if (Rho < dmin) { Rho = dmin; iWas = 1; }
    if (P < pmin) { P = pmin; iWas = 1; }
    if (iWas == 1) {
      u[s + URHO] = Rho;
      u[s + UMOMX] = Rho * Ux; u[s + UMOMY] = Rho * Uy; u[s + UMOMZ] = Rho * Uz;
      u[s + UENG] = P / E.hgamma + Rho * UxU;
    }
  }
}

// MAXWAVESPEED (srcIdeal/PatchIdealMHDF.ChF:5-63) + MAXWAVESPEED_E
// (srcIdeal/PatchEulerF.ChF:5-51)
double max_wave_speed_cell(const Eq& E, const double* u) {
  double rho = u[URHO];
  double invrho = one / rho;
  double ru = u[UMOMX], rv = u[UMOMY], rw = u[UMOMZ];
  double uu = ru * invrho, v = rv * invrho, w = rw * invrho;
  double ke = ru * uu + rv * v + rw * w;
  double bx = u[UBX], by = u[UBY], bz = u[UBZ];
  double b = d_1_4PI * (bx * bx + by * by + bz * bz);
  double p = E.hgamma * (u[UENG] - half * (ke + b));
  p = std::max(p, E.smallp);
  double a2 = E.gamma * p * invrho;
  double as2 = half * (a2 + invrho * b);
  double as4 = as2 * as2;
  double bn2 = d_1_4PI * bx * bx * invrho;
  double cf2 = as2 + std::sqrt(as4 - a2 * bn2);
  double umax = std::fabs(uu) + std::sqrt(cf2);
  bn2 = d_1_4PI * by * by * invrho;
  cf2 = as2 + std::sqrt(as4 - a2 * bn2);
  umax = std::max(std::fabs(v) + std::sqrt(cf2), umax);
  bn2 = d_1_4PI * bz * bz * invrho;
  cf2 = as2 + std::sqrt(as4 - a2 * bn2);
  umax = std::max(std::fabs(w) + std::sqrt(cf2), umax);
  double cmax = std::max(umax, zero);  // dtFab.setVal(0) then max
  for (int f = 1; f < E.nf; ++f) {
    int s = E.iRhoN + (f - 1) * WNUM_E;
    double r = std::max(u[s], E.smallr);
    double ir = one / r;
    double nru = u[s + 1], nrv = u[s + 2], nrw = u[s + 3];
    double nu = nru * ir, nv = nrv * ir, nw = nrw * ir;
    double nke = nru * nu + nrv * nv + nrw * nw;
    double np = E.hgamma * (u[s + 4] - half * nke);
    np = std::max(np, E.smallp);
    double a = std::sqrt(E.gamma * np * ir);
    double um = std::max(std::max(std::fabs(nu), std::fabs(nv)), std::fabs(nw)) + a;
    cmax = std::max(um, cmax);
  }
  return cmax;
}

// =================================================== EulerStep pieces ======
struct FaceSet {
  Fab WM[3], WP[3];
  Fab Bn[3];
};

void fill_ghost(const Eq& E, Fab& W, int dir, double dx) {
  if (E.P->problem == MFM_PROBLEM_RIEMANN) riemann_fill_ghost(E, W, dir);
  else if (E.P->problem == MFM_PROBLEM_HELIO) helio_fill_ghost(E, W, dir);
  else if (E.P->problem == MFM_PROBLEM_SWLISM) swlism_fill_ghost(E, W, dir, dx);
}

// ========================================= characteristic reconstruction ==
// *FACESLOPES (src/reconstructionF.ChF): limited left / right slopes of a
// cell from its one-sided differences dSL = W - W(i-1), dSR = W(i+1) - W
// (in characteristic variables when char_reconstruction is on).
void face_slopes(int lim, double dSL, double dSR, double* SL, double* SR) {
  double S = zero, SLOPL = zero, SLOPR = zero;
  switch (lim) {
    case MFM_LIM_CD:  // CDFACESLOPES :1081-1094
      *SL = half * dSL; *SR = half * dSR;
      return;
    case MFM_LIM_MINMOD:  // MINMODFACESLOPES :157-177
      S = half * minmod(dSR, dSL);
      break;
    case MFM_LIM_SUPERBEE:  // SUPERBEEFACESLOPES :356-388
      if (dSL * dSR > zero) {
        double AR = std::fabs(dSR), AL = std::fabs(dSL);
        if (AL <= half * AR) S = dSL;
        else if (AL <= AR) S = half * dSR;
        else if (AL <= AR + AR) S = half * dSL;
        else S = dSR;
      }
      break;
    case MFM_LIM_HYPERBEE:  // HYPERBEEFACESLOPES :466-490
      if (dSL * dSR > zero) {
        double A = one + std::sqrt(dSR / dSL);
        S = dSR / (A * A);
        S = S + S;
      }
      break;
    case MFM_LIM_HARMONIC:  // HARMONICFACESLOPES :526-548
      if (dSL * dSR > zero) S = dSR * dSL / (dSL + dSR);
      break;
    case MFM_LIM_VANALBADA: {  // VANALBADAFACESLOPES :591-620
      const double dSave = 1.0e-6, dSaveS = dSave + dSave;
      if (dSL * dSR > zero)
        S = half * (dSR * dSL + dSave) * (dSR + dSL) / (dSR * dSR + dSL * dSL + dSaveS);
      break;
    }
    case MFM_LIM_MC:  // MCFACESLOPES :664-694
      if (dSL * dSR > zero) {
        double AR = std::fabs(dSR), AL = std::fabs(dSL);
        if (3.0 * AL <= AR) S = dSL;
        else if (AL <= 3.0 * AR) S = 0.25 * (dSR + dSL);
        else S = dSR;
      }
      break;
    case MFM_LIM_VANLEER: {  // VANLEERFACESLOPES :843-881, kappa = 1/3
      const double dSLim = one / 3.0;
      const double BT = (3.0 - dSLim) / (one - dSLim);
      const double SL1 = 0.25 * (one - dSLim), SL2 = 0.25 * (one + dSLim);
      if (dSL * dSR > zero) {
        double SLm = std::min(std::fabs(dSR), std::fabs(BT * dSL));
        double SRm = std::min(std::fabs(dSL), std::fabs(BT * dSR));
        SLOPL = SL1 * SLm + SL2 * SRm;
        SLOPR = SL1 * SRm + SL2 * SLm;
        if (dSR < zero) { SLOPL = -SLOPL; SLOPR = -SLOPR; }
      }
      *SL = SLOPL; *SR = SLOPR;
      return;
    }
    case MFM_LIM_KOREN: {  // KORENFACESLOPES :920-948
      const double dSave = 1.0e-10;
      if (dSL * dSR > zero) {
        double H = dSL * dSR / (two * (dSL * dSL + dSR * dSR + dSave) - dSL * dSR);
        SLOPL = (dSL + half * dSR) * H;
        SLOPR = (dSR + half * dSL) * H;
      }
      *SL = SLOPL; *SR = SLOPR;
      return;
    }
    case MFM_LIM_VENKAT: {  // VENKATAKRISHNANFACESLOPES :990-1020
      const double dSave = 1.0e-10, dSave2 = dSave * dSave;
      if (dSL * dSR > zero) {
        double H = (dSR + dSL) / (6.0 * (dSR * dSR + dSL * dSL + dSave2));
        SLOPL = (dSR + two * dSL) * (dSL + dSave) * H;
        SLOPR = (dSL + two * dSR) * (dSR + dSave) * H;
      }
      *SL = SLOPL; *SR = SLOPR;
      return;
    }
    case MFM_LIM_WENO3: {  // WENO3FACESLOPES :1138-1167
      const double dSave = 1.0e-6;
      double BR = dSR * dSR + dSave, AR = one / (BR * BR);
      double BL = dSL * dSL + dSave, AL = one / (BL * BL);
      double H = AR / (AR + two * AL);
      *SL = half * (dSL + H * (dSR - dSL));
      H = AL / (AL + two * AR);
      *SR = half * (dSR - H * (dSR - dSL));
      return;
    }
    case MFM_LIM_WENO3YC: {  // WENO3YCFACESLOPES :1214-1247
      const double dSave = 1.0e-6;
      double BR = dSR * dSR + dSave, BL = dSL * dSL + dSave;
      double D2P = dSR - dSL;
      D2P = D2P * D2P;
      double AL = one + D2P / BL, AR = one + D2P / BR;
      double H = AR / (AR + two * AL);
      *SL = half * (dSL + H * (dSR - dSL));
      H = AL / (AL + two * AR);
      *SR = half * (dSR - H * (dSR - dSL));
      return;
    }
    default:  // slopes::faceSlopes leaves unlimited slopes alone
      *SL = dSL; *SR = dSR;
      return;
  }
  *SL = S; *SR = S;
}

// Eigen-structure of the ideal-MHD system at one cell in direction dir, as
// CHARANALYSISF / CHARSYNTHESISF (srcIdeal/PatchIdealMHDF.ChF:1167-1505)
// form it from the cell's primitive state.  The fast/slow weights switch to
// the degenerate branch below dfs_eps: 1e-12 in the analysis, 1e-8 in the
// synthesis.
struct CharBasis {
  int n, t1, t2, bn, bt1, bt2;
  double Rho, cc, c, af, as, alphaF, alphaS, betaY, betaZ, BX;
};

CharBasis char_basis(const Eq& E, const double* w, int dir, double dfs_eps) {
  CharBasis b;
  b.n = 1 + dir; b.t1 = 1 + (dir + 1) % 3; b.t2 = 1 + (dir + 2) % 3;
  b.bn = 5 + dir; b.bt1 = 5 + (dir + 1) % 3; b.bt2 = 5 + (dir + 2) % 3;
  double PGas = std::max(E.smallp, w[4]);
  double Rho = std::max(E.smallr, w[0]);
  double BX = w[b.bn], BY = w[b.bt1], BZ = w[b.bt2];
  double cc = E.gamma * PGas / Rho;
  double c = std::sqrt(cc);
  double BX2 = BX * BX, BY2 = BY * BY, BZ2 = BZ * BZ;
  double BYZ2 = BY2 + BZ2;
  double BxBH = d_1_8PI * (BX2 + BYZ2);
  double bb = d_1_4PI * BX2 / Rho;
  double bs = std::sqrt(bb);
  double aas, as, aaf, af, betaY, betaZ;
  if (BYZ2 < E.smallB * E.smallB) {
    betaY = d_SQRT_2;
    betaZ = d_SQRT_2;
    if (bb > cc) { aaf = bb; aas = cc; af = bs; as = c; }
    else { aaf = cc; aas = bb; af = c; as = bs; }
  } else {
    double hyz = std::sqrt(BYZ2);
    betaY = BY / hyz;
    betaZ = BZ / hyz;
    double szb = cc + (BxBH + BxBH) / Rho;
    double h = std::sqrt(szb * szb - four * cc * bb);
    aas = half * (szb - h);
    as = std::sqrt(aas);
    aaf = half * (szb + h);
    af = std::sqrt(aaf);
  }
  double dfs = aaf - aas, alphaF, alphaS;
  if (dfs < dfs_eps) { alphaF = one; alphaS = zero; }
  else if ((cc - aas) <= zero) { alphaF = zero; alphaS = one; }
  else if ((aaf - cc) <= zero) { alphaF = one; alphaS = zero; }
  else {
    alphaF = std::sqrt((cc - aas) / dfs);
    alphaS = std::sqrt((aaf - cc) / dfs);
  }
  b.Rho = Rho; b.cc = cc; b.c = c; b.af = af; b.as = as;
  b.alphaF = alphaF; b.alphaS = alphaS; b.betaY = betaY; b.betaZ = betaZ; b.BX = BX;
  return b;
}

// CHARANALYSISF (:1252-1325): primitive differences d (components 0..7,
// Cartesian order) -> characteristic variables a (entropy, div(B), two
// Alfven, two fast, two slow).
void char_analysis(const CharBasis& b, const double* d, double* a) {
  // SIGN(one, BX) taken with a negative zero counting as positive, the rule
  // of the synthesis's BX < 0 test: with the IEEE sign of -0.0 the two
  // transforms stop being inverse where the normal field is exactly zero,
  // and the result would hang on the sign of that zero, which the order of
  // the update's floating-point operations decides.
  double siBX = b.BX < zero ? -one : one;
  double dP = d[4], dRho = d[0], dUn = d[b.n], dUt1 = d[b.t1], dUt2 = d[b.t2];
  double dBn = d[b.bn], dBt1 = d[b.bt1], dBt2 = d[b.bt2];
  a[0] = dRho - dP / b.cc;
  a[1] = dBn;
  double h1 = d_SQRT_2 * (b.betaY * dUt2 - b.betaZ * dUt1);
  double h2 = siBX * (b.betaZ * dBt1 - b.betaY * dBt2) / (two * std::sqrt(two * d_PI * b.Rho));
  a[2] = h1 + h2;
  a[3] = h1 - h2;
  h1 = dP / (two * b.Rho * b.cc);
  h2 = dUn / (two * b.cc);
  double h3 = (b.betaY * dBt1 + b.betaZ * dBt2) / (four * b.c * std::sqrt(d_PI * b.Rho));
  double h4 = siBX * (b.betaY * dUt1 + b.betaZ * dUt2) / (two * b.cc);
  const double aF = b.alphaF, aS = b.alphaS, af = b.af, as = b.as;
  a[4] = aF * (h1 + af * h2) + aS * (h3 - as * h4);
  a[5] = aF * (h1 - af * h2) + aS * (h3 + as * h4);
  a[6] = aS * (h1 + as * h2) - aF * (h3 - af * h4);
  a[7] = aS * (h1 - as * h2) - aF * (h3 + af * h4);
}

// CHARSYNTHESISF (:1438-1502): characteristic variables -> primitive
// differences.
void char_synthesis(const CharBasis& b, const double* a, double* d) {
  double betaYX = b.betaY, betaZX = b.betaZ;
  if (b.BX < zero) { betaYX = -b.betaY; betaZX = -b.betaZ; }
  double alpha1 = a[0], alpha2 = a[1], alpha3 = a[2], alpha4 = a[3];
  double alpha5 = a[4], alpha6 = a[5], alpha7 = a[6], alpha8 = a[7];
  alpha3 = d_SQRT_2 * alpha3;
  alpha4 = d_SQRT_2 * alpha4;
  double alph3p4 = alpha3 + alpha4, alph3m4 = alpha3 - alpha4;
  double alph5p6 = alpha5 + alpha6, alph5m6 = alpha5 - alpha6;
  double alph7p8 = alpha7 + alpha8, alph7m8 = alpha7 - alpha8;
  double DN = two * std::sqrt(d_PI * b.Rho);
  const double aF = b.alphaF, aS = b.alphaS;
  double HRho = b.Rho * (aF * alph5p6 + aS * alph7p8);
  double HU = aS * b.as * alph5m6 - aF * b.af * alph7m8;
  double HB = b.c * (aS * alph5p6 - aF * alph7p8);
  d[0] = alpha1 + HRho;
  d[b.n] = aF * b.af * alph5m6 + aS * b.as * alph7m8;
  d[b.t1] = -b.betaZ * alph3p4 - betaYX * HU;
  d[b.t2] = b.betaY * alph3p4 - betaZX * HU;
  d[4] = b.cc * HRho;
  d[b.bn] = alpha2;
  d[b.bt1] = DN * (betaZX * alph3m4 + b.betaY * HB);
  d[b.bt2] = DN * (-betaYX * alph3m4 + b.betaZ * HB);
}

// minmod3 (src/reconstructionF.ChF:16-26)
inline double minmod3(double A, double B, double C) {
  double signA = fsign1(A);
  return signA * std::max(zero, std::min(std::min(std::fabs(A), signA * B), signA * C));
}

// Variables whose face values Reconstruction::checkTVDCondition clips
// (slopes::setTVDCheck from mhdam.check_tvd, main/MHDAM.cpp:942-966: 1
// flags RHO..PRES and BX..BZ and falls through to 2, which flags RHO and
// PRES).
bool tvd_checked(const mfm_params* P, int v) {
  if (P->check_tvd == 1) return v < 8;
  if (P->check_tvd == 2) return v == WRHO || v == WPRES;
  return false;
}

// Reconstruction::faceValues, characteristic branch
// (src/Reconstruction.cpp:178-204): INITIALSLOPES for every component,
// EqSysMHDMF::charAnalysis on the plasma components 0..7
// (srcIdeal/EqSysMHDMF.cpp:309-322), each component's faceSlopes, then
// charSynthesis + FROMSLOPESTOVALUES (srcIdeal/EqSysMHDMF.cpp:325-346),
// then checkTVDCondition (src/Reconstruction.cpp:236-250) on the variables
// mhdam.check_tvd flags.
void char_face_values(const Eq& E, const Fab& W, Fab& WM, Fab& WP, int dir,
                      const Box& b) {
  int io = dir == 0, jo = dir == 1, ko = dir == 2;
  PFORBOX(b, i, j, k) {
    double w[32], dL[32], dR[32], SL[32], SR[32];
    for (int v = 0; v < E.nc; ++v) {  // INITIALSLOPES
      double WPc = W(i, j, k, v);
      w[v] = WPc;
      dR[v] = W(i + io, j + jo, k + ko, v) - WPc;
      dL[v] = WPc - W(i - io, j - jo, k - ko, v);
    }
    CharBasis ba = char_basis(E, w, dir, 1.0e-12);
    double cL[8], cR[8];
    char_analysis(ba, dL, cL);
    char_analysis(ba, dR, cR);
    for (int v = 0; v < 8; ++v) { dL[v] = cL[v]; dR[v] = cR[v]; }
    for (int v = 0; v < E.nc; ++v) face_slopes(E.limiter(v), dL[v], dR[v], &SL[v], &SR[v]);
    CharBasis bs = char_basis(E, w, dir, 1.0e-8);
    double pL[8], pR[8];
    char_synthesis(bs, SL, pL);
    char_synthesis(bs, SR, pR);
    for (int v = 0; v < 8; ++v) { SL[v] = pL[v]; SR[v] = pR[v]; }
    for (int v = 0; v < E.nc; ++v) {  // FROMSLOPESTOVALUES
      WM(i, j, k, v) = w[v] - SL[v];
      WP(i, j, k, v) = w[v] + SR[v];
    }
    for (int v = 0; v < E.nc; ++v) {  // checkTVDCondition -> CHECKTVD :1248-1278
      if (!tvd_checked(E.P, v)) continue;
      double WPc = w[v];
      double dWR = W(i + io, j + jo, k + ko, v) - WPc;
      double dWL = WPc - W(i - io, j - jo, k - ko, v);
      double dSL = WPc - WM(i, j, k, v);
      double dSR = WP(i, j, k, v) - WPc;
      WM(i, j, k, v) = WPc - minmod3(dWL, dWR, dSL);
      WP(i, j, k, v) = WPc + minmod3(dWL, dWR, dSR);
    }
  }
}

// Reconstruction::faceValues (src/Reconstruction.cpp:131-206), per variable
void face_values(const Eq& E, const Fab& W, Fab& WM, Fab& WP, int dir,
                 const Box& b) {
  int io = dir == 0, jo = dir == 1, ko = dir == 2;
  if (E.P->char_reconstruction) {
    char_face_values(E, W, WM, WP, dir, b);
    return;
  }
  for (int v = 0; v < E.nc; ++v) {
    int lim = E.limiter(v);
    PFORBOX(b, i, j, k) {
      double l, r;
      limit(lim, W(i - io, j - jo, k - ko, v), W(i, j, k, v),
            W(i + io, j + jo, k + ko, v), &l, &r);
      WM(i, j, k, v) = l;
      WP(i, j, k, v) = r;
    }
  }
}

// EulerStep (src/PatchMHDAM.cpp:495-709)
void euler_step(const Eq& E, Fab& U, Fab* Fout, Fab* divBout, double dt,
                double dx, double time, const Box& box) {
  const int dims = E.dims;
  Box UBox = U.b;
  Box WBoxB = E.clip(UBox);
  Fab W;
  W.define(UBox, E.nc, std::numeric_limits<double>::quiet_NaN());
  cons_to_prim(E, W, U, WBoxB);
  for (int d = 0; d < dims; ++d) fill_ghost(E, W, d, dx);

  Box valueBox = E.clip(box);
  Box slopeBox = grow(valueBox, 1, dims);

  // preprocessing -> defineRegions (srcIdeal/PatchMHDMF.cpp:1447-1458)
  IFab REG;
  bool ce = (E.nf >= 2 && (E.P->problem == MFM_PROBLEM_HELIO ||
                           E.P->problem == MFM_PROBLEM_SWLISM));
  if (ce) {
    REG.define(slopeBox);
    FORBOX(slopeBox, i, j, k) {
      double w[32];
      for (int c = 0; c < E.nc; ++c) w[c] = W(i, j, k, c);
      REG(i, j, k) = region_cell(E, w, i, j, k, dx);
    }
  }

  Fab WM[3], WP[3], Bn[3];
  for (int d = 0; d < dims; ++d) {
    WM[d].define(slopeBox, E.nc, 0.0);
    WP[d].define(slopeBox, E.nc, 0.0);
    face_values(E, W, WM[d], WP[d], d, slopeBox);
    // checkPositivity: main/MHDAM.cpp:978-987 only queries the flag, so the
    // check is never enabled in the reference; nothing to do.
  }

  Box faceBox[3], fluxBox[3];
  for (int d = 0; d < dims; ++d) {
    Box fb = valueBox;
    fb = grow_dir(fb, d, 1);
    fb = E.clip(fb);
    fb = grow_dir(fb, d, -1);
    fb.hi[d] += 1;  // surroundingNodes
    faceBox[d] = fb;
    Box xb = valueBox;
    xb.hi[d] += 1;
    fluxBox[d] = xb;
    Fout[d].define(fluxBox[d], E.nc, 0.0);
    Bn[d].define(fluxBox[d], 1, 0.0);
  }

  for (int d = 0; d < dims; ++d) {
    Fab& F = Fout[d];
    int io = d == 0, jo = d == 1, ko = d == 2;
    // computeBn (srcIdeal/PatchMHDMF.cpp:509-621), method 1 = COMPUTEBN
    int meth = E.P->divb_method;
    if (meth == 1) {
      FORBOX(faceBox[d], i, j, k) {
        Bn[d](i, j, k, 0) = half * (WP[d](i - io, j - jo, k - ko, WBX + d) +
                                    WM[d](i, j, k, WBX + d));
      }
      if (E.physical(d)) {
        Box bb = fluxBox[d];
        if (valueBox.lo[d] == E.P->domain_lo[d]) {
          Box s = bb; s.hi[d] = s.lo[d];
          FORBOX(s, i, j, k) Bn[d](i, j, k, 0) = WM[d](i, j, k, WBX + d);
        }
        if (valueBox.hi[d] == E.P->domain_hi[d]) {
          Box s = bb; s.lo[d] = s.hi[d];
          FORBOX(s, i, j, k) Bn[d](i, j, k, 0) = WP[d](i - io, j - jo, k - ko, WBX + d);
        }
      }
    } else if (meth == 2) {  // COMPUTEBNCD :404-426
      FORBOX(faceBox[d], i, j, k) {
        Bn[d](i, j, k, 0) = half * (W(i - io, j - jo, k - ko, WBX + d) + W(i, j, k, WBX + d));
      }
    }
    // F.copy(Bn) into comp WBX+d (src/PatchMHDAM.cpp:609-610), then
    // fluxesRP overwrites it.
    FORBOX(faceBox[d], i, j, k) F(i, j, k, WBX + d) = Bn[d](i, j, k, 0);
    // fluxesRP (srcIdeal/PatchMHDMF.cpp:408-440)
    PFORBOX(faceBox[d], i, j, k) {
      double wl[32], wr[32], out[32];
      for (int c = 0; c < E.nc; ++c) {
        wl[c] = WP[d](i - io, j - jo, k - ko, c);
        wr[c] = WM[d](i, j, k, c);
        out[c] = F(i, j, k, c);
      }
      plasma_flux_face(E, wl, wr, d, out);
      for (int f = 1; f < E.nf; ++f)
        neutral_flux_face(E, wl, wr, d, E.iRhoN + (f - 1) * WNUM_E, out);
      for (int c = 0; c < E.nc; ++c) F(i, j, k, c) = out[c];
    }
    // fluxBC (problem), Lo then Hi
    if (E.physical(d)) {
      for (int side = 0; side < 2; ++side) {
        bool touches = side == 0 ? valueBox.lo[d] == E.P->domain_lo[d]
                                 : valueBox.hi[d] == E.P->domain_hi[d];
        if (!touches) continue;
        Box s = fluxBox[d];
        if (side == 0) s.hi[d] = s.lo[d]; else s.lo[d] = s.hi[d];
        FORBOX(s, i, j, k) {
          double wl[32], wr[32], out[32], bn;
          for (int c = 0; c < E.nc; ++c) {
            wl[c] = WP[d](i - io, j - jo, k - ko, c);
            wr[c] = WM[d](i, j, k, c);
            out[c] = F(i, j, k, c);
          }
          bn = Bn[d](i, j, k, 0);
          if (E.P->problem == MFM_PROBLEM_RIEMANN) {
            riemann_flux_bc_face(E, side, out, &bn);
          } else if (E.P->problem == MFM_PROBLEM_HELIO) {
            // m_RS->fluxes + m_RSGD->fluxes first (overwritten below)
            plasma_flux_face(E, wl, wr, d, out);
            for (int f = 1; f < E.nf; ++f)
              neutral_flux_face(E, wl, wr, d, E.iRhoN + (f - 1) * WNUM_E, out);
            helio_bc_face(E, side == 0 ? -1 : 1, d, dx, i, j, k, wl, wr, out, &bn);
          } else if (E.P->problem == MFM_PROBLEM_SWLISM) {
            // m_RS->fluxes + m_RSGD->fluxes, then SWLISMBC
            plasma_flux_face(E, wl, wr, d, out);
            for (int f = 1; f < E.nf; ++f)
              neutral_flux_face(E, wl, wr, d, E.iRhoN + (f - 1) * WNUM_E, out);
            swlism_bc_face(E, side == 0 ? -1 : 1, d, wl, wr, out, &bn);
          }
          for (int c = 0; c < E.nc; ++c) F(i, j, k, c) = out[c];
          Bn[d](i, j, k, 0) = bn;
        }
      }
    }
  }

  // finalUpdate -> UPDATECONS (src/PatchMHDAMF.ChF:542-588), dir order 0,1,2
  double invvol = one / dx;
  for (int d = 0; d < dims; ++d) {
    int io = d == 0, jo = d == 1, ko = d == 2;
    Fab& F = Fout[d];
    for (int c = 0; c < E.nc; ++c) {
      PFORBOX(valueBox, i, j, k) {
        U(i, j, k, c) = U(i, j, k, c) -
                        dt * invvol * (F(i + io, j + jo, k + ko, c) - F(i, j, k, c));
      }
    }
  }

  // addExplicitSources (srcIdeal/PatchMHDMF.cpp:1149-1389)
  Fab divB;
  divB.define(valueBox, 1, 0.0);
  for (int d = 0; d < dims; ++d) {  // COMPUTEDIVB :472-492
    int io = d == 0, jo = d == 1, ko = d == 2;
    FORBOX(valueBox, i, j, k) {
      divB(i, j, k, 0) = divB(i, j, k, 0) + Bn[d](i + io, j + jo, k + ko, 0) -
                         Bn[d](i, j, k, 0);
    }
  }
  const bool axisym = E.P->coord_sys == MFM_COORD_AXISYMMETRIC;
  if (axisym) {  // AXISYMMETRICDBZBYDZ (srcIdeal/PatchIdealMHDF.ChF:661-679)
    FORBOX(valueBox, i, j, k) {
      double y = dx * (j + half);
      double By = W(i, j, k, WBY);
      divB(i, j, k, 0) = divB(i, j, k, 0) + dx * (By / y);
    }
  }
  if (E.P->eight_wave) {  // SOURCE8WAVES :681-722 + ADDSOURCES UMOMX..UBZ
    double dtbydx = dt / dx;
    PFORBOX(valueBox, i, j, k) {
      double dB = divB(i, j, k, 0);
      dB = -dB * dtbydx;
      double Ux = W(i, j, k, WVELX), Uy = W(i, j, k, WVELY), Uz = W(i, j, k, WVELZ);
      double Bx = d_1_4PI * W(i, j, k, WBX), By = d_1_4PI * W(i, j, k, WBY),
             Bz = d_1_4PI * W(i, j, k, WBZ);
      double S[8];
      S[UMOMX] = dB * Bx; S[UMOMY] = dB * By; S[UMOMZ] = dB * Bz;
      S[UENG] = dB * (Bx * Ux + By * Uy + Bz * Uz);
      S[UBX] = dB * Ux; S[UBY] = dB * Uy; S[UBZ] = dB * Uz;
      for (int c = UMOMX; c <= UBZ; ++c) U(i, j, k, c) = U(i, j, k, c) + S[c];
    }
    if (divBout && divBout->p) {
      FORBOX(valueBox, i, j, k) {
        if (i >= divBout->b.lo[0] && i <= divBout->b.hi[0] &&
            j >= divBout->b.lo[1] && j <= divBout->b.hi[1] &&
            k >= divBout->b.lo[2] && k <= divBout->b.hi[2])
          (*divBout)(i, j, k, 0) = divB(i, j, k, 0);
      }
    }
  }
  if (axisym) {
    // geometric sources (srcIdeal/PatchMHDMF.cpp:1237-1276) + ADDSOURCES
    // from URHO to UBZ (one fluid) or the last neutral's UENG
    PFORBOX(valueBox, i, j, k) {
      double w[32], sa[32];
      for (int c = 0; c < E.nc; ++c) w[c] = W(i, j, k, c);
      axisym_source_cell(E, w, j, dt, dx, sa);
      for (int c = 0; c < E.nc; ++c) U(i, j, k, c) = U(i, j, k, c) + sa[c];
    }
  }
  if (E.dedner) {  // PSIPARABOLIC (src/DednerF.ChF:65-87)
    double coeff = std::exp(-dt * E.ch * E.ch / (E.cp * E.cp));
    FORBOX(valueBox, i, j, k) U(i, j, k, E.iPsi) = coeff * U(i, j, k, E.iPsi);
  }
  // m_PhPr->explicitSource: charge exchange (swlism/HeliosphericProblem.cpp:
  // 1394-1539); ADDSOURCES from UMOMX to the last neutral's UENG
  if (ce) {
    PFORBOX(valueBox, i, j, k) {
      double w[32], s[32];
      for (int c = 0; c < E.nc; ++c) w[c] = W(i, j, k, c);
      charge_exchange_cell(E, w, REG(i, j, k), dt, s);
      for (int c = UMOMX; c < E.nc; ++c) U(i, j, k, c) = U(i, j, k, c) + s[c];
    }
  }
  // problem postprocessing (HeliosphericProblem::postprocessing :2399-2476)
  if (E.P->problem == MFM_PROBLEM_HELIO) helio_reinit(E, U, dx, dt + time, valueBox);
  if (E.P->problem == MFM_PROBLEM_SWLISM) swlism_reinit(E, U, dx, valueBox);
  // PatchMHDMF::postprocessing (srcIdeal/PatchMHDMF.cpp:1392-1440)
  PFORBOX(valueBox, i, j, k) {
    double u[32];
    for (int c = 0; c < E.nc; ++c) u[c] = U(i, j, k, c);
    floors_cell(E, u);
    for (int c = 0; c < E.nc; ++c) U(i, j, k, c) = u[c];
  }
}

// FLUXESHANCOCK (srcIdeal/PatchIdealMHDF.ChF:235-297): physical MHD flux of
// one primitive state
void phys_flux_mhd(const Eq& E, const double* w, int idir, double* out) {
  Rot r = rot(0, idir);
  double rho = std::max(E.smallr, w[WRHO]);
  double u = w[r.n], v = w[r.t1], ww = w[r.t2];
  double pp = std::max(E.smallp, w[WPRES]);
  double bx = w[r.bn], by = w[r.bt1], bz = w[r.bt2];
  double b = d_1_8PI * (bx * bx + by * by + bz * bz);
  double ke = half * (u * u + v * v + ww * ww);
  double bu = bx * u + by * v + bz * ww;
  double e = pp * E.invhgamma + rho * ke;
  pp = pp + b;
  e = e + b;
  out[URHO] = rho * u;
  out[UMOMX + idir] = rho * u * u + pp - d_1_4PI * bx * bx;
  out[UMOMX + (idir + 1) % 3] = rho * u * v - d_1_4PI * bx * by;
  out[UMOMX + (idir + 2) % 3] = rho * u * ww - d_1_4PI * bx * bz;
  out[UENG] = u * (e + pp) - d_1_4PI * bx * bu;
  out[UBX + idir] = zero;
  out[UBX + (idir + 1) % 3] = u * by - v * bx;
  out[UBX + (idir + 2) % 3] = u * bz - ww * bx;
}

// FLUXESHANCOCK_E (srcIdeal/PatchEulerF.ChF:417-468): physical Euler flux
// of the fluid at iRho (conserved) / iRhoP (primitive)
void phys_flux_euler(const Eq& E, const double* w, int idir, int iRho, int iRhoP,
                     double* out) {
  const int iMOMX = iRho + 1, iENG = iRho + 4, iVELX = iRhoP + 1, iPRES = iRhoP + 4;
  double rho = std::max(E.smallr, w[iRhoP]);
  double u = w[iVELX + idir], v = w[iVELX + (idir + 1) % 3], ww = w[iVELX + (idir + 2) % 3];
  double p = std::max(E.smallp, w[iPRES]);
  double ke = half * (u * u + v * v + ww * ww);
  double e = p * E.invhgamma + rho * ke;
  out[iRho] = rho * u;
  out[iMOMX + idir] = rho * u * u + p;
  out[iMOMX + (idir + 1) % 3] = rho * u * v;
  out[iMOMX + (idir + 2) % 3] = rho * u * ww;
  out[iENG] = u * (e + p);
}

// PatchMHDMF::computeBn method 1 (srcIdeal/PatchMHDMF.cpp:509-621) on the
// faces of `faces`, then the physical-boundary faces of the Bn FAB box
// `bnbox` take the face value of the cell inside; method 2 COMPUTEBNCD.
void compute_bn(const Eq& E, const Fab& W, const Fab& WM, const Fab& WP, Fab& Bn, int d,
                const Box& faces, const Box& bnbox) {
  int io = d == 0, jo = d == 1, ko = d == 2;
  int meth = E.P->divb_method;
  if (meth == 1) {
    FORBOX(faces, i, j, k)
      Bn(i, j, k, 0) = half * (WP(i - io, j - jo, k - ko, WBX + d) + WM(i, j, k, WBX + d));
    if (E.physical(d)) {
      if (bnbox.lo[d] == E.P->domain_lo[d]) {
        Box s = bnbox; s.hi[d] = s.lo[d];
        FORBOX(s, i, j, k) Bn(i, j, k, 0) = WM(i, j, k, WBX + d);
      }
      if (bnbox.hi[d] == E.P->domain_hi[d] + 1) {
        Box s = bnbox; s.lo[d] = s.hi[d];
        FORBOX(s, i, j, k) Bn(i, j, k, 0) = WP(i - io, j - jo, k - ko, WBX + d);
      }
    }
  } else if (meth == 2) {
    FORBOX(faces, i, j, k)
      Bn(i, j, k, 0) = half * (W(i - io, j - jo, k - ko, WBX + d) + W(i, j, k, WBX + d));
  }
}

// addExplicitSources (srcIdeal/PatchMHDMF.cpp:1149-1389) on valueBox with
// the primitive state W, face Bn and regions REG (8-wave, PSIPARABOLIC,
// charge exchange), as in euler_step.
void explicit_sources(const Eq& E, Fab& U, const Fab& W, const Fab* Bn, IFab& REG,
                      bool ce, double dt, double dx, const Box& valueBox, Fab* divBout) {
  const int dims = E.dims;
  Fab divB;
  divB.define(valueBox, 1, 0.0);
  for (int d = 0; d < dims; ++d) {  // COMPUTEDIVB :472-492
    int io = d == 0, jo = d == 1, ko = d == 2;
    FORBOX(valueBox, i, j, k) {
      divB(i, j, k, 0) = divB(i, j, k, 0) + Bn[d](i + io, j + jo, k + ko, 0) -
                         Bn[d](i, j, k, 0);
    }
  }
  if (E.P->eight_wave) {  // SOURCE8WAVES :681-722 + ADDSOURCES UMOMX..UBZ
    double dtbydx = dt / dx;
    PFORBOX(valueBox, i, j, k) {
      double dB = divB(i, j, k, 0);
      dB = -dB * dtbydx;
      double Ux = W(i, j, k, WVELX), Uy = W(i, j, k, WVELY), Uz = W(i, j, k, WVELZ);
      double Bx = d_1_4PI * W(i, j, k, WBX), By = d_1_4PI * W(i, j, k, WBY),
             Bz = d_1_4PI * W(i, j, k, WBZ);
      double S[8];
      S[UMOMX] = dB * Bx; S[UMOMY] = dB * By; S[UMOMZ] = dB * Bz;
      S[UENG] = dB * (Bx * Ux + By * Uy + Bz * Uz);
      S[UBX] = dB * Ux; S[UBY] = dB * Uy; S[UBZ] = dB * Uz;
      for (int c = UMOMX; c <= UBZ; ++c) U(i, j, k, c) = U(i, j, k, c) + S[c];
    }
    if (divBout && divBout->p) {
      FORBOX(valueBox, i, j, k) {
        if (i >= divBout->b.lo[0] && i <= divBout->b.hi[0] &&
            j >= divBout->b.lo[1] && j <= divBout->b.hi[1] &&
            k >= divBout->b.lo[2] && k <= divBout->b.hi[2])
          (*divBout)(i, j, k, 0) = divB(i, j, k, 0);
      }
    }
  }
  if (E.dedner) {  // PSIPARABOLIC (src/DednerF.ChF:65-87)
    double coeff = std::exp(-dt * E.ch * E.ch / (E.cp * E.cp));
    FORBOX(valueBox, i, j, k) U(i, j, k, E.iPsi) = coeff * U(i, j, k, E.iPsi);
  }
  if (ce) {
    PFORBOX(valueBox, i, j, k) {
      double w[32], s[32];
      for (int c = 0; c < E.nc; ++c) w[c] = W(i, j, k, c);
      charge_exchange_cell(E, w, REG(i, j, k), dt, s);
      for (int c = UMOMX; c < E.nc; ++c) U(i, j, k, c) = U(i, j, k, c) + s[c];
    }
  }
}

// updateState with time_approximation = 1: predictorHancock on a copy of U
// with dt/2, then correctorHancock on U with dt (src/PatchMHDAM.cpp:1288-1300,
// 858-1189).  The predictor is cell-local in time: physical fluxes of the
// cell's own face values (FLUXESHANCOCK), no Riemann problem; its state moves
// the face values by W(U_pred) - W(U) (UPDATEVALUESONFACES), and the
// corrector solves the Riemann problems on those.
void hancock_step(const Eq& E, Fab& U, Fab* Fout, Fab* divBout, double dt, double dx,
                  double time, const Box& box) {
  const int dims = E.dims;
  const Box UBox = U.b;
  const Box WBoxB = E.clip(UBox);
  const bool ce = (E.nf >= 2 && E.P->problem == MFM_PROBLEM_HELIO);
  const double invvol = one / dx;
  // ---------------------------------------------------------- predictor --
  const double dtp = 0.5 * dt;
  Fab UP;
  UP.define(UBox, E.nc, 0.0);
  std::memcpy(UP.p, U.p, sizeof(double) * (size_t)U.sc * E.nc);
  Fab W;
  W.define(UBox, E.nc, std::numeric_limits<double>::quiet_NaN());
  cons_to_prim(E, W, UP, WBoxB);
  for (int d = 0; d < dims; ++d) fill_ghost(E, W, d, dx);
  const Box valueP = E.clip(grow(box, 1, dims));
  const Box slopeBox = grow(valueP, 1, dims);
  IFab REG;
  if (ce) {
    REG.define(slopeBox);
    FORBOX(slopeBox, i, j, k) {
      double w[32];
      for (int c = 0; c < E.nc; ++c) w[c] = W(i, j, k, c);
      REG(i, j, k) = region_cell(E, w, i, j, k, dx);
    }
  }
  Fab WM[3], WP[3], Bn[3], FM[3], FP[3];
  Box fluxP[3];
  for (int d = 0; d < dims; ++d) {
    WM[d].define(slopeBox, E.nc, 0.0);
    WP[d].define(slopeBox, E.nc, 0.0);
    face_values(E, W, WM[d], WP[d], d, slopeBox);
    FM[d].define(valueP, E.nc, 0.0);
    FP[d].define(valueP, E.nc, 0.0);
    PFORBOX(valueP, i, j, k) {  // fluxesHancock (srcIdeal/PatchMHDMF.cpp:307-406)
      double wm[32], wp[32], fm[32], fp[32];
      for (int c = 0; c < E.nc; ++c) { wm[c] = WM[d](i, j, k, c); wp[c] = WP[d](i, j, k, c); }
      phys_flux_mhd(E, wm, d, fm);
      phys_flux_mhd(E, wp, d, fp);
      for (int f = 1; f < E.nf; ++f) {
        const int s = E.iRhoN + (f - 1) * WNUM_E;
        phys_flux_euler(E, wm, d, s, s, fm);
        phys_flux_euler(E, wp, d, s, s, fp);
      }
      for (int c = 0; c < E.nc; ++c) { FM[d](i, j, k, c) = fm[c]; FP[d](i, j, k, c) = fp[c]; }
    }
    fluxP[d] = valueP;
    fluxP[d].hi[d] += 1;
    Bn[d].define(fluxP[d], 1, 0.0);
    compute_bn(E, W, WM[d], WP[d], Bn[d], d, fluxP[d], fluxP[d]);
  }
  for (int d = 0; d < dims; ++d) {  // updateHancock -> UPDATECONSHANCOCK (areas = 1)
    for (int c = 0; c < E.nc; ++c) {
      PFORBOX(valueP, i, j, k) {
        UP(i, j, k, c) = UP(i, j, k, c) -
                         dtp * invvol * (one * FP[d](i, j, k, c) - one * FM[d](i, j, k, c));
      }
    }
  }
  explicit_sources(E, UP, W, Bn, REG, ce, dtp, dx, valueP, nullptr);
  Fab WPred;
  WPred.define(UBox, E.nc, std::numeric_limits<double>::quiet_NaN());
  cons_to_prim(E, WPred, UP, WBoxB);
  for (int d = 0; d < dims; ++d) fill_ghost(E, WPred, d, dx);
  for (int d = 0; d < dims; ++d) {  // UPDATEVALUESONFACES (src/PatchMHDAMF.ChF:640-661)
    for (int c = 0; c < E.nc; ++c) {
      FORBOX(slopeBox, i, j, k) {
        double aux = WPred(i, j, k, c) - W(i, j, k, c);
        WM[d](i, j, k, c) = WM[d](i, j, k, c) + aux;
        WP[d](i, j, k, c) = WP[d](i, j, k, c) + aux;
      }
    }
    compute_bn(E, W, WM[d], WP[d], Bn[d], d, fluxP[d], fluxP[d]);
  }
  // ---------------------------------------------------------- corrector --
  const Box valueBox = E.clip(box);
  IFab REGc;
  if (ce) {
    REGc.define(valueBox);
    FORBOX(valueBox, i, j, k) {
      double w[32];
      for (int c = 0; c < E.nc; ++c) w[c] = WPred(i, j, k, c);
      REGc(i, j, k) = region_cell(E, w, i, j, k, dx);
    }
  }
  for (int d = 0; d < dims; ++d) {
    int io = d == 0, jo = d == 1, ko = d == 2;
    Box faceBox = grow_dir(valueBox, d, 1);
    faceBox = E.clip(faceBox);
    faceBox = grow_dir(faceBox, d, -1);
    faceBox.hi[d] += 1;
    Box fluxBox = valueBox;
    fluxBox.hi[d] += 1;
    Fout[d].define(fluxBox, E.nc, 0.0);
    Fab& F = Fout[d];
    FORBOX(fluxBox, i, j, k) F(i, j, k, WBX + d) = Bn[d](i, j, k, 0);
    PFORBOX(faceBox, i, j, k) {
      double wl[32], wr[32], out[32];
      for (int c = 0; c < E.nc; ++c) {
        wl[c] = WP[d](i - io, j - jo, k - ko, c);
        wr[c] = WM[d](i, j, k, c);
        out[c] = F(i, j, k, c);
      }
      plasma_flux_face(E, wl, wr, d, out);
      for (int f = 1; f < E.nf; ++f)
        neutral_flux_face(E, wl, wr, d, E.iRhoN + (f - 1) * WNUM_E, out);
      for (int c = 0; c < E.nc; ++c) F(i, j, k, c) = out[c];
    }
    if (E.physical(d)) {
      for (int side = 0; side < 2; ++side) {
        bool touches = side == 0 ? valueBox.lo[d] == E.P->domain_lo[d]
                                 : valueBox.hi[d] == E.P->domain_hi[d];
        if (!touches) continue;
        Box sb = fluxBox;
        if (side == 0) sb.hi[d] = sb.lo[d]; else sb.lo[d] = sb.hi[d];
        FORBOX(sb, i, j, k) {
          double wl[32], wr[32], out[32], bn;
          for (int c = 0; c < E.nc; ++c) {
            wl[c] = WP[d](i - io, j - jo, k - ko, c);
            wr[c] = WM[d](i, j, k, c);
            out[c] = F(i, j, k, c);
          }
          bn = Bn[d](i, j, k, 0);
          if (E.P->problem == MFM_PROBLEM_RIEMANN) {
            riemann_flux_bc_face(E, side, out, &bn);
          } else if (E.P->problem == MFM_PROBLEM_HELIO) {
            plasma_flux_face(E, wl, wr, d, out);
            for (int f = 1; f < E.nf; ++f)
              neutral_flux_face(E, wl, wr, d, E.iRhoN + (f - 1) * WNUM_E, out);
            helio_bc_face(E, side == 0 ? -1 : 1, d, dx, i, j, k, wl, wr, out, &bn);
          }
          for (int c = 0; c < E.nc; ++c) F(i, j, k, c) = out[c];
          Bn[d](i, j, k, 0) = bn;
        }
      }
    }
  }
  for (int d = 0; d < dims; ++d) {  // finalUpdate -> UPDATECONS
    int io = d == 0, jo = d == 1, ko = d == 2;
    Fab& F = Fout[d];
    for (int c = 0; c < E.nc; ++c) {
      PFORBOX(valueBox, i, j, k) {
        U(i, j, k, c) = U(i, j, k, c) -
                        dt * invvol * (F(i + io, j + jo, k + ko, c) - F(i, j, k, c));
      }
    }
  }
  explicit_sources(E, U, WPred, Bn, REGc, ce, dt, dx, valueBox, divBout);
  if (E.P->problem == MFM_PROBLEM_HELIO) helio_reinit(E, U, dx, dt + time, valueBox);
  PFORBOX(valueBox, i, j, k) {
    double u[32];
    for (int c = 0; c < E.nc; ++c) u[c] = U(i, j, k, c);
    floors_cell(E, u);
    for (int c = 0; c < E.nc; ++c) U(i, j, k, c) = u[c];
  }
}

// PatchMHDMF::computeDt + PhysProblem/HeliosphericProblem::computeDt
// (swlism/HeliosphericProblem.cpp:3809-3860); the exclusion radius is
// m_R0dtIgnore = m_R0 - 2 dx of the level (:497).
double compute_dt(const Eq& E, const Fab& U, const Box& box, double dx,
                  int cell[3]) {
  double dtMin = std::numeric_limits<double>::max();
  bool first = true;
  // the inner-boundary exclusion of HeliosphericProblem::computeDt (R0 - 2dx
  // of the level) and SWLISMProblem::computeDt (m_R0dtIgnore = R0 - dx of
  // the coarsest level, swlism/SWLISMProblem.cpp:153, 1539-1582)
  double ign = E.P->problem == MFM_PROBLEM_HELIO    ? E.P->sun_r0 - 2.0 * dx
               : E.P->problem == MFM_PROBLEM_SWLISM ? E.P->r0_dt_ignore
                                                    : -1.0;
  double ign2 = ign * ign;
  cell[0] = box.lo[0]; cell[1] = box.lo[1]; cell[2] = box.lo[2];
  FORBOX(box, i, j, k) {
    double u[32];
    for (int c = 0; c < E.nc; ++c) u[c] = U(i, j, k, c);
    double cmax = max_wave_speed_cell(E, u);
    double dt = dx / cmax;  // FArrayBox::invert(dx)
    if (ign > 0.0) {
      double x = (i + 0.5) * dx - E.P->sun_xc, y = (j + 0.5) * dx - E.P->sun_yc;
      double z = (E.dims == 3) ? (k + 0.5) * dx - E.P->sun_zc : 0.0;
      double dist2 = (E.dims == 3) ? x * x + y * y + z * z : x * x + y * y;
      if ((dist2 > ign2) && (dt < dtMin)) {
        dtMin = dt; cell[0] = i; cell[1] = j; cell[2] = k;
      }
    } else {
      if (first || dt < dtMin) {  // BaseFab::minIndex: first minimum
        dtMin = dt; cell[0] = i; cell[1] = j; cell[2] = k; first = false;
      }
    }
  }
  return dtMin;
}

int check_args(const mfm_params* p) {
  if (!p || p->abi_version != MFM_ABI_VERSION) return MFM_ERR_ARG;
  if (p->nfluids < 1 || p->nfluids > 4) return MFM_ERR_UNSUPPORTED;
  // the factory's MHD solvers (src/RiemannSolver.cpp:39-104) but VFRoe (5),
  // and the HLLC flux (21)
  if (!((p->flux >= 0 && p->flux <= 4) || p->flux == 11 || p->flux == 13 ||
        p->flux == MFM_FLUX_HLLC))
    return MFM_ERR_UNSUPPORTED;
  if (p->dedner && p->nfluids != 1 && p->problem != MFM_PROBLEM_NONE) return MFM_ERR_UNSUPPORTED;
  if (p->spacedim != 2 && p->spacedim != 3) return MFM_ERR_ARG;
  if (p->time_approx != MFM_TIME_RK2 && p->time_approx != MFM_TIME_FIRST_ORDER &&
      p->time_approx != MFM_TIME_HANCOCK)
    return MFM_ERR_UNSUPPORTED;
  // FLUXESHANCOCK has no psi flux: the GLM system is not run with Hancock
  if (p->time_approx == MFM_TIME_HANCOCK && p->dedner) return MFM_ERR_UNSUPPORTED;
  // coord_sys = 1 with the 2-D SW/LISM problem (RK2 or first order, no GLM)
  if (p->coord_sys != MFM_COORD_CARTESIAN &&
      !(p->coord_sys == MFM_COORD_AXISYMMETRIC && p->spacedim == 2 && !p->dedner &&
        p->time_approx != MFM_TIME_HANCOCK &&
        (p->problem == MFM_PROBLEM_SWLISM || p->problem == MFM_PROBLEM_NONE)))
    return MFM_ERR_UNSUPPORTED;
  if (p->problem == MFM_PROBLEM_SWLISM &&
      (p->spacedim != 2 || p->dedner || p->time_approx == MFM_TIME_HANCOCK))
    return MFM_ERR_UNSUPPORTED;
  if (p->check_tvd < 0 || p->check_tvd > 2) return MFM_ERR_ARG;
  if (p->char_reconstruction)  // no first-order (unlimited) plasma slopes
    for (int v = 0; v < 8; ++v)
      if (p->limiter[v] < MFM_LIM_CD || p->limiter[v] > MFM_LIM_WENO3YC)
        return MFM_ERR_UNSUPPORTED;
  return MFM_OK;
}

}  // namespace

// ============================================================ C entry ======
extern "C" {

int oracle_euler_step(const mfm_params* p, double ch, double cp, mfm_fab Uf,
                      mfm_box box, double dt, double dx, double time,
                      mfm_fab* F, mfm_fab* divB, int nthreads) {
  int rc = check_args(p);
  if (rc) return rc;
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#endif
  Eq E(p, ch, cp);
  Fab U; U.wrap(Uf);
  Fab Fo[3];
  Fab dB; if (divB) dB.wrap(*divB);
  euler_step(E, U, Fo, divB ? &dB : nullptr, dt, dx, time, mkbox(box));
  if (F) {
    for (int d = 0; d < E.dims; ++d) {
      Fab Fd; Fd.wrap(F[d]);
      FORBOX(Fd.b, i, j, k)
        for (int c = 0; c < E.nc; ++c) Fd(i, j, k, c) = Fo[d](i, j, k, c);
    }
  }
  return MFM_OK;
}

int oracle_patch_update(const mfm_params* p, double ch, double cp, mfm_fab Uf,
                        mfm_box validb, double dt, double dx, double time,
                        mfm_fab* F, mfm_fab* divB, double* new_dt,
                        int min_cell[3], int nthreads) {
  int rc = check_args(p);
  if (rc) return rc;
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#endif
  Eq E(p, ch, cp);
  Fab U; U.wrap(Uf);
  Box box = mkbox(validb);
  // PatchMHDAM::numGhostCells (src/PatchMHDAM.cpp:280-286)
  int ng = p->time_approx == MFM_TIME_RK2 ? 4 : (p->time_approx == MFM_TIME_HANCOCK ? 3 : 2);
  if (!contains(U.b, grow(box, ng, E.dims)) || U.nc != E.nc) return MFM_ERR_SHAPE;
  Fab dB; if (divB) dB.wrap(*divB);
  Fab Fo[3];
  if (p->time_approx == MFM_TIME_RK2) {
    // src/PatchMHDAM.cpp:1239-1285
    Fab Uold; Uold.define(U.b, E.nc, 0.0);
    std::memcpy(Uold.p, U.p, sizeof(double) * (size_t)U.sc * E.nc);
    Fab Up; Up.define(U.b, E.nc, 0.0);
    std::memcpy(Up.p, U.p, sizeof(double) * (size_t)U.sc * E.nc);
    Fab Fp[3];
    euler_step(E, Up, Fp, divB ? &dB : nullptr, dt, dx, time, grow(box, 2, E.dims));
    std::memcpy(U.p, Up.p, sizeof(double) * (size_t)U.sc * E.nc);
    euler_step(E, U, Fo, divB ? &dB : nullptr, 0.5 * dt, dx, time, box);
    for (int d = 0; d < E.dims; ++d) {  // a_F += FPredictor; a_F *= 0.5
      FORBOX(Fo[d].b, i, j, k)
        for (int c = 0; c < E.nc; ++c)
          Fo[d](i, j, k, c) = (Fo[d](i, j, k, c) + Fp[d](i, j, k, c)) * 0.5;
    }
    // AVERAGESTATE (src/PatchMHDAMF.ChF:289-316) over all components
    for (int c = 0; c < E.nc; ++c) {
      PFORBOX(box, i, j, k) {
        U(i, j, k, c) = U(i, j, k, c) + half * (Uold(i, j, k, c) - Up(i, j, k, c));
      }
    }
    // ghost cells: the reference leaves U*, the library leaves U^n; restore
    // U^n outside the valid box so both contracts coincide.
    FORBOX(U.b, i, j, k) {
      bool in = i >= box.lo[0] && i <= box.hi[0] && j >= box.lo[1] &&
                j <= box.hi[1] && k >= box.lo[2] && k <= box.hi[2];
      if (!in)
        for (int c = 0; c < E.nc; ++c) U(i, j, k, c) = Uold(i, j, k, c);
    }
  } else if (p->time_approx == MFM_TIME_HANCOCK) {
    hancock_step(E, U, Fo, divB ? &dB : nullptr, dt, dx, time, box);
  } else {
    euler_step(E, U, Fo, divB ? &dB : nullptr, dt, dx, time, box);
  }
  if (F) {
    for (int d = 0; d < E.dims; ++d) {
      Fab Fd; Fd.wrap(F[d]);
      FORBOX(Fd.b, i, j, k)
        for (int c = 0; c < E.nc; ++c) Fd(i, j, k, c) = Fo[d](i, j, k, c);
    }
  }
  double dtn = compute_dt(E, U, box, dx, min_cell);
  if (new_dt) *new_dt = dtn;
  if (!(dtn > 0.0) || !std::isfinite(dtn)) return MFM_ERR_NUMERIC;
  return MFM_OK;
}

int oracle_compute_dt(const mfm_params* p, mfm_fab Uf, mfm_box box, double dx,
                      double* dt, int min_cell[3]) {
  int rc = check_args(p);
  if (rc) return rc;
  Eq E(p, 1.0, 1.0);
  Fab U; U.wrap(Uf);
  *dt = compute_dt(E, U, mkbox(box), dx, min_cell);
  return MFM_OK;
}

// HELIOINIT (swlism/HeliosphericF.ChF:1150-1361), ORSZAGINIT
// (problems/OrszagF.ChF:27-79), RIEMANNINIT (problems/RiemannF.ChF:65-127)
int oracle_init_problem(const mfm_params* p, mfm_fab Uf, mfm_box bx, double dx) {
  int rc = check_args(p);
  if (rc) return rc;
  Eq E(p, 1.0, 1.0);
  Fab U; U.wrap(Uf);
  Box box = mkbox(bx);
  const double hgamma = E.hgamma;
  const bool d3 = E.dims == 3;
  if (p->problem == MFM_PROBLEM_HELIO) {
    double sinTILT = std::sin(p->sun_tilt), cosTILT = std::cos(p->sun_tilt);
    double sunBeta = p->omega * eos_AU / (p->sun_v * p->lism_v_dim);
    FORBOX(box, i, j, k) {
      double x = (i + half) * dx - p->sun_xc, y = (j + half) * dx - p->sun_yc;
      double z = d3 ? (k + half) * dx - p->sun_zc : zero;
      double r = d3 ? x * x + y * y + z * z : x * x + y * y;
      r = std::sqrt(r);
      double xy = std::sqrt(x * x + y * y);
      double cosT = z / r, sinT = std::sqrt(one - cosT * cosT);
      double cosF = x / xy, sinF = y / xy;
      if (r > p->init_r) {
        double rr = p->init_r / r;
        rr = std::max(rr, 0.0);
        rr = rr * rr * rr;
        double vel = sinT * cosF * p->lism_ux + sinT * sinF * p->lism_uy + cosT * p->lism_uz;
        double ux = p->lism_ux - half * rr * (3.0 * sinT * cosF * vel - p->lism_ux);
        double uy = p->lism_uy - half * rr * (3.0 * sinT * sinF * vel - p->lism_uy);
        double uz = p->lism_uz - half * rr * (3.0 * cosT * vel - p->lism_uz);
        double pp = one / (E.gamma * p->lism_m * p->lism_m);
        double b = sinT * cosF * p->lism_bx + sinT * sinF * p->lism_by + cosT * p->lism_bz;
        double Bx = p->lism_bx - half * rr * (3.0 * sinT * cosF * b - p->lism_bx);
        double By = p->lism_by - half * rr * (3.0 * sinT * sinF * b - p->lism_by);
        double Bz = p->lism_bz - half * rr * (3.0 * cosT * b - p->lism_bz);
        b = d_1_4PI * (Bx * Bx + By * By + Bz * Bz);
        double ke = ux * ux + uy * uy + uz * uz;
        double e0 = pp / hgamma + half * (ke + b);
        U(i, j, k, URHO) = one;
        U(i, j, k, UMOMX) = ux; U(i, j, k, UMOMY) = uy; U(i, j, k, UMOMZ) = uz;
        U(i, j, k, UENG) = e0;
        U(i, j, k, UBX) = Bx; U(i, j, k, UBY) = By; U(i, j, k, UBZ) = Bz;
      } else {
        double rr = r / p->sun_int_bc_radius;
        double theta = std::acos(cosT);
        double cosBr = std::cos(sunBeta * r), sinBr = std::sin(sunBeta * r);
        double sinB = sinF * cosBr + cosF * sinBr;
        double cosB = cosF * cosBr - sinF * sinBr;
        double thetaS = sinTILT * sinB /
                        std::sqrt((cosTILT * cosB) * (cosTILT * cosB) + sinB * sinB);
        thetaS = std::asin(thetaS) + d_PI_2;
        double rho = p->sun_rho / (rr * rr);
        double ux = p->sun_v * sinT * cosF, uy = p->sun_v * sinT * sinF,
               uz = p->sun_v * cosT;
        double pp = p->sun_p / std::pow(rr, two * E.gamma);
        double br = p->dir_br_n * p->sun_b / (rr * rr);
        double bp = -sunBeta * br * r * sinT;
        double bxx = br * sinT * cosF - bp * sinF;
        double byy = br * sinT * sinF + bp * cosF;
        double bzz = br * cosT;
        double b = bxx * bxx + byy * byy + bzz * bzz;
        double e0 = pp / hgamma + half * (rho * p->sun_v * p->sun_v + d_1_4PI * b);
        U(i, j, k, URHO) = rho;
        U(i, j, k, UMOMX) = rho * ux; U(i, j, k, UMOMY) = rho * uy;
        U(i, j, k, UMOMZ) = rho * uz;
        U(i, j, k, UENG) = e0;
        double sg = (p->sun_b_monopolar == 1 || theta < thetaS) ? one : -one;
        U(i, j, k, UBX) = sg == one ? bxx : -bxx;
        U(i, j, k, UBY) = sg == one ? byy : -byy;
        U(i, j, k, UBZ) = sg == one ? bzz : -bzz;
      }
      if (E.nf > 1) {
        int s = E.iRhoN;
        U(i, j, k, s + URHO) = p->net_rho;
        U(i, j, k, s + UMOMX) = p->net_rho * p->lism_ux;
        U(i, j, k, s + UMOMY) = p->net_rho * p->lism_uy;
        U(i, j, k, s + UMOMZ) = p->net_rho * p->lism_uz;
        U(i, j, k, s + UENG) = p->net_p / hgamma + half * p->net_rho;
      }
      if (E.nf > 2) {
        int s = E.iRhoN + WNUM_E;
        U(i, j, k, s + URHO) = netFluid2Dens;
        U(i, j, k, s + UMOMX) = zero; U(i, j, k, s + UMOMY) = zero;
        U(i, j, k, s + UMOMZ) = zero;
        U(i, j, k, s + UENG) = netFluid2Pres / hgamma;
      }
      if (E.nf > 3) {
        int s = E.iRhoN + 2 * WNUM_E;
        U(i, j, k, s + URHO) = netFluid3Dens;
        U(i, j, k, s + UMOMX) = zero; U(i, j, k, s + UMOMY) = zero;
        U(i, j, k, s + UMOMZ) = zero;
        U(i, j, k, s + UENG) = netFluid3Pres / hgamma;
      }
    }
    return MFM_OK;
  }
  if (p->problem == MFM_PROBLEM_SWLISM) {
    // SWLISMINIT (swlism/SWLISMF.ChF:106-255): the LISM potential flow past
    // the sphere initR outside it, the radial solar wind inside
    const int nf = E.nf, iRhoN = E.iRhoN;
    FORBOX(box, i, j, k) {
      double x = (i + half) * dx - p->sun_xc, y = (j + half) * dx - p->sun_yc;
      double r = std::sqrt(x * x + y * y);
      if (r > p->init_r) {
        double sinT = y / r, cosT = x / r;
        double h = p->init_r / r;
        h = h * h * h;
        double ux = one - (one - 1.5 * sinT * sinT) * h;
        double uy = -1.5 * sinT * cosT * h;
        double uz = zero;
        double pr = one / (E.gamma * p->lism_m * p->lism_m);
        double Bx = p->lism_bx * ux, By = p->lism_bx * uy, Bz = zero;
        double b = d_1_4PI * (Bx * Bx + By * By + Bz * Bz);
        double ke = ux * ux + uy * uy + uz * uz;
        double e0 = pr / hgamma + half * (ke + b);
        U(i, j, k, URHO) = one;
        U(i, j, k, UMOMX) = ux; U(i, j, k, UMOMY) = uy; U(i, j, k, UMOMZ) = uz;
        U(i, j, k, UENG) = e0;
        U(i, j, k, UBX) = Bx; U(i, j, k, UBY) = By; U(i, j, k, UBZ) = Bz;
        if (nf > 1) {
          e0 = p->net_p / hgamma + half * p->net_rho * ke;
          U(i, j, k, iRhoN + URHO) = p->net_rho;
          U(i, j, k, iRhoN + UMOMX) = p->net_rho * ux;
          U(i, j, k, iRhoN + UMOMY) = p->net_rho * uy;
          U(i, j, k, iRhoN + UMOMZ) = p->net_rho * uz;
          U(i, j, k, iRhoN + UENG) = e0;
        }
      } else {
        double rho = p->sun_rho / (r * r);
        double ux = p->sun_v * x / r, uy = p->sun_v * y / r;
        double pr = p->sun_p / std::pow(r, two * E.gamma);
        double e0 = pr / hgamma + half * rho * p->sun_v * p->sun_v;
        U(i, j, k, URHO) = rho;
        U(i, j, k, UMOMX) = rho * ux; U(i, j, k, UMOMY) = rho * uy; U(i, j, k, UMOMZ) = zero;
        U(i, j, k, UENG) = e0;
        U(i, j, k, UBX) = zero; U(i, j, k, UBY) = zero; U(i, j, k, UBZ) = zero;
        if (nf > 1) {
          U(i, j, k, iRhoN + URHO) = p->net_rho;
          U(i, j, k, iRhoN + UMOMX) = p->net_rho * one;
          U(i, j, k, iRhoN + UMOMY) = zero;
          U(i, j, k, iRhoN + UMOMZ) = zero;
          U(i, j, k, iRhoN + UENG) = p->net_p / hgamma + half * p->net_rho;
        }
      }
      if (nf > 2) {
        int sh = iRhoN + WNUM_E;
        U(i, j, k, sh + URHO) = netFluid2Dens;
        U(i, j, k, sh + UMOMX) = zero; U(i, j, k, sh + UMOMY) = zero; U(i, j, k, sh + UMOMZ) = zero;
        U(i, j, k, sh + UENG) = netFluid2Pres / hgamma;
      }
      if (nf > 3) {
        int sh = iRhoN + 2 * WNUM_E;
        U(i, j, k, sh + URHO) = netFluid3Dens;
        U(i, j, k, sh + UMOMX) = zero; U(i, j, k, sh + UMOMY) = zero; U(i, j, k, sh + UMOMZ) = zero;
        U(i, j, k, sh + UENG) = netFluid3Pres / hgamma;
      }
    }
    return MFM_OK;
  }
  if (p->problem == MFM_PROBLEM_RIEMANN) {
    const double* L = p->rp_left;
    const double* R = p->rp_right;
    double b = d_1_4PI * (L[5] * L[5] + L[6] * L[6] + L[7] * L[7]);
    double ke = L[1] * L[1] + L[2] * L[2] + L[3] * L[3];
    double eL = L[4] / hgamma + half * (L[0] * ke + b);
    b = d_1_4PI * (R[5] * R[5] + R[6] * R[6] + R[7] * R[7]);
    ke = R[1] * R[1] + R[2] * R[2] + R[3] * R[3];
    double eR = R[4] / hgamma + half * (R[0] * ke + b);
    FORBOX(box, i, j, k) {
      double x = (i + half) * dx;
      const double* S = (x < p->rp_start_x) ? L : R;
      double e = (x < p->rp_start_x) ? eL : eR;
      U(i, j, k, URHO) = S[0];
      U(i, j, k, UMOMX) = S[0] * S[1]; U(i, j, k, UMOMY) = S[0] * S[2];
      U(i, j, k, UMOMZ) = S[0] * S[3];
      U(i, j, k, UENG) = e;
      U(i, j, k, UBX) = S[5]; U(i, j, k, UBY) = S[6]; U(i, j, k, UBZ) = S[7];
      if (E.nc > 8) U(i, j, k, 8) = zero;
    }
    return MFM_OK;
  }
  // MFM_PROBLEM_NONE: Orszag-Tang vortex (ORSZAGINIT)
  double p0 = d_1_4PI * E.gamma;
  double r0 = p0 * E.gamma;
  FORBOX(box, i, j, k) {
    double x = (i + half) * dx, y = (j + half) * dx;
    double Ux = -std::sin(d_2PI * y), Uy = std::sin(d_2PI * x);
    double Bx = -std::sin(d_2PI * y), By = std::sin((d_2PI + d_2PI) * x);
    double b = d_1_4PI * (Bx * Bx + By * By);
    double ke = Ux * Ux + Uy * Uy;
    double e0 = p0 / hgamma + half * (r0 * ke + b);
    U(i, j, k, URHO) = r0;
    U(i, j, k, UMOMX) = r0 * Ux; U(i, j, k, UMOMY) = r0 * Uy; U(i, j, k, UMOMZ) = zero;
    U(i, j, k, UENG) = e0;
    U(i, j, k, UBX) = Bx; U(i, j, k, UBY) = By; U(i, j, k, UBZ) = zero;
    if (E.dedner) U(i, j, k, E.iPsi) = zero;
  }
  return MFM_OK;
}

// The problem's physical-boundary ghost fill of direction dir on a primitive
// FAB (PhysBC::fillGhostCells of the problem), dx the cell size.
int oracle_fill_ghost(const mfm_params* p, mfm_fab Wf, int dir, double dx) {
  int rc = check_args(p);
  if (rc) return rc;
  Eq E(p, 1.0, 1.0);
  Fab W; W.wrap(Wf);
  fill_ghost(E, W, dir, dx);
  return MFM_OK;
}
// SWLISMREINIT on the conserved FAB U over box b, dx the cell size.
int oracle_swlism_reinit(const mfm_params* p, mfm_fab Uf, mfm_box b, double dx) {
  int rc = check_args(p);
  if (rc) return rc;
  Eq E(p, 1.0, 1.0);
  Fab U; U.wrap(Uf);
  swlism_reinit(E, U, dx, mkbox(b));
  return MFM_OK;
}
int oracle_fill_periodic(const mfm_params* p, mfm_fab Uf, mfm_box validb) {
  Fab U; U.wrap(Uf);
  Box v = mkbox(validb);
  int dims = p->spacedim;
  long n[3] = {v.n(0), v.n(1), v.n(2)};
  FORBOX(U.b, i, j, k) {
    int ii = i, jj = j, kk = k;
    int idx[3] = {i, j, k};
    for (int d = 0; d < dims; ++d) {
      long o = idx[d] - v.lo[d];
      o = ((o % n[d]) + n[d]) % n[d];
      idx[d] = (int)(v.lo[d] + o);
    }
    ii = idx[0]; jj = idx[1]; kk = idx[2];
    if (ii == i && jj == j && kk == k) continue;
    for (int c = 0; c < U.nc; ++c) U(i, j, k, c) = U(ii, jj, kk, c);
  }
  return MFM_OK;
}

// ---- pointwise entry points ------------------------------------------------
void oracle_flux_roe8(const mfm_params* p, const double* wl, const double* wr,
                      int idir, int averaging, int lax_friedrichs, double* out) {
  Eq E(p, 1.0, 1.0);
  E.iAveraging = averaging;
  E.iLaxFriedrix = lax_friedrichs;
  Rot r = rot(0, idir);
  double PGasL = std::max(E.smallp, wl[r.p]), RhoL = std::max(E.smallr, wl[r.rho]);
  double PGasR = std::max(E.smallp, wr[r.p]), RhoR = std::max(E.smallr, wr[r.rho]);
  double f[8];
  roe8_core(E, averaging, lax_friedrichs, PGasL, RhoL, wl[r.n], wl[r.t1],
            wl[r.t2], wl[r.bn], wl[r.bt1], wl[r.bt2], PGasR, RhoR, wr[r.n],
            wr[r.t1], wr[r.t2], wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  out[r.rho] = f[0]; out[r.n] = f[1]; out[r.t1] = f[2]; out[r.t2] = f[3];
  out[r.p] = f[4]; out[r.bn] = f[5]; out[r.bt1] = f[6]; out[r.bt2] = f[7];
}
void oracle_flux_rusanov(const mfm_params* p, const double* wl,
                         const double* wr, int idir, int averaging, double* out) {
  Eq E(p, 1.0, 1.0);
  Rot r = rot(0, idir);
  double PGasL = std::max(E.smallp, wl[r.p]), RhoL = std::max(E.smallr, wl[r.rho]);
  double PGasR = std::max(E.smallp, wr[r.p]), RhoR = std::max(E.smallr, wr[r.rho]);
  double f[8];
  rusanov_core(E, averaging, PGasL, RhoL, wl[r.n], wl[r.t1], wl[r.t2], wl[r.bn],
               wl[r.bt1], wl[r.bt2], PGasR, RhoR, wr[r.n], wr[r.t1], wr[r.t2],
               wr[r.bn], wr[r.bt1], wr[r.bt2], f);
  out[r.rho] = f[0]; out[r.n] = f[1]; out[r.t1] = f[2]; out[r.t2] = f[3];
  out[r.p] = f[4]; out[r.bn] = f[5]; out[r.bt1] = f[6]; out[r.bt2] = f[7];
}
void oracle_flux_dedner(const mfm_params* p, double ch, const double* wl,
                        const double* wr, int idir, int averaging, double* out) {
  mfm_params q = *p;
  q.dedner = 1;
  q.nfluids = 1;
  q.flux = averaging ? MFM_FLUX_ROE8_ROEAVG : MFM_FLUX_ROE8;
  Eq E(&q, ch, 1.0);
  plasma_flux_face(E, wl, wr, idir, out);
}
// The plasma solver the patch update selects from p->flux / p->dedner
// (RiemannSolver::make, src/RiemannSolver.cpp:39-104), at one face.
void oracle_flux_face(const mfm_params* p, double ch, const double* wl,
                      const double* wr, int idir, double* out) {
  Eq E(p, ch, 1.0);
  plasma_flux_face(E, wl, wr, idir, out);
}
void oracle_flux_vanleer(const mfm_params* p, const double* wl,
                         const double* wr, int idir, double* out) {
  Eq E(p, 1.0, 1.0);
  neutral_flux_face(E, wl, wr, idir, 0, out);
}
void oracle_helio_bflux(const mfm_params* p, int radial, const double* win,
                        const double* wout, const double* dxyz, double* out) {
  Eq E(p, 1.0, 1.0);
  helio_bflux(E, radial != 0, win[0], win[1], win[2], win[3], win[4], win[5],
              win[6], win[7], wout[0], wout[1], wout[2], wout[3], wout[4],
              wout[5], wout[6], wout[7], dxyz[0], dxyz[1], dxyz[2], out, &out[8],
              &out[10]);
  out[9] = std::isnan(out[10]) ? 0.0 : 1.0;
}
void oracle_helio_bc_face(const mfm_params* p, int lohisign, int dir, double dx,
                          const int* ijk, const double* wleft,
                          const double* wright, double* f, double* bn) {
  Eq E(p, 1.0, 1.0);
  helio_bc_face(E, lohisign, dir, dx, ijk[0], ijk[1], ijk[2], wleft, wright, f, bn);
}
void oracle_bfluxgas(const mfm_params* p, const double* win,
                     const double* wout, double* f5) {
  Eq E(p, 1.0, 1.0);
  swlism_bfluxgas(E, win[0], win[1], win[2], win[3], win[4], wout[0], wout[1],
                  wout[2], wout[3], wout[4], f5);
}
void oracle_limiter(int lim, double wm, double w0, double wp, double* out) {
  limit(lim, wm, w0, wp, &out[0], &out[1]);
}
// Face values of the primitive FAB W over box b in direction dir
// (Reconstruction::faceValues with the limiters and the characteristic
// switch of p).  W must hold b grown by one cell in dir.
int oracle_face_values(const mfm_params* p, mfm_fab Wf, int dir, mfm_box b,
                       mfm_fab WMf, mfm_fab WPf) {
  int rc = check_args(p);
  if (rc) return rc;
  Eq E(p, 1.0, 1.0);
  Fab W, WM, WP;
  W.wrap(Wf); WM.wrap(WMf); WP.wrap(WPf);
  face_values(E, W, WM, WP, dir, mkbox(b));
  return MFM_OK;
}
void oracle_face_slopes(int lim, double dsl, double dsr, double* out) {
  face_slopes(lim, dsl, dsr, &out[0], &out[1]);
}
// which = 0: analysis of d (8 primitive differences) with the basis of the
// primitive state w; which = 1: synthesis.
void oracle_char_transform(const mfm_params* p, const double* w, int dir, int which,
                           const double* in, double* out) {
  Eq E(p, 1.0, 1.0);
  if (which == 0) char_analysis(char_basis(E, w, dir, 1.0e-12), in, out);
  else char_synthesis(char_basis(E, w, dir, 1.0e-8), in, out);
}
void oracle_cons_to_prim(const mfm_params* p, const double* u, double* w) {
  Eq E(p, 1.0, 1.0);
  cons_to_prim_cell(E, u, w);
}
void oracle_physical_flux(const mfm_params* p, const double* w, int idir,
                          double* out) {
  // FLUXESHANCOCK (srcIdeal/PatchIdealMHDF.ChF:235-297)
  Eq E(p, 1.0, 1.0);
  Rot r = rot(0, idir);
  double rho = std::max(E.smallr, w[WRHO]);
  double u = w[r.n], v = w[r.t1], ww = w[r.t2];
  double pp = std::max(E.smallp, w[WPRES]);
  double bx = w[r.bn], by = w[r.bt1], bz = w[r.bt2];
  double b = d_1_8PI * (bx * bx + by * by + bz * bz);
  double ke = half * (u * u + v * v + ww * ww);
  double bu = bx * u + by * v + bz * ww;
  double e = pp * E.invhgamma + rho * ke;
  pp = pp + b;
  e = e + b;
  int inormc = UMOMX + idir, itanc1 = UMOMX + (idir + 1) % 3,
      itanc2 = UMOMX + (idir + 2) % 3;
  int inorcB = UBX + idir, itacB1 = UBX + (idir + 1) % 3,
      itacB2 = UBX + (idir + 2) % 3;
  out[URHO] = rho * u;
  out[inormc] = rho * u * u + pp - d_1_4PI * bx * bx;
  out[itanc1] = rho * u * v - d_1_4PI * bx * by;
  out[itanc2] = rho * u * ww - d_1_4PI * bx * bz;
  out[UENG] = u * (e + pp) - d_1_4PI * bx * bu;
  out[inorcB] = zero;
  out[itacB1] = u * by - v * bx;
  out[itacB2] = u * bz - ww * bx;
}
void oracle_charge_exchange(const mfm_params* p, const double* w, int reg,
                            double dt, double* s) {
  Eq E(p, 1.0, 1.0);
  charge_exchange_cell(E, w, reg, dt, s);
}
int oracle_region(const mfm_params* p, const double* w, int i, int j, int k,
                  double dx) {
  Eq E(p, 1.0, 1.0);
  return region_cell(E, w, i, j, k, dx);
}
double oracle_max_wave_speed(const mfm_params* p, const double* u) {
  Eq E(p, 1.0, 1.0);
  return max_wave_speed_cell(E, u);
}

// The geometric sources of coord_sys = 1 of one cell in row j.
void oracle_axisym_source(const mfm_params* p, const double* w, int j, double dt,
                          double dx, double* s) {
  Eq E(p, 1.0, 1.0);
  axisym_source_cell(E, w, j, dt, dx, s);
}

void oracle_floors(const mfm_params* p, double* u) {
  Eq E(p, 1.0, 1.0);
  floors_cell(E, u);
}

}  // extern "C"
