// mfm_device.hpp -- device-side physics of the multi-fluid MHD patch update.
//
// Per-cell / per-face functions for gfx950.  TF is the arithmetic type of
// reconstruction and Riemann fluxes (double in the parity build, float in the
// fast build); cons->prim recovery, floors, sources and the state update are
// always FP64.  Paths below are relative to MHDAM-3.2.4/ of the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "../../include/mfm.h"

namespace mfm {

// srcIdeal/EosCommon.fh:13-37
constexpr double k1_4PI = 0.079577471545947667884441881686257181;
constexpr double k1_8PI = 0.039788735772973833942220940843128591;
constexpr double kSQRT_2 = 0.707106781186547524400844362104849039;  // 1/sqrt(2)
constexpr double kPI = 3.14159265358979323846264338327950288;
constexpr double k2PI = 6.28318530717958647692528676655900576;
constexpr double k4PI = 12.566370614359172953850573533118;
constexpr double kPI_2 = 1.57079632679489661923132169163975144;
constexpr double kAU = 1.5e13;
constexpr double kNet2Dens = 1.0e-6, kNet2Pres = 1.0e-8;
constexpr double kNet3Dens = 1.0e-7, kNet3Pres = 1.0e-9;

constexpr int kMaxComp = 24;  // 8 + psi + 3*5

// Compile-time equation system: NF fluids (plasma + NF-1 neutrals), DED = 1
// adds the GLM psi component.  Every kernel is instantiated per system so
// all component loops unroll and per-cell vectors live in registers.
template <int NF_, int DED_>
struct Sys {
  static constexpr int NF = NF_;
  static constexpr int DED = DED_;
  static constexpr int NC = 8 + DED_ + 5 * (NF_ - 1);
  static constexpr int IRN = 8 + DED_;  // first neutral component
};

// Constants of one context, passed by value to every kernel (replaces the
// /eos/, /RiemSol/, /Dedner/, /chargeexcm/ and problem COMMON blocks).
struct Consts {
  int dims, nf, nc, dedner, iPsi, iRhoN;
  int flux_kind;      // 0 Roe8, 1 Rusanov, 2 Roe7, 3 HLLC (+ the GLM block with psi)
  int iAveraging;
  int eight_wave, divb_method, problem;
  int8_t lim[kMaxComp];
  int plasma_lim;     // limiter shared by every plasma variable (psi incl.), or -1
  int char_rec;       // plasma slopes limited in characteristic variables
  int tvd_mask;       // bit v: CHECKTVD clips the face values of variable v
  double gamma, hgamma, invhgamma, smallr, smallp, del2, smallB;
  double ch, cp;
  int dom_lo[3], dom_hi[3], phys[3];  // phys[d]: non-periodic physical dir
  // Riemann problem
  double rpL[8], rpR[8];
  // heliospheric
  double rp_start_x;
  double sun_xc, sun_yc, sun_zc, sun_r0, sun_rho, sun_p, sun_v, sun_b,
      sun_tilt, sun_int_bc_radius, dir_br_n, omega, lism_v_dim, init_r;
  int sun_b_monopolar;
  double lism_ux, lism_uy, lism_uz, lism_p, lism_m, lism_t, lism_bx, lism_by,
      lism_bz, net_rho, net_p, tmlim, reg_safe_z_top, reg_safe_z_bot;
  double ce_scale_vel, ce_sigma, c13493, c0531;
  int ce_on;
  int ce_mf;          // CHARGE_EXCHANGE_MF: neutral k is born where REG == k
  int axisym;         // coord_sys = 1: geometric sources, r-term of div(B)
  double r0_dt_ignore;  // SW/LISM: the dt exclusion radius
  // uniform factors of HELIOREINIT_DEFAULT, formed once on the host
  double sin_tilt, cos_tilt, sun_beta;
};

template <typename T>
__device__ __forceinline__ T tsqrt(T x) { return sqrt(x); }
template <typename T>
__device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T>
__device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
// Fortran MAX(a, b) with b = floor: returns a if a > b else b
template <typename T>
__device__ __forceinline__ T fmaxr(T b, T a) { return a > b ? a : b; }

// In FP32 a rounding error can push a mathematically non-negative radicand
// below zero; the FP64 build keeps the reference arithmetic untouched.
template <typename T>
__device__ __forceinline__ T nonneg(T x) {
  if constexpr (std::is_same<T, float>::value) return x > 0.f ? x : 0.f;
  return x;
}

// ------------------------------------------------------------ limiters --
// src/reconstructionF.ChF: face values of a cell (left = W-, right = W+)
// from its neighbours wl, wr along the sweep direction.
template <typename T>
__device__ __forceinline__ void limit(int lim, T wl, T w0, T wr, T& left,
                                      T& right) {
  const T dWR = wr - w0, dWL = w0 - wl;
  const T z = T(0), h = T(0.5);
  switch (lim) {
    case MFM_LIM_MINMOD: {  // minmod, :126-155
      T s = copysign(T(1), dWR);
      T m = s * tmax(z, tmin(fabs(dWR), s * dWL));
      T sl = h * m;
      left = w0 - sl; right = w0 + sl;
      return;
    }
    case MFM_LIM_VANLEER: {  // :794-841 (kappa = 1/3)
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      const T BT = T((3.0 - 1.0 / 3.0) / (1.0 - 1.0 / 3.0));
      const T SL1 = T(0.25 * (1.0 - 1.0 / 3.0)), SL2 = T(0.25 * (1.0 + 1.0 / 3.0));
      T SL = tmin(fabs(dWR), fabs(BT * dWL));
      T SR = tmin(fabs(dWL), fabs(BT * dWR));
      T pl = SL1 * SL + SL2 * SR, pr = SL1 * SR + SL2 * SL;
      if (dWR < z) { pl = -pl; pr = -pr; }
      left = w0 - pl; right = w0 + pr;
      return;
    }
    case MFM_LIM_MC: {  // :622-662
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      T AR = fabs(dWR), AL = fabs(dWL), s;
      if (T(3) * AL <= AR) s = dWL;
      else if (AL <= T(3) * AR) s = T(0.25) * (dWR + dWL);
      else s = dWR;
      left = w0 - s; right = w0 + s;
      return;
    }
    case MFM_LIM_SUPERBEE: {  // :258-300
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      T AR = fabs(dWR), AL = fabs(dWL), s;
      if (AL <= h * AR) s = dWL;
      else if (AL <= AR) s = h * dWR;
      else if (AL <= AR + AR) s = h * dWL;
      else s = dWR;
      left = w0 - s; right = w0 + s;
      return;
    }
    case MFM_LIM_VANALBADA: {  // :550-589
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      const T dS = T(1.0e-6), dSS = T(1.0e-6) + T(1.0e-6);
      T s = h * (dWR * dWL + dS) * (dWR + dWL) / (dWR * dWR + dWL * dWL + dSS);
      left = w0 - s; right = w0 + s;
      return;
    }
    case MFM_LIM_HARMONIC: {  // :492-524
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      T s = dWR / (T(1) + dWR / dWL);
      left = w0 - s; right = w0 + s;
      return;
    }
    case MFM_LIM_HYPERBEE: {  // :430-464
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      T A = T(1) + tsqrt(dWR / dWL);
      T s = dWR / (A * A);
      s = s + s;
      left = w0 - s; right = w0 + s;
      return;
    }
    case MFM_LIM_KOREN: {  // :883-918
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      T H = dWL * dWR / (T(2) * (dWL * dWL + dWR * dWR + T(1.0e-10)) - dWL * dWR);
      left = w0 - (dWL + h * dWR) * H;
      right = w0 + (dWR + h * dWL) * H;
      return;
    }
    case MFM_LIM_VENKAT: {  // :950-988
      if (dWL * dWR <= z) { left = w0; right = w0; return; }
      const T dS = T(1.0e-10), dS2 = T(1.0e-10 * 1.0e-10);
      T H = (dWR + dWL) / (T(6) * (dWR * dWR + dWL * dWL + dS2));
      left = w0 - (dWR + T(2) * dWL) * (dWL + dS) * H;
      right = w0 + (dWL + T(2) * dWR) * (dWR + dS) * H;
      return;
    }
    case MFM_LIM_WENO3: {  // :1096-1136
      T BR = dWR * dWR + T(1.0e-6), AR = T(1) / (BR * BR);
      T BL = dWL * dWL + T(1.0e-6), AL = T(1) / (BL * BL);
      T H = AR / (AR + T(2) * AL);
      left = w0 - h * (dWL + H * (dWR - dWL));
      H = AL / (AL + T(2) * AR);
      right = w0 + h * (dWR - H * (dWR - dWL));
      return;
    }
    case MFM_LIM_WENO3YC: {  // :1169-1212
      T BR = dWR * dWR + T(1.0e-6), BL = dWL * dWL + T(1.0e-6);
      T D2P = dWR - dWL;
      D2P = D2P * D2P;
      T AL = T(1) + D2P / BL, AR = T(1) + D2P / BR;
      T H = AR / (AR + T(2) * AL);
      left = w0 - h * (dWL + H * (dWR - dWL));
      H = AL / (AL + T(2) * AR);
      right = w0 + h * (dWR - H * (dWR - dWL));
      return;
    }
    case MFM_LIM_CD:  // :1056-1079 (W- takes the +1 neighbour, as written)
      left = h * (w0 + wr);
      right = h * (w0 + wl);
      return;
    default:
      left = w0; right = w0;
      return;
  }
}

// ----------------------------------------- characteristic reconstruction --
// *FACESLOPES (src/reconstructionF.ChF): limited slopes SL (lo side) and SR
// (hi side) of a cell from its one-sided differences dSL = W - W(i-1),
// dSR = W(i+1) - W, taken in characteristic variables.
template <typename T>
__device__ __forceinline__ void face_slopes(int lim, T dSL, T dSR, T& SL, T& SR) {
  const T z = T(0), h = T(0.5);
  T s = z;
  switch (lim) {
    case MFM_LIM_MINMOD: {  // MINMODFACESLOPES :157-177
      T sg = copysign(T(1), dSR);
      s = h * (sg * tmax(z, tmin(fabs(dSR), sg * dSL)));
      break;
    }
    case MFM_LIM_VANLEER: {  // VANLEERFACESLOPES :843-881 (kappa = 1/3)
      T pl = z, pr = z;
      if (dSL * dSR > z) {
        const T BT = T((3.0 - 1.0 / 3.0) / (1.0 - 1.0 / 3.0));
        const T SL1 = T(0.25 * (1.0 - 1.0 / 3.0)), SL2 = T(0.25 * (1.0 + 1.0 / 3.0));
        T a = tmin(fabs(dSR), fabs(BT * dSL));
        T b = tmin(fabs(dSL), fabs(BT * dSR));
        pl = SL1 * a + SL2 * b;
        pr = SL1 * b + SL2 * a;
        if (dSR < z) { pl = -pl; pr = -pr; }
      }
      SL = pl; SR = pr;
      return;
    }
    case MFM_LIM_MC:  // MCFACESLOPES :664-694
      if (dSL * dSR > z) {
        T AR = fabs(dSR), AL = fabs(dSL);
        if (T(3) * AL <= AR) s = dSL;
        else if (AL <= T(3) * AR) s = T(0.25) * (dSR + dSL);
        else s = dSR;
      }
      break;
    case MFM_LIM_SUPERBEE:  // SUPERBEEFACESLOPES :356-388
      if (dSL * dSR > z) {
        T AR = fabs(dSR), AL = fabs(dSL);
        if (AL <= h * AR) s = dSL;
        else if (AL <= AR) s = h * dSR;
        else if (AL <= AR + AR) s = h * dSL;
        else s = dSR;
      }
      break;
    case MFM_LIM_VANALBADA:  // VANALBADAFACESLOPES :591-620
      if (dSL * dSR > z) {
        const T dS = T(1.0e-6), dSS = T(1.0e-6) + T(1.0e-6);
        s = h * (dSR * dSL + dS) * (dSR + dSL) / (dSR * dSR + dSL * dSL + dSS);
      }
      break;
    case MFM_LIM_HARMONIC:  // HARMONICFACESLOPES :526-548
      if (dSL * dSR > z) s = dSR * dSL / (dSL + dSR);
      break;
    case MFM_LIM_HYPERBEE:  // HYPERBEEFACESLOPES :466-490
      if (dSL * dSR > z) {
        T A = T(1) + tsqrt(dSR / dSL);
        s = dSR / (A * A);
        s = s + s;
      }
      break;
    case MFM_LIM_KOREN: {  // KORENFACESLOPES :920-948
      T pl = z, pr = z;
      if (dSL * dSR > z) {
        T H = dSL * dSR / (T(2) * (dSL * dSL + dSR * dSR + T(1.0e-10)) - dSL * dSR);
        pl = (dSL + h * dSR) * H;
        pr = (dSR + h * dSL) * H;
      }
      SL = pl; SR = pr;
      return;
    }
    case MFM_LIM_VENKAT: {  // VENKATAKRISHNANFACESLOPES :990-1020
      T pl = z, pr = z;
      if (dSL * dSR > z) {
        const T dS = T(1.0e-10), dS2 = T(1.0e-10 * 1.0e-10);
        T H = (dSR + dSL) / (T(6) * (dSR * dSR + dSL * dSL + dS2));
        pl = (dSR + T(2) * dSL) * (dSL + dS) * H;
        pr = (dSL + T(2) * dSR) * (dSR + dS) * H;
      }
      SL = pl; SR = pr;
      return;
    }
    case MFM_LIM_WENO3: {  // WENO3FACESLOPES :1138-1167
      T BR = dSR * dSR + T(1.0e-6), AR = T(1) / (BR * BR);
      T BL = dSL * dSL + T(1.0e-6), AL = T(1) / (BL * BL);
      T H = AR / (AR + T(2) * AL);
      SL = h * (dSL + H * (dSR - dSL));
      H = AL / (AL + T(2) * AR);
      SR = h * (dSR - H * (dSR - dSL));
      return;
    }
    case MFM_LIM_WENO3YC: {  // WENO3YCFACESLOPES :1214-1247
      T BR = dSR * dSR + T(1.0e-6), BL = dSL * dSL + T(1.0e-6);
      T D2P = dSR - dSL;
      D2P = D2P * D2P;
      T AL = T(1) + D2P / BL, AR = T(1) + D2P / BR;
      T H = AR / (AR + T(2) * AL);
      SL = h * (dSL + H * (dSR - dSL));
      H = AL / (AL + T(2) * AR);
      SR = h * (dSR - H * (dSR - dSL));
      return;
    }
    case MFM_LIM_CD:  // CDFACESLOPES :1081-1094
      SL = h * dSL; SR = h * dSR;
      return;
    default:  // no limiter: the slopes stay as they are
      SL = dSL; SR = dSR;
      return;
  }
  SL = s; SR = s;
}

// minmod3 (src/reconstructionF.ChF:16-26)
template <typename T>
__device__ __forceinline__ T minmod3(T A, T B, T C) {
  const T sg = copysign(T(1), A);
  return sg * tmax(T(0), tmin(tmin(fabs(A), sg * B), sg * C));
}

// CHECKTVD (src/reconstructionF.ChF:1248-1278) on one face value: the
// slope W - W- (lo) or W+ - W (hi) is clipped to the cell's one-sided
// differences.
template <typename T>
__device__ __forceinline__ T tvd_lo(T wm, T w0, T dWL, T dWR) {
  return w0 - minmod3(dWL, dWR, w0 - wm);
}
template <typename T>
__device__ __forceinline__ T tvd_hi(T wp, T w0, T dWL, T dWR) {
  return w0 + minmod3(dWL, dWR, wp - w0);
}

// Eigen-structure of the ideal-MHD system at one cell along dir, formed from
// the cell's primitive state as in CHARANALYSISF / CHARSYNTHESISF
// (srcIdeal/PatchIdealMHDF.ChF:1167-1505).  The fast / slow weights take the
// degenerate branch below 1e-12 in the analysis (aF, aS) and below 1e-8 in
// the synthesis (aF2, aS2).
template <typename T>
struct CharBasis {
  T Rho, cc, c, af, as, aF, aS, aF2, aS2, betaY, betaZ, BX;
};

template <typename T>
__device__ __forceinline__ void char_alpha(T dfs, T cc, T aas, T aaf, T eps, T& aF, T& aS) {
  if (dfs < eps) { aF = T(1); aS = T(0); }
  else if ((cc - aas) <= T(0)) { aF = T(0); aS = T(1); }
  else if ((aaf - cc) <= T(0)) { aF = T(1); aS = T(0); }
  else { aF = tsqrt((cc - aas) / dfs); aS = tsqrt((aaf - cc) / dfs); }
}

template <typename T, int dir>
__device__ __forceinline__ CharBasis<T> char_basis(const Consts& C, const T* w) {
  constexpr int bn = 5 + dir, bt1 = 5 + (dir + 1) % 3, bt2 = 5 + (dir + 2) % 3;
  CharBasis<T> B;
  const T PGas = fmaxr((T)C.smallp, w[4]);
  const T Rho = fmaxr((T)C.smallr, w[0]);
  const T BX = w[bn], BY = w[bt1], BZ = w[bt2];
  const T cc = T(C.gamma) * PGas / Rho;
  const T c = tsqrt(cc);
  const T BX2 = BX * BX, BY2 = BY * BY, BZ2 = BZ * BZ;
  const T BYZ2 = BY2 + BZ2;
  const T BxBH = T(k1_8PI) * (BX2 + BYZ2);
  const T bb = T(k1_4PI) * BX2 / Rho;
  const T b = tsqrt(bb);
  T aas, as, aaf, af;
  if (BYZ2 < T(C.smallB) * T(C.smallB)) {
    B.betaY = T(kSQRT_2);
    B.betaZ = T(kSQRT_2);
    if (bb > cc) { aaf = bb; aas = cc; af = b; as = c; }
    else { aaf = cc; aas = bb; af = c; as = b; }
  } else {
    const T hyz = tsqrt(BYZ2);
    B.betaY = BY / hyz;
    B.betaZ = BZ / hyz;
    const T szb = cc + (BxBH + BxBH) / Rho;
    const T h = tsqrt(nonneg(szb * szb - T(4) * cc * bb));
    aas = T(0.5) * (szb - h);
    as = tsqrt(nonneg(aas));
    aaf = T(0.5) * (szb + h);
    af = tsqrt(aaf);
  }
  const T dfs = aaf - aas;
  char_alpha<T>(dfs, cc, aas, aaf, T(1.0e-12), B.aF, B.aS);
  char_alpha<T>(dfs, cc, aas, aaf, T(1.0e-8), B.aF2, B.aS2);
  B.Rho = Rho; B.cc = cc; B.c = c; B.af = af; B.as = as; B.BX = BX;
  return B;
}

// CHARANALYSISF (:1252-1325): primitive differences d (plasma components in
// Cartesian order) -> characteristic variables a (entropy, div(B), two
// Alfven, two fast, two slow).
template <typename T, int dir>
__device__ __forceinline__ void char_analysis(const CharBasis<T>& B, const T* d, T* a) {
  constexpr int n = 1 + dir, t1 = 1 + (dir + 1) % 3, t2 = 1 + (dir + 2) % 3;
  constexpr int bn = 5 + dir, bt1 = 5 + (dir + 1) % 3, bt2 = 5 + (dir + 2) % 3;
  // SIGN(one, BX) with -0.0 counted as positive, as the synthesis's BX < 0
  // test does, so that both transforms agree where the field is exactly zero
  const T two = T(2), siBX = B.BX < T(0) ? T(-1) : T(1);
  const T dP = d[4], dRho = d[0], dUn = d[n], dUt1 = d[t1], dUt2 = d[t2];
  const T dBn = d[bn], dBt1 = d[bt1], dBt2 = d[bt2];
  a[0] = dRho - dP / B.cc;
  a[1] = dBn;
  T h1 = T(kSQRT_2) * (B.betaY * dUt2 - B.betaZ * dUt1);
  T h2 = siBX * (B.betaZ * dBt1 - B.betaY * dBt2) / (two * tsqrt(T(k2PI) * B.Rho));
  a[2] = h1 + h2;
  a[3] = h1 - h2;
  h1 = dP / (two * B.Rho * B.cc);
  h2 = dUn / (two * B.cc);
  const T h3 = (B.betaY * dBt1 + B.betaZ * dBt2) / (T(4) * B.c * tsqrt(T(kPI) * B.Rho));
  const T h4 = siBX * (B.betaY * dUt1 + B.betaZ * dUt2) / (two * B.cc);
  const T aF = B.aF, aS = B.aS, af = B.af, as = B.as;
  a[4] = aF * (h1 + af * h2) + aS * (h3 - as * h4);
  a[5] = aF * (h1 - af * h2) + aS * (h3 + as * h4);
  a[6] = aS * (h1 + as * h2) - aF * (h3 - af * h4);
  a[7] = aS * (h1 - as * h2) - aF * (h3 + af * h4);
}

// CHARSYNTHESISF (:1438-1502): characteristic variables -> primitive
// differences.
template <typename T, int dir>
__device__ __forceinline__ void char_synthesis(const CharBasis<T>& B, const T* a, T* d) {
  constexpr int n = 1 + dir, t1 = 1 + (dir + 1) % 3, t2 = 1 + (dir + 2) % 3;
  constexpr int bn = 5 + dir, bt1 = 5 + (dir + 1) % 3, bt2 = 5 + (dir + 2) % 3;
  const T betaYX = B.BX < T(0) ? -B.betaY : B.betaY;
  const T betaZX = B.BX < T(0) ? -B.betaZ : B.betaZ;
  const T alpha3 = T(kSQRT_2) * a[2], alpha4 = T(kSQRT_2) * a[3];
  const T a3p4 = alpha3 + alpha4, a3m4 = alpha3 - alpha4;
  const T a5p6 = a[4] + a[5], a5m6 = a[4] - a[5];
  const T a7p8 = a[6] + a[7], a7m8 = a[6] - a[7];
  const T DN = T(2) * tsqrt(T(kPI) * B.Rho);
  const T aF = B.aF2, aS = B.aS2;
  const T HRho = B.Rho * (aF * a5p6 + aS * a7p8);
  const T HU = aS * B.as * a5m6 - aF * B.af * a7m8;
  const T HB = B.c * (aS * a5p6 - aF * a7p8);
  d[0] = a[0] + HRho;
  d[n] = aF * B.af * a5m6 + aS * B.as * a7m8;
  d[t1] = -B.betaZ * a3p4 - betaYX * HU;
  d[t2] = B.betaY * a3p4 - betaZX * HU;
  d[4] = B.cc * HRho;
  d[bn] = a[1];
  d[bt1] = DN * (betaZX * a3m4 + B.betaY * HB);
  d[bt2] = DN * (-betaYX * a3m4 + B.betaZ * HB);
}

// ---------------------------------------------------- cons -> prim (FP64) --
// CONSTOPRIM / CONSTOPRIM_MF (srcIdeal/PatchIdealMHDF.ChF:143-189,
// srcIdeal/PatchMHDMFF.ChF:5-80): the multi-fluid version floors rho.
template <class S>
__device__ __forceinline__ void cons_to_prim(const Consts& C, const double* u,
                                             double* w) {
  double rho = u[0];
  if (S::NF > 1) rho = fmaxr(C.smallr, u[0]);
  const double ir = 1.0 / rho;
  const double vx = u[1] * ir, vy = u[2] * ir, vz = u[3] * ir;
  const double ke = u[1] * vx + u[2] * vy + u[3] * vz;
  const double b = k1_4PI * (u[5] * u[5] + u[6] * u[6] + u[7] * u[7]);
  double p = C.hgamma * (u[4] - 0.5 * (ke + b));
  p = fmaxr(C.smallp, p);
  w[0] = rho; w[1] = vx; w[2] = vy; w[3] = vz; w[4] = p;
  w[5] = u[5]; w[6] = u[6]; w[7] = u[7];
  if (S::DED) w[8] = u[8];
#pragma unroll
  for (int f = 1; f < S::NF; ++f) {
    const int s = S::IRN + 5 * (f - 1);
    const double r = fmaxr(C.smallr, u[s]);
    const double jr = 1.0 / r;
    const double ax = u[s + 1] * jr, ay = u[s + 2] * jr, az = u[s + 3] * jr;
    const double k2 = u[s + 1] * ax + u[s + 2] * ay + u[s + 3] * az;
    double q = C.hgamma * (u[s + 4] - 0.5 * k2);
    q = fmaxr(C.smallp, q);
    w[s] = r; w[s + 1] = ax; w[s + 2] = ay; w[s + 3] = az; w[s + 4] = q;
  }
}

// ---------------------------------------------------------- MHD fluxes --
// Rotated plasma state: p, rho, un, ut1, ut2, bn, bt1, bt2.
template <typename T>
struct MState { T p, rho, u, v, w, bx, by, bz; };

// Differences (left - right) of the conserved variables (cons) and the
// primitive-like jumps the left eigenvectors act on (out): the du block of
// RIEMANNF / RIEMANN7F (src/Roe8WaveF.ChF).
template <typename T, bool seven>
__device__ __forceinline__ void roe_jumps(const Consts& C, const MState<T>& L,
                                          const MState<T>& R, T Rho_, T U_, T V_,
                                          T W_, T BX_, T BY_, T BZ_, T kk_, T* out,
                                          T* cons) {
  using D = T;
  const D half = D(0.5), c4 = D(k1_4PI), c8 = D(k1_8PI), hg = D(C.hgamma);
  const D Lr = L.rho, Lu = L.u, Lv = L.v, Lw = L.w, Lp = L.p, Lbx = L.bx, Lby = L.by, Lbz = L.bz;
  const D Rr = R.rho, Ru = R.u, Rv = R.v, Rw = R.w, Rp = R.p, Rbx = R.bx, Rby = R.by, Rbz = R.bz;
  const D Rho = Rho_, U = U_, V = V_, W = W_, BX = BX_, BY = BY_, BZ = BZ_, kk = kk_;
  const D RUL = Lr * Lu, RVL = Lr * Lv, RWL = Lr * Lw;
  const D BxBL = c8 * (Lbx * Lbx + Lby * Lby + Lbz * Lbz);
  const D RUxUL = half * (RUL * Lu + RVL * Lv + RWL * Lw);
  const D EL = Lp / hg + RUxUL + BxBL;
  const D RUR = Rr * Ru, RVR = Rr * Rv, RWR = Rr * Rw;
  const D BxBR = c8 * (Rbx * Rbx + Rby * Rby + Rbz * Rbz);
  const D RUxUR = half * (RUR * Ru + RVR * Rv + RWR * Rw);
  const D ER = Rp / hg + RUxUR + BxBR;
  D du1 = Lr - Rr, du2 = RUL - RUR, du3 = RVL - RVR, du4 = RWL - RWR;
  D du5 = EL - ER;
  const D du6 = Lbx - Rbx, du7 = Lby - Rby, du8 = Lbz - Rbz;
  cons[0] = T(du1); cons[1] = T(du2); cons[2] = T(du3); cons[3] = T(du4);
  cons[4] = T(du5); cons[5] = T(du6); cons[6] = T(du7); cons[7] = T(du8);
  if constexpr (!seven) {
    du5 = -hg * (-kk * du1 + U * du2 + V * du3 + W * du4 - du5 +
                 (BX * du6 + BY * du7 + BZ * du8) * c4);
  } else {
    du5 = -hg * (-kk * du1 + U * du2 + V * du3 + W * du4 - du5 +
                 (BY * du7 + BZ * du8) * c4);
  }
  du2 = (-U * du1 + du2) / Rho;
  du3 = (-V * du1 + du3) / Rho;
  du4 = (-W * du1 + du4) / Rho;
  out[0] = T(du1); out[1] = T(du2); out[2] = T(du3); out[3] = T(du4);
  out[4] = T(du5); out[5] = T(du6); out[6] = T(du7); out[7] = T(du8);
}

// 8-wave Roe flux (RIEMANNF, src/Roe8WaveF.ChF:5-601), left/right already
// clamped to (smallp, smallr).  Only the non-zero eigenvector entries are
// formed; the sums keep the reference's term order, so the FP64 build is
// bit-compatible with the dense form.  out order: rho, mn, mt1, mt2, E, bn,
// bt1, bt2.  seven = true gives RIEMANN7F (src/Roe7WaveF.ChF:5-558): no B_n
// wave, out[5] = 0.
template <typename T, bool seven>
__device__ __forceinline__ void roe_flux(const Consts& C, const MState<T>& L,
                                         const MState<T>& R, T* out) {
  const T half = T(0.5), one = T(1), zero = T(0), two = T(2);
  const T g = T(C.gamma), hg = T(C.hgamma);
  const T c4 = T(k1_4PI), c8 = T(k1_8PI), rs2 = T(kSQRT_2);

  const T RUL = L.rho * L.u, RVL = L.rho * L.v, RWL = L.rho * L.w;
  const T BxBL = c8 * (L.bx * L.bx + L.by * L.by + L.bz * L.bz);
  const T RUxUL = half * (RUL * L.u + RVL * L.v + RWL * L.w);
  const T PL = L.p + BxBL;
  const T EL = L.p / hg + RUxUL + BxBL;
  const T AL2 = g * L.p / L.rho;
  const T RUR = R.rho * R.u, RVR = R.rho * R.v, RWR = R.rho * R.w;
  const T BxBR = c8 * (R.bx * R.bx + R.by * R.by + R.bz * R.bz);
  const T RUxUR = half * (RUR * R.u + RVR * R.v + RWR * R.w);
  const T PR = R.p + BxBR;
  const T ER = R.p / hg + RUxUR + BxBR;
  const T AR2 = g * R.p / R.rho;
  T ccmi, ccma;
  if (AL2 > AR2) { ccmi = AR2; ccma = AL2; } else { ccmi = AL2; ccma = AR2; }

  T Rho, U, V, W, BX, BY, BZ, BX2, BYZ2, BxBH, kk, cc;
  if (C.iAveraging == 1) {
    const T HL = (EL + PL) / L.rho, HR = (ER + PR) / R.rho;
    const T cL = one / (one + tsqrt(R.rho / L.rho));
    const T cR = one - cL;
    Rho = R.rho * cL + L.rho * cR;
    U = L.u * cL + R.u * cR;
    V = L.v * cL + R.v * cR;
    W = L.w * cL + R.w * cR;
    const T HAV = HL * cL + HR * cR;
    BX = R.bx * cL + L.bx * cR;
    BY = R.by * cL + L.by * cR;
    BZ = R.bz * cL + L.bz * cR;
    const T BXAV = half * (L.bx + R.bx);
    BX2 = BXAV * BXAV;
    BYZ2 = BY * BY + BZ * BZ;
    BxBH = c8 * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    const T dbx = R.bx - L.bx, dby = R.by - L.by, dbz = R.bz - L.bz;
    const T X = c8 * (dbx * dbx + dby * dby + dbz * dbz) * cL * cL / L.rho;
    cc = hg * (HAV - kk - c4 * (BX * BXAV + BYZ2) / Rho) + (two - g) * X;
  } else {
    Rho = half * (L.rho + R.rho);
    U = half * (L.u + R.u); V = half * (L.v + R.v); W = half * (L.w + R.w);
    BX = half * (L.bx + R.bx); BY = half * (L.by + R.by); BZ = half * (L.bz + R.bz);
    const T P = half * (PL + PR);
    BX2 = BX * BX;
    BYZ2 = BY * BY + BZ * BZ;
    BxBH = c8 * (BX2 + BYZ2);
    kk = half * (U * U + V * V + W * W);
    cc = g * (P - BxBH) / Rho;
  }
  const T bb = c4 * BX2 / Rho;
  const T b = tsqrt(bb);
  if (cc < ccmi) cc = ccmi;
  if (cc > ccma) cc = ccma;
  const T c = tsqrt(cc);

  T hyy, hzz, aaf, aas, af, as;
  const T sB = T(C.smallB);
  if (BYZ2 < sB * sB) {
    hyy = rs2; hzz = rs2;
    if (bb > cc) { aaf = bb; aas = cc; af = b; as = c; }
    else { aaf = cc; aas = bb; af = c; as = b; }
  } else {
    const T hyz = tsqrt(BYZ2);
    hyy = BY / hyz; hzz = BZ / hyz;
    const T szb = cc + (BxBH + BxBH) / Rho;
    const T hh = tsqrt(nonneg(szb * szb - T(4) * cc * bb));
    aas = half * (szb - hh);
    if (aas <= zero) { aas = zero; as = zero; }
    else if (aas >= cc) { aas = cc; as = c; }
    else as = tsqrt(aas);
    aaf = half * (szb + hh);
    if (aaf <= cc) { aaf = cc; af = c; }
    else af = tsqrt(aaf);
  }

  // |lambda| with the entropy fix
  T A1, A3, A4, A5, A6, A7, A8;  // A2 == A1 (both |U|)
  {
    const T eigmax = fabs(U) + fabs(V) + fabs(W) + af;
    const T del = T(C.del2) * eigmax;
    const T dl2 = del * del;
    auto fix = [&](T e) -> T {
      const T zz = fabs(e);
      return zz >= del ? zz : half * (zz * zz + dl2) / del;
    };
    A1 = fix(U); A3 = fix(U + b); A4 = fix(U - b);
    A5 = fix(U + af); A6 = fix(U - af); A7 = fix(U + as); A8 = fix(U - as);
  }
  const T dfs = aaf - aas;
  T alf, als;
  if (dfs < T(1.0e-8)) { alf = one; als = zero; }
  else if ((cc - aas) <= zero) { alf = zero; als = one; }
  else if ((aaf - cc) <= zero) { alf = one; als = zero; }
  else { alf = tsqrt((cc - aas) / dfs); als = tsqrt((aaf - cc) / dfs); }
  const T sih = (BX >= zero) ? one : -one;

  T du[8], dcons[8];
  roe_jumps<T, seven>(C, L, R, Rho, U, V, W, BX, BY, BZ, kk, du, dcons);
  const T du1 = du[0], du2 = du[1], du3 = du[2], du4 = du[3], du5 = du[4];
  const T du6 = du[5], du7 = du[6], du8 = du[7];
  const T sqPiRho = tsqrt(T(kPI) * Rho), sq2PiRho = tsqrt(T(k2PI) * Rho);

  // non-zero entries of the left/right eigenvector matrices
  const T ic = one / cc;
  const T aff = alf * af, ass = als * as;
  const T hzs = hzz * rs2, hys = hyy * rs2;
  const T q7 = half * hzz * sih / sq2PiRho, q8 = half * hyy * sih / sq2PiRho;
  const T csp = c * sqPiRho;
  // fast/slow rows
  const T so52 = half * alf * af / cc;
  const T so53 = -half * als * as * hyy * sih / cc;
  const T so54 = -half * als * as * hzz * sih / cc;
  const T so55 = half * alf / (cc * Rho);
  const T so57 = T(0.25) * als * hyy / csp, so58 = T(0.25) * als * hzz / csp;
  const T so72 = half * als * as / cc;
  const T so73 = half * alf * af * hyy * sih / cc;
  const T so74 = half * alf * af * hzz * sih / cc;
  const T so75 = half * als / (Rho * cc);
  const T so77 = -T(0.25) * alf * hyy / csp, so78 = -T(0.25) * alf * hzz / csp;

  const T ra_f = Rho * alf, ra_s = Rho * als;
  const T sp35 = -ass * hyy * sih, sp37 = aff * hyy * sih;
  const T sp45 = -ass * hzz * sih, sp47 = aff * hzz * sih;
  const T sp55 = alf * Rho * cc, sp57 = als * Rho * cc;
  const T sp73 = hzz * sq2PiRho * sih, sp83 = -hyy * sq2PiRho * sih;
  const T sp75 = two * als * sqPiRho * c * hyy, sp77 = -two * alf * sqPiRho * c * hyy;
  const T sp85 = two * als * sqPiRho * c * hzz, sp87 = -two * alf * sqPiRho * c * hzz;

  T ra1, ra2, ra3, ra4, ra5, ra6, ra7, ra8;
  if constexpr (std::is_same<T, float>::value) {
    // FP32: |Lambda| = A1 I + D.  The A1 part is A1 times the conserved
    // jumps (R L = I); only D = |Lambda| - A1 I goes through the
    // eigenvectors.  With the reference's form, the entropy wave and the
    // acoustic waves carry opposite dp/c^2 terms that cancel, which at
    // large Mach number (c^2 << u^2, e.g. a pressure on its floor) FP32
    // cannot resolve.  Pairs of waves are combined as half-sums and
    // half-differences of D, each formed without cancellation.
    const T del = T(C.del2) * (fabs(U) + fabs(V) + fabs(W) + af);
    auto fix = [&](T e) -> T {
      const T zz = fabs(e);
      return zz >= del ? zz : half * (zz * zz + del * del) / del;
    };
    auto fixdiff = [&](T sh) -> T {  // fix(U + sh) - fix(U)
      const T e1 = U + sh, z1 = fabs(e1), z0 = fabs(U);
      if (z1 >= del && z0 >= del) {
        if ((e1 >= zero) == (U >= zero)) return U >= zero ? sh : -sh;
        return z1 - z0;
      }
      if (z1 < del && z0 < del) return half * sh * (U + U + sh) / del;
      return fix(e1) - fix(U);
    };
    const T d3 = fixdiff(b), d4 = fixdiff(-b), d5 = fixdiff(af), d6 = fixdiff(-af);
    const T d7 = fixdiff(as), d8 = fixdiff(-as);
    const T Pa = -hzs * du3 + hys * du4, Qa = q7 * du7 + (-q8) * du8;
    const T Xf = so52 * du2 + so53 * du3 + so54 * du4;
    const T Yf = so55 * du5 + so57 * du7 + so58 * du8;
    const T Xs = so72 * du2 + so73 * du3 + so74 * du4;
    const T Ys = so75 * du5 + so77 * du7 + so78 * du8;
    const T l3 = d3 * (Pa + Qa), l4 = d4 * (Pa - Qa);
    const T s56 = (d5 - d6) * Xf + (d5 + d6) * Yf, m56 = (d5 + d6) * Xf + (d5 - d6) * Yf;
    const T s78 = (d7 - d8) * Xs + (d7 + d8) * Ys, m78 = (d7 + d8) * Xs + (d7 - d8) * Ys;
    ra1 = ra_f * s56 + ra_s * s78;
    ra2 = aff * m56 + ass * m78;
    ra3 = -hzs * (l3 + l4) + sp35 * m56 + sp37 * m78;
    ra4 = hys * (l3 + l4) + sp45 * m56 + sp47 * m78;
    ra5 = sp55 * s56 + sp57 * s78;
    ra7 = sp73 * (l3 - l4) + sp75 * s56 + sp77 * s78;
    ra8 = sp83 * (l3 - l4) + sp85 * s56 + sp87 * s78;
    ra6 = zero;
    ra5 = ra1 * kk + Rho * (U * ra2 + V * ra3 + W * ra4) + ra5 / hg +
          c4 * (BY * ra7 + BZ * ra8);
    ra2 = U * ra1 + Rho * ra2;
    ra3 = V * ra1 + Rho * ra3;
    ra4 = W * ra1 + Rho * ra4;
    ra1 = ra1 + A1 * dcons[0];
    ra2 = ra2 + A1 * dcons[1];
    ra3 = ra3 + A1 * dcons[2];
    ra4 = ra4 + A1 * dcons[3];
    ra5 = ra5 + A1 * dcons[4];
    if constexpr (!seven) ra6 = A1 * dcons[5];
    ra7 = ra7 + A1 * dcons[6];
    ra8 = ra8 + A1 * dcons[7];
  } else {
    const T so15 = -ic;
    const T lu1 = A1 * (du1 + so15 * du5);
    const T lu3 = A3 * (-hzs * du3 + hys * du4 + q7 * du7 + (-q8) * du8);
    const T lu4 = A4 * (-hzs * du3 + hys * du4 + (-q7) * du7 + q8 * du8);
    const T lu5 = A5 * (so52 * du2 + so53 * du3 + so54 * du4 + so55 * du5 +
                        so57 * du7 + so58 * du8);
    const T lu6 = A6 * ((-so52) * du2 + (-so53) * du3 + (-so54) * du4 +
                        so55 * du5 + so57 * du7 + so58 * du8);
    const T lu7 = A7 * (so72 * du2 + so73 * du3 + so74 * du4 + so75 * du5 +
                        so77 * du7 + so78 * du8);
    const T lu8 = A8 * ((-so72) * du2 + (-so73) * du3 + (-so74) * du4 +
                        so75 * du5 + so77 * du7 + so78 * du8);

    ra1 = lu1 + ra_f * lu5 + ra_f * lu6 + ra_s * lu7 + ra_s * lu8;
    ra2 = aff * lu5 + (-aff) * lu6 + ass * lu7 + (-ass) * lu8;
    ra3 = -hzs * lu3 + (-hzs) * lu4 + sp35 * lu5 + (-sp35) * lu6 + sp37 * lu7 +
            (-sp37) * lu8;
    ra4 = hys * lu3 + hys * lu4 + sp45 * lu5 + (-sp45) * lu6 + sp47 * lu7 +
            (-sp47) * lu8;
    ra5 = sp55 * lu5 + sp55 * lu6 + sp57 * lu7 + sp57 * lu8;
    ra7 = sp73 * lu3 + (-sp73) * lu4 + sp75 * lu5 + sp75 * lu6 + sp77 * lu7 +
            sp77 * lu8;
    ra8 = sp83 * lu3 + (-sp83) * lu4 + sp85 * lu5 + sp85 * lu6 + sp87 * lu7 +
            sp87 * lu8;
    ra6 = zero;
    if constexpr (!seven) {
      ra6 = A1 * du6;  // lu2 = |U| * du6, sp62 = 1
      ra5 = ra1 * kk + Rho * (U * ra2 + V * ra3 + W * ra4) + ra5 / hg +
            c4 * (BX * ra6 + BY * ra7 + BZ * ra8);
    } else {
      ra5 = ra1 * kk + Rho * (U * ra2 + V * ra3 + W * ra4) + ra5 / hg +
            c4 * (BY * ra7 + BZ * ra8);
    }

    ra2 = U * ra1 + Rho * ra2;
    ra3 = V * ra1 + Rho * ra3;
    ra4 = W * ra1 + Rho * ra4;
  }

  const T fl1 = RUL, fr1 = RUR;
  const T fl2 = L.u * RUL - L.bx * L.bx * c4 + PL;
  const T fr2 = R.u * RUR - R.bx * R.bx * c4 + PR;
  const T fl3 = L.u * RVL - L.bx * L.by * c4, fr3 = R.u * RVR - R.bx * R.by * c4;
  const T fl4 = L.u * RWL - L.bx * L.bz * c4, fr4 = R.u * RWR - R.bx * R.bz * c4;
  const T fl5 = L.u * (EL + PL) - L.bx * (L.u * L.bx + L.v * L.by + L.w * L.bz) * c4;
  const T fr5 = R.u * (ER + PR) - R.bx * (R.u * R.bx + R.v * R.by + R.w * R.bz) * c4;
  const T fl7 = L.u * L.by - L.bx * L.v, fr7 = R.u * R.by - R.bx * R.v;
  const T fl8 = L.u * L.bz - L.bx * L.w, fr8 = R.u * R.bz - R.bx * R.w;
  out[0] = half * (fl1 + fr1 + ra1);
  out[1] = half * (fl2 + fr2 + ra2);
  out[2] = half * (fl3 + fr3 + ra3);
  out[3] = half * (fl4 + fr4 + ra4);
  out[4] = half * (fl5 + fr5 + ra5);
  out[5] = seven ? zero : half * (zero + zero + ra6);
  out[6] = half * (fl7 + fr7 + ra7);
  out[7] = half * (fl8 + fr8 + ra8);
}

// Rusanov flux (RUSANOVF, src/RusanovF.ChF:4-197)
template <typename T>
__device__ __forceinline__ void rusanov_flux(const Consts& C,
                                             const MState<T>& L,
                                             const MState<T>& R, T* out) {
  const T half = T(0.5), one = T(1), two = T(2);
  const T g = T(C.gamma), hg = T(C.hgamma);
  const T c4 = T(k1_4PI), c8 = T(k1_8PI);
  const T RUL = L.rho * L.u, RVL = L.rho * L.v, RWL = L.rho * L.w;
  const T BxBL = c8 * (L.bx * L.bx + L.by * L.by + L.bz * L.bz);
  const T RUxUL = half * (RUL * L.u + RVL * L.v + RWL * L.w);
  const T PL = L.p + BxBL, EL = L.p / hg + RUxUL + BxBL;
  const T BXL4 = L.bx * c4;
  const T RUR = R.rho * R.u, RVR = R.rho * R.v, RWR = R.rho * R.w;
  const T BxBR = c8 * (R.bx * R.bx + R.by * R.by + R.bz * R.bz);
  const T RUxUR = half * (RUR * R.u + RVR * R.v + RWR * R.w);
  const T PR = R.p + BxBR, ER = R.p / hg + RUxUR + BxBR;
  const T BXR4 = R.bx * c4;
  T Rho, U, BX2, BxBH, A2;
  if (C.iAveraging == 1) {
    const T HL = (EL + PL) / L.rho, HR = (ER + PR) / R.rho;
    const T cL = one / (one + tsqrt(R.rho / L.rho));
    const T cR = one - cL;
    Rho = R.rho * cL + L.rho * cR;
    U = L.u * cL + R.u * cR;
    const T V = L.v * cL + R.v * cR, W = L.w * cL + R.w * cR;
    const T HAV = HL * cL + HR * cR;
    const T BX = R.bx * cL + L.bx * cR, BY = R.by * cL + L.by * cR,
            BZ = R.bz * cL + L.bz * cR;
    const T BXAV = half * (L.bx + R.bx);
    BX2 = BXAV * BXAV;
    const T BYZ2 = BY * BY + BZ * BZ;
    BxBH = c8 * (BX2 + BYZ2);
    const T kk = half * (U * U + V * V + W * W);
    const T dbx = R.bx - L.bx, dby = R.by - L.by, dbz = R.bz - L.bz;
    const T X = c8 * (dbx * dbx + dby * dby + dbz * dbz) * cL * cL / L.rho;
    A2 = hg * (HAV - kk - c4 * (BX * BXAV + BYZ2) / Rho) + (two - g) * X;
  } else {
    Rho = half * (L.rho + R.rho);
    U = half * (L.u + R.u);
    const T BX = half * (L.bx + R.bx), BY = half * (L.by + R.by),
            BZ = half * (L.bz + R.bz);
    const T PGas = half * (L.p + R.p);
    BX2 = BX * BX;
    BxBH = c8 * (BX2 + BY * BY + BZ * BZ);
    A2 = g * PGas / Rho;
  }
  const T bb = c4 * BX2 / Rho;
  const T szb = A2 + (BxBH + BxBH) / Rho;
  const T hh = tsqrt(nonneg(szb * szb - T(4) * A2 * bb));
  const T af = tsqrt(half * (szb + hh));
  const T em = fabs(U) + af;
  out[0] = half * (RUL + RUR + em * (L.rho - R.rho));
  out[1] = half * ((L.u * RUL - BXL4 * L.bx + PL) + (R.u * RUR - BXR4 * R.bx + PR) +
                   em * (RUL - RUR));
  out[2] = half * ((L.u * RVL - BXL4 * L.by) + (R.u * RVR - BXR4 * R.by) +
                   em * (RVL - RVR));
  out[3] = half * ((L.u * RWL - BXL4 * L.bz) + (R.u * RWR - BXR4 * R.bz) +
                   em * (RWL - RWR));
  out[4] = half * ((L.u * (EL + PL) - BXL4 * (L.u * L.bx + L.v * L.by + L.w * L.bz)) +
                   (R.u * (ER + PR) - BXR4 * (R.u * R.bx + R.v * R.by + R.w * R.bz)) +
                   em * (EL - ER));
  out[5] = half * (T(0) + T(0) + em * (L.bx - R.bx));
  out[6] = half * ((L.u * L.by - L.bx * L.v) + (R.u * R.by - R.bx * R.v) +
                   em * (L.by - R.by));
  out[7] = half * ((L.u * L.bz - L.bx * L.w) + (R.u * R.bz - R.bx * R.w) +
                   em * (L.bz - R.bz));
}

// HLLC flux of ideal MHD (S. Li, J. Comput. Phys. 203 (2005) 344-357), in
// the reference's Gaussian units (magnetic pressure B^2/8pi, Maxwell
// stress B B/4pi).  No counterpart in the reference's solvers.  Outer
// speeds S_L = min(u_L - c_fL, u_R - c_fR), S_R = max(u_L + c_fL,
// u_R + c_fR) (normal fast speeds); the star region takes the HLL state's
// field B* = B_hll and contact speed S_M = (rho u)_hll / rho_hll, the total
// pressure p_T* = rho_L (S_L - u_L)(S_M - u_L) + p_TL - (B_xL^2 - B_x*^2)/4pi,
// and on each side K
//   rho*_K = rho_K (S_K - u_K) / (S_K - S_M),   (rho u)*_K = rho*_K S_M,
//   (rho v_t)*_K = [rho_K v_tK (S_K - u_K) - (B_x* B_t* - B_xK B_tK)/4pi] / (S_K - S_M),
//   E*_K = [E_K (S_K - u_K) - p_TK u_K + p_T* S_M
//           + (B_xK (v_K . B_K) - B_x* (v* . B*))/4pi] / (S_K - S_M),
// v* = (S_M, v_hll, w_hll); F = F_K + S_K (U*_K - U_K) on the side of the
// contact that holds x/t = 0, F_L or F_R beyond the outer waves.  Inputs
// clamped to (smallp, smallr) like the other solvers'; out order as
// roe_flux's.
template <typename T>
__device__ __forceinline__ void hllc_flux(const Consts& C, const MState<T>& L,
                                          const MState<T>& R, T* out) {
  const T half = T(0.5), zero = T(0);
  const T g = T(C.gamma), hg = T(C.hgamma);
  const T c4 = T(k1_4PI), c8 = T(k1_8PI);
  // conserved states and physical fluxes of both sides
  T UL[8], UR[8], FL[8], FR[8];
  T ptL, ptR;
  {
    const T b2 = L.bx * L.bx + L.by * L.by + L.bz * L.bz;
    ptL = L.p + c8 * b2;
    const T E = L.p / hg + half * L.rho * (L.u * L.u + L.v * L.v + L.w * L.w) + c8 * b2;
    const T vb = L.u * L.bx + L.v * L.by + L.w * L.bz;
    UL[0] = L.rho; UL[1] = L.rho * L.u; UL[2] = L.rho * L.v; UL[3] = L.rho * L.w;
    UL[4] = E; UL[5] = L.bx; UL[6] = L.by; UL[7] = L.bz;
    FL[0] = UL[1];
    FL[1] = UL[1] * L.u + ptL - c4 * L.bx * L.bx;
    FL[2] = UL[1] * L.v - c4 * L.bx * L.by;
    FL[3] = UL[1] * L.w - c4 * L.bx * L.bz;
    FL[4] = (E + ptL) * L.u - c4 * L.bx * vb;
    FL[5] = zero;
    FL[6] = L.u * L.by - L.v * L.bx;
    FL[7] = L.u * L.bz - L.w * L.bx;
  }
  {
    const T b2 = R.bx * R.bx + R.by * R.by + R.bz * R.bz;
    ptR = R.p + c8 * b2;
    const T E = R.p / hg + half * R.rho * (R.u * R.u + R.v * R.v + R.w * R.w) + c8 * b2;
    const T vb = R.u * R.bx + R.v * R.by + R.w * R.bz;
    UR[0] = R.rho; UR[1] = R.rho * R.u; UR[2] = R.rho * R.v; UR[3] = R.rho * R.w;
    UR[4] = E; UR[5] = R.bx; UR[6] = R.by; UR[7] = R.bz;
    FR[0] = UR[1];
    FR[1] = UR[1] * R.u + ptR - c4 * R.bx * R.bx;
    FR[2] = UR[1] * R.v - c4 * R.bx * R.by;
    FR[3] = UR[1] * R.w - c4 * R.bx * R.bz;
    FR[4] = (E + ptR) * R.u - c4 * R.bx * vb;
    FR[5] = zero;
    FR[6] = R.u * R.by - R.v * R.bx;
    FR[7] = R.u * R.bz - R.w * R.bx;
  }
  // normal fast speeds: c_f^2 = (a^2 + b^2 + sqrt((a^2 + b^2)^2 - 4 a^2 b_n^2)) / 2
  auto cfast = [&](const MState<T>& S) {
    const T a2 = g * S.p / S.rho;
    const T b2 = c4 * (S.bx * S.bx + S.by * S.by + S.bz * S.bz) / S.rho;
    const T bn2 = c4 * S.bx * S.bx / S.rho;
    const T s = a2 + b2;
    return tsqrt(half * (s + tsqrt(tmax(zero, s * s - T(4) * a2 * bn2))));
  };
  const T cfL = cfast(L), cfR = cfast(R);
  const T SL = tmin(L.u - cfL, R.u - cfR), SR = tmax(L.u + cfL, R.u + cfR);
  if (SL >= zero) {
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = FL[c];
    return;
  }
  if (SR <= zero) {
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = FR[c];
    return;
  }
  // the HLL state (divisions, not reciprocals: the FP64 build gives the
  // oracle's bits)
  T H[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) H[c] = (SR * UR[c] - SL * UL[c] - (FR[c] - FL[c])) / (SR - SL);
  const T SM = H[1] / H[0];
  const T vh = H[2] / H[0], wh = H[3] / H[0];
  const T bx = H[5], by = H[6], bz = H[7];
  const T ptS = L.rho * (SL - L.u) * (SM - L.u) + ptL - c4 * (L.bx * L.bx - bx * bx);
  const bool left = SM >= zero;
  const MState<T>& K = left ? L : R;
  const T* UK = left ? UL : UR;
  const T* FK = left ? FL : FR;
  const T SK = left ? SL : SR, ptK = left ? ptL : ptR;
  const T sk = SK - K.u, den = SK - SM;
  T S[8];
  S[0] = K.rho * sk / den;
  S[1] = S[0] * SM;
  S[2] = (K.rho * K.v * sk - c4 * (bx * by - K.bx * K.by)) / den;
  S[3] = (K.rho * K.w * sk - c4 * (bx * bz - K.bx * K.bz)) / den;
  S[4] = (UK[4] * sk - ptK * K.u + ptS * SM +
          c4 * (K.bx * (K.u * K.bx + K.v * K.by + K.w * K.bz) - bx * (SM * bx + vh * by + wh * bz))) /
         den;
  S[5] = bx; S[6] = by; S[7] = bz;
#pragma unroll
  for (int c = 0; c < 8; ++c) out[c] = FK[c] + SK * (S[c] - UK[c]);
}

// van Leer flux-vector splitting for one neutral fluid (VANLEERF /
// VANLEERRS, src/VanLeerRSF.ChF:4-240).  Inputs rotated, already clamped:
// p, rho, un, ut1, ut2.  out: rho, mn, mt1, mt2, E.
template <typename T>
__device__ __forceinline__ void vanleer_flux(const Consts& C, T pL, T rL, T uL,
                                             T vL, T wL, T pR, T rR, T uR,
                                             T vR, T wR, T* out) {
  const T g = T(C.gamma), hg = T(C.hgamma), half = T(0.5), zero = T(0);
  T p1, p2, p3, p4, p5;  // F+ from the left state
  {
    const T RUxU = half * (rL * uL * uL + rL * vL * vL + rL * wL * wL);
    const T E = pL / hg + RUxU;
    const T a = tsqrt(g * pL / rL);
    if (uL <= -a) { p1 = p2 = p3 = p4 = p5 = zero; }
    else if (uL >= a) {
      p1 = rL * uL; p2 = p1 * uL + pL; p3 = p1 * vL; p4 = p1 * wL;
      p5 = uL * (E + pL);
    } else {
      const T UA = (-uL + (a + a)) / g;
      const T H = (E + pL) / rL;
      const T HU = uL + a;
      p1 = T(0.25) * rL * HU * HU / a;
      p2 = p1 * (uL + UA); p3 = p1 * vL; p4 = p1 * wL; p5 = p1 * H;
    }
  }
  T m1, m2, m3, m4, m5;  // F- from the right state
  {
    const T RUxU = half * (rR * uR * uR + rR * vR * vR + rR * wR * wR);
    const T E = pR / hg + RUxU;
    const T a = tsqrt(g * pR / rR);
    if (uR >= a) { m1 = m2 = m3 = m4 = m5 = zero; }
    else if (uR <= -a) {
      m1 = rR * uR; m2 = m1 * uR + pR; m3 = m1 * vR; m4 = m1 * wR;
      m5 = uR * (E + pR);
    } else {
      const T UA = (-uR - (a + a)) / g;
      const T H = (E + pR) / rR;
      const T HU = uR - a;
      m1 = -T(0.25) * rR * HU * HU / a;
      m2 = m1 * (uR + UA); m3 = m1 * vR; m4 = m1 * wR; m5 = m1 * H;
    }
  }
  out[0] = m1 + p1; out[1] = m2 + p2; out[2] = m3 + p3; out[3] = m4 + p4;
  out[4] = m5 + p5;
}

}  // namespace mfm
