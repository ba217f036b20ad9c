// mfm_fill.hip -- device ghost fill of a level: rectangle copies (the
// same-level exchange, and the coarse data gathered under the fine ghost
// cells with its time interpolation) and the piecewise-linear coarse-fine
// interpolation of PiecewiseLinearFillPatchMHDAM.  Everything is FP64 and
// HBM-bound; built without FMA contraction so that every value is rounded
// as in the reference's separate multiply and add.
#include "mfm_fill.hpp"

namespace mfm {
namespace {
constexpr int kFB = 256;

__device__ __forceinline__ long fab_off(const FillFab& F, int i, int j, int k) {
  return (long)(i - F.lo[0]) +
         (long)F.n[0] * ((long)(j - F.lo[1]) + (long)F.n[1] * (long)(k - F.lo[2]));
}

// Copy of the nc components of one cell, in batches of 8 loads issued
// before their stores so that every lane keeps several loads in flight.
// mode 0: dst = new; 1: dst = old; 2: dst = new*a + old*b (b = 1 - a), the
// order of timeInterp (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:489-496).
__device__ __forceinline__ void copy_cell(double* dp, long dcs, const double* sp,
                                          long scs, const double* op, long ocs,
                                          int nc, double a, double b, int mode) {
  constexpr int kBatch = 8;
  double v[kBatch], w[kBatch];
  for (int c0 = 0; c0 < nc; c0 += kBatch) {
    const int m = nc - c0 < kBatch ? nc - c0 : kBatch;
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      if (q < m) v[q] = (mode == 1 ? op[(c0 + q) * ocs] : sp[(c0 + q) * scs]);
    if (mode == 2) {
#pragma unroll
      for (int q = 0; q < kBatch; ++q)
        if (q < m) w[q] = op[(c0 + q) * ocs];
#pragma unroll
      for (int q = 0; q < kBatch; ++q)
        if (q < m) {
          const double x = v[q] * a;
          const double y = w[q] * b;
          v[q] = x + y;
        }
    }
#pragma unroll
    for (int q = 0; q < kBatch; ++q)
      if (q < m) dp[(c0 + q) * dcs] = v[q];
  }
}

__global__ __launch_bounds__(kFB) void k_rect_copy(
    const CopyRect* __restrict__ R, int nr, const FillFab* __restrict__ Sn,
    const FillFab* __restrict__ So, const FillFab* __restrict__ D, int nc,
    double a, double b, int mode) {
  const int bx = blockIdx.x;
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (R[mid].blk0 <= bx) lo = mid; else hi = mid - 1;
  }
  const CopyRect r = R[lo];
  const long t = (long)(bx - r.blk0) * kFB + threadIdx.x;
  const long n01 = (long)r.n[0] * r.n[1];
  if (t >= n01 * r.n[2]) return;
  const int k = (int)(t / n01);
  const long rem = t - (long)k * n01;
  const int j = (int)(rem / r.n[0]);
  const int i = (int)(rem - (long)j * r.n[0]);
  const FillFab d = D[r.dst];
  double* dp = d.p + fab_off(d, r.dlo[0] + i, r.dlo[1] + j, r.dlo[2] + k);
  const int si = r.slo[0] + i, sj = r.slo[1] + j, sk = r.slo[2] + k;
  const double* sp = nullptr;
  const double* op = nullptr;
  long scs = 0, ocs = 0;
  if (mode != 1) { const FillFab s = Sn[r.src]; sp = s.p + fab_off(s, si, sj, sk); scs = s.cs; }
  if (mode != 0) { const FillFab o = So[r.src]; op = o.p + fab_off(o, si, sj, sk); ocs = o.cs; }
  copy_cell(dp, d.cs, sp, scs, op, ocs, nc, a, b, mode);
}

// The same three modes for a list of single cells (the coarse cells the
// interpolation reads, gathered into the work FABs).
__global__ __launch_bounds__(kFB) void k_cell_copy(
    const GatherCell* __restrict__ G, long n, const FillFab* __restrict__ Sn,
    const FillFab* __restrict__ So, const FillFab* __restrict__ D, int nc,
    double a, double b, int mode) {
  const long t = (long)blockIdx.x * kFB + threadIdx.x;
  if (t >= n) return;
  const GatherCell g = G[t];
  const FillFab d = D[g.dst];
  const double* sp = nullptr;
  const double* op = nullptr;
  long scs = 0, ocs = 0;
  if (mode != 1) { const FillFab s = Sn[g.src]; sp = s.p + fab_off(s, g.s[0], g.s[1], g.s[2]); scs = s.cs; }
  if (mode != 0) { const FillFab o = So[g.src]; op = o.p + fab_off(o, g.s[0], g.s[1], g.s[2]); ocs = o.cs; }
  copy_cell(d.p + g.doff, d.cs, sp, scs, op, ocs, nc, a, b, mode);
}

// minmod of src/reconstructionF.ChF:4-14: sign(1,A) max(0, min(|A|, sign(1,A) B))
__device__ __forceinline__ double minmod_ab(double A, double B) {
  const double s = copysign(1.0, A);
  return s * fmax(0.0, fmin(fabs(A), s * B));
}

// interp_divB_0 (srcIdeal/FineInterpMHDAMF.ChF:449-555) on the slopes of
// (Bx, By, Bz) of one coarse cell: s[d] is d(B_d)/dx_d.
__device__ void divb_fix(double s[3], int dims) {
  const double eps = 1.0e-8;
  const double dbx = s[0], dby = s[1], dbz = dims == 3 ? s[2] : 0.0;
  const double fact = -1.0 / (dims == 3 ? 3.0 : 2.0);
  const double divb = dims == 3 ? dbx + dby + dbz : dbx + dby;
  double del = fact * divb;
  double adel = fabs(del);
  if (!(adel > eps)) return;
  double dx = minmod_ab(-dbx, del), dy = minmod_ab(-dby, del);
  double dz = dims == 3 ? minmod_ab(-dbz, del) : 0.0;
  double del2 = dims == 3 ? dx + dy + dz : dx + dy;
  if (fabs(del2) < fabs(divb)) {
    if (dims == 2) {
      dx = minmod_ab(dbx, -dby);
      dy = -dx;
      dx = dx - dbx;
      dy = dy - dby;
    } else if (fabs(dx) < adel) {
      del = -0.5 * (dx + divb);
      dy = minmod_ab(-dby, del);
      dz = minmod_ab(-dbz, del);
      del2 = dx + dy + dz;
      if (fabs(del2) < fabs(divb)) {
        if (fabs(dy) < adel) dz = -(divb + dx + dy);
        else dy = -(divb + dx + dz);
      }
    } else if (fabs(dy) < adel) {
      del = -0.5 * (dy + divb);
      dx = minmod_ab(-dbx, del);
      dz = minmod_ab(-dbz, del);
      del2 = dx + dy + dz;
      if (fabs(del2) < fabs(divb)) {
        if (fabs(dx) < adel) dz = -(divb + dx + dy);
        else dx = -(divb + dy + dz);
      }
    } else {
      del = -0.5 * (dz + divb);
      adel = fabs(del);
      dx = minmod_ab(-dbx, del);
      dy = minmod_ab(-dby, del);
      del2 = dx + dy + dz;
      if (fabs(del2) < fabs(divb)) {
        if (fabs(dx) < adel) dy = -(divb + dx + dz);
        else dx = -(divb + dy + dz);
      }
    }
  }
  s[0] = s[0] + dx;
  s[1] = s[1] + dy;
  if (dims == 3) s[2] = s[2] + dz;
}

struct CellCtx {
  const double* w;   // work FAB, component 0, at the cell
  long wcs;
  int st[3];         // strides of the work FAB
  int dims;
  uint8_t fl;
};

// Slopes of computeSlopes (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:591-
// 760, constant spacing) and interplimit_v (FineInterpMHDAMF.ChF:235-385).
__device__ __forceinline__ double slopes(const CellCtx& X, const InterpArgs& A,
                                         int c, double s[3]) {
  const double* w = X.w + c * X.wcs;
  const double w0 = w[0];
  s[2] = 0.0;
  for (int d = 0; d < X.dims; ++d) {
    const int kind = (X.fl >> (2 * d)) & 3;
    const int st = X.st[d];
    if (kind == kSlopeCentred) s[d] = A.c_centre * (w[st] - w[-st]);
    else if (kind == kSlopeLo) s[d] = A.c_side * (w0 - w[-st]);
    else s[d] = A.c_side * (w[st] - w0);
  }
  if (X.fl & kFillLimit) {
    double smax = w0, smin = w0;
    const int kz = X.dims == 3 ? 1 : 0;
    for (int kk = -kz; kk <= kz; ++kk)
      for (int jj = -1; jj <= 1; ++jj)
        for (int ii = -1; ii <= 1; ++ii) {
          const double v = w[ii + jj * X.st[1] + kk * X.st[2]];
          smax = fmax(smax, v);
          smin = fmin(smin, v);
        }
    double dsum = fabs(s[0] * A.dxc) + fabs(s[1] * A.dxc);
    if (X.dims == 3) dsum = dsum + fabs(s[2] * A.dxc);
    dsum = 0.5 * dsum;
    double eta = fmin(smax - w0, w0 - smin);
    if (eta <= A.tiny * fabs(smax)) eta = 0.0;
    else if (dsum > eta) eta = eta / dsum;
    else eta = 1.0;
    for (int d = 0; d < X.dims; ++d) s[d] = eta * s[d];
  }
  return w0;
}

// fine = coarse value, then + coeff (xf - xc)_d slope_d per direction
// (fillConstantInterp :537-585 and incrementLinearInterp :800-880).
__device__ __forceinline__ void write_children(const InterpArgs& A,
                                               const FillFab& F, const int f0[3],
                                               const int f1[3], const int ci[3],
                                               int c, double w0, const double s[3]) {
  double xc[3];
  for (int d = 0; d < 3; ++d) xc[d] = A.dxc * ((double)ci[d] + 0.5);
  double* base = F.p + c * F.cs;
  for (int k = f0[2]; k <= f1[2]; ++k) {
    const double tz = A.dims == 3 ? A.coeff * (A.dxf * ((double)k + 0.5) - xc[2]) : 0.0;
    for (int j = f0[1]; j <= f1[1]; ++j) {
      const double ty = A.coeff * (A.dxf * ((double)j + 0.5) - xc[1]);
      for (int i = f0[0]; i <= f1[0]; ++i) {
        const double tx = A.coeff * (A.dxf * ((double)i + 0.5) - xc[0]);
        double v = w0 + tx * s[0];
        v = v + ty * s[1];
        if (A.dims == 3) v = v + tz * s[2];
        base[fab_off(F, i, j, k)] = v;
      }
    }
  }
}

__global__ __launch_bounds__(kFB) void k_fill_interp(InterpArgs A) {
  const long t = (long)blockIdx.x * kFB + threadIdx.x;
  if (t >= A.ncell) return;
  const int2 e = A.cell[t];
  const FillPatch P = A.P[e.x];
  const FillFab F = A.U[e.x];
  const int wo = e.y;
  const int n01 = P.wn[0] * P.wn[1];
  const int ck = wo / n01, cj = (wo - ck * n01) / P.wn[0];
  const int ci0 = wo - ck * n01 - cj * P.wn[0];
  const int ci[3] = {P.wlo[0] + ci0, P.wlo[1] + cj, P.wlo[2] + ck};
  int f0[3], f1[3];
  for (int d = 0; d < 3; ++d) {
    if (d < A.dims) {
      f0[d] = max(ci[d] * A.r, P.glo[d]);
      f1[d] = min(ci[d] * A.r + A.r - 1, P.ghi[d]);
    } else {
      f0[d] = f1[d] = ci[d];
    }
  }
  CellCtx X;
  X.w = A.work + P.woff + wo;
  X.wcs = (long)n01 * P.wn[2];
  X.st[0] = 1; X.st[1] = P.wn[0]; X.st[2] = n01;
  X.dims = A.dims;
  X.fl = A.flag[t];
  for (int c = 0; c < A.nc; ++c) {
    if (A.ibx >= 0 && c >= A.ibx && c < A.ibx + A.dims) continue;
    double s[3];
    const double w0 = slopes(X, A, c, s);
    write_children(A, F, f0, f1, ci, c, w0, s);
  }
  if (A.ibx >= 0) {
    // the slopes of B_d along d are corrected together, then each component
    // is written with its own three slopes
    double sb[3][3], wb[3];
    for (int q = 0; q < A.dims; ++q) wb[q] = slopes(X, A, A.ibx + q, sb[q]);
    if (X.fl & kFillLimit) {
      double dd[3] = {sb[0][0], sb[1][1], A.dims == 3 ? sb[2][2] : 0.0};
      divb_fix(dd, A.dims);
      for (int q = 0; q < A.dims; ++q) sb[q][q] = dd[q];
    }
    for (int q = 0; q < A.dims; ++q)
      write_children(A, F, f0, f1, ci, A.ibx + q, wb[q], sb[q]);
  }
}
}  // namespace

hipError_t launch_rect_copy(const CopyRect* R, int nr, int nblk,
                            const FillFab* src_new, const FillFab* src_old,
                            const FillFab* dst, int nc, double a, double b,
                            int mode, hipStream_t s) {
  if (nr <= 0 || nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_rect_copy, dim3((unsigned)nblk), dim3(kFB), 0, s, R, nr,
                     src_new, src_old, dst, nc, a, b, mode);
  return hipGetLastError();
}

hipError_t launch_cell_copy(const GatherCell* G, long n, const FillFab* src_new,
                            const FillFab* src_old, const FillFab* dst, int nc,
                            double a, double b, int mode, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_copy, dim3((unsigned)((n + kFB - 1) / kFB)), dim3(kFB), 0, s,
                     G, n, src_new, src_old, dst, nc, a, b, mode);
  return hipGetLastError();
}

hipError_t launch_fill_interp(const InterpArgs& A, hipStream_t s) {
  if (A.ncell <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_interp, dim3((unsigned)((A.ncell + kFB - 1) / kFB)),
                     dim3(kFB), 0, s, A);
  return hipGetLastError();
}

}  // namespace mfm
