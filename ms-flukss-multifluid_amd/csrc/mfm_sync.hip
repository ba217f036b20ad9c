// mfm_sync.hip -- coarse-fine synchronisation of a level pair on the device:
// the flux register of the coarse-fine interface (increments from the coarse
// and fine face fluxes, reflux into the coarse cells next to the fine
// region) and the average of the fine level down onto the coarse cells it
// covers.  FP64, no FMA contraction.
#include "mfm_fill.hpp"

namespace mfm {
namespace {
constexpr int kSB = 256;

__device__ __forceinline__ long off3(const FillFab& F, int i, int j, int k) {
  return (long)(i - F.lo[0]) +
         (long)F.n[0] * ((long)(j - F.lo[1]) + (long)F.n[1] * (long)(k - F.lo[2]));
}

// Register values are stored component-major: reg[c * ne + e].
__global__ __launch_bounds__(kSB) void k_reg_zero(double* reg, long n) {
  const long t = (long)blockIdx.x * kSB + threadIdx.x;
  if (t < n) reg[t] = 0.0;
}

// reg -= scale * F_coarse at the face (the coarse side of the register)
__global__ __launch_bounds__(kSB) void k_reg_coarse(const RegFace* __restrict__ E, long ne,
                                                   const FillFab* __restrict__ F, int dims,
                                                   double* __restrict__ reg, int nc,
                                                   double scale) {
  const long t = (long)blockIdx.x * kSB + threadIdx.x;
  if (t >= ne) return;
  const RegFace e = E[t];
  const FillFab f = F[e.cb * dims + e.d];
  int fc[3] = {e.c[0], e.c[1], e.c[2]};
  if (e.sigma > 0) fc[e.d] += 1;
  const double* p = f.p + off3(f, fc[0], fc[1], fc[2]);
  for (int c = 0; c < nc; ++c) {
    const double x = scale * p[c * f.cs];
    reg[c * ne + t] = reg[c * ne + t] - x;
  }
}

// reg += (scale / r^(dims-1)) * F_fine over the fine faces covering the
// coarse face, the first transverse direction fastest.
__global__ __launch_bounds__(kSB) void k_reg_fine(const RegFace* __restrict__ E, long ne,
                                                 const FillFab* __restrict__ F, int dims,
                                                 int r, double* __restrict__ reg, int nc,
                                                 double scale) {
  const long t = (long)blockIdx.x * kSB + threadIdx.x;
  if (t >= ne) return;
  const RegFace e = E[t];
  const FillFab f = F[e.fb * dims + e.d];
  int tr[2] = {-1, -1}, nt = 0;
  for (int d = 0; d < dims; ++d)
    if (d != e.d) tr[nt++] = d;
  double denom = 1.0;
  for (int q = 0; q < dims - 1; ++q) denom *= r;
  const double sc = scale / denom;
  const int n2 = nt > 1 ? r : 1;
  for (int c = 0; c < nc; ++c) {
    double acc = reg[c * ne + t];
    for (int q2 = 0; q2 < n2; ++q2)
      for (int q1 = 0; q1 < r; ++q1) {
        int fi[3] = {e.f0[0], e.f0[1], e.f0[2]};
        fi[tr[0]] += q1;
        if (nt > 1) fi[tr[1]] += q2;
        const double x = sc * f.p[off3(f, fi[0], fi[1], fi[2]) + c * f.cs];
        acc = acc + x;
      }
    reg[c * ne + t] = acc;
  }
}

// U += (scale * sigma * invvol) * reg for each register face of the cell,
// in the face list order (direction, then low face before high face).
__global__ __launch_bounds__(kSB) void k_reflux(const RegCell* __restrict__ C, long ncell,
                                               const RegFace* __restrict__ E, long ne,
                                               const FillFab* __restrict__ U,
                                               const double* __restrict__ reg, int nc,
                                               double scale, double invvol) {
  const long t = (long)blockIdx.x * kSB + threadIdx.x;
  if (t >= ncell) return;
  const RegCell g = C[t];
  const RegFace e0 = E[g.e0];
  const FillFab u = U[e0.cb];
  double* p = u.p + off3(u, e0.c[0], e0.c[1], e0.c[2]);
  for (int c = 0; c < nc; ++c) {
    double v = p[c * u.cs];
    for (int e = g.e0; e < g.e1; ++e) {
      const double f = scale * (double)E[e].sigma * invvol;
      v = v + f * reg[c * ne + e];
    }
    p[c * u.cs] = v;
  }
}

// averagewithvolumes (src/CoarseAverageExtF.ChF:10-50) into the coarse cell
__global__ __launch_bounds__(kSB) void k_average_down(const AvgRect* __restrict__ R, int nr,
                                                     const FillFab* __restrict__ Uc,
                                                     const FillFab* __restrict__ Uf, int nc,
                                                     int dims, int r, double vol) {
  const int bx = blockIdx.x;
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (R[mid].blk0 <= bx) lo = mid; else hi = mid - 1;
  }
  const AvgRect a = R[lo];
  const long t = (long)(bx - a.blk0) * kSB + threadIdx.x;
  const long n01 = (long)a.n[0] * a.n[1];
  if (t >= n01 * a.n[2]) return;
  const int k = (int)(t / n01);
  const long rem = t - (long)k * n01;
  const int j = (int)(rem / a.n[0]);
  const int i = (int)(rem - (long)j * a.n[0]);
  const int ic[3] = {a.lo[0] + i, a.lo[1] + j, a.lo[2] + k};
  const FillFab uc = Uc[a.cb], uf = Uf[a.fb];
  const int rz = dims == 3 ? r : 1;
  const int fz0 = dims == 3 ? ic[2] * r : ic[2];
  for (int c = 0; c < nc; ++c) {
    double sum = 0.0, volc = 0.0;
    for (int kk = 0; kk < rz; ++kk)
      for (int jj = 0; jj < r; ++jj)
        for (int ii = 0; ii < r; ++ii) {
          const double x = vol * uf.p[off3(uf, ic[0] * r + ii, ic[1] * r + jj, fz0 + kk) + c * uf.cs];
          sum = sum + x;
          volc = volc + vol;
        }
    uc.p[off3(uc, ic[0], ic[1], ic[2]) + c * uc.cs] = sum / volc;
  }
}
}  // namespace

hipError_t launch_reg_zero(double* reg, long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reg_zero, dim3((unsigned)((n + kSB - 1) / kSB)), dim3(kSB), 0, s, reg, n);
  return hipGetLastError();
}

hipError_t launch_reg_coarse(const RegFace* E, long ne, const FillFab* F, int dims,
                             double* reg, int nc, double scale, hipStream_t s) {
  if (ne <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reg_coarse, dim3((unsigned)((ne + kSB - 1) / kSB)), dim3(kSB), 0, s,
                     E, ne, F, dims, reg, nc, scale);
  return hipGetLastError();
}

hipError_t launch_reg_fine(const RegFace* E, long ne, const FillFab* F, int dims, int r,
                           double* reg, int nc, double scale, hipStream_t s) {
  if (ne <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reg_fine, dim3((unsigned)((ne + kSB - 1) / kSB)), dim3(kSB), 0, s,
                     E, ne, F, dims, r, reg, nc, scale);
  return hipGetLastError();
}

hipError_t launch_reflux(const RegCell* C, long ncell, const RegFace* E, long ne,
                         const FillFab* U, const double* reg, int nc, double scale,
                         double invvol, hipStream_t s) {
  if (ncell <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_reflux, dim3((unsigned)((ncell + kSB - 1) / kSB)), dim3(kSB), 0, s,
                     C, ncell, E, ne, U, reg, nc, scale, invvol);
  return hipGetLastError();
}

hipError_t launch_average_down(const AvgRect* R, int nr, int nblk, const FillFab* Uc,
                               const FillFab* Uf, int nc, int dims, int r, double vol,
                               hipStream_t s) {
  if (nr <= 0 || nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_average_down, dim3((unsigned)nblk), dim3(kSB), 0, s, R, nr, Uc, Uf,
                     nc, dims, r, vol);
  return hipGetLastError();
}

}  // namespace mfm
