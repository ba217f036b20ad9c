// mfm_ceiling.hip -- HBM ceilings measured with the library's own streaming
// kernels (read, copy, triad), the denominators the patch-update kernels
// are compared against besides the 8 TB/s of the data sheet.  No
// counterpart in the reference (a CPU code).
//
// Each thread moves 16-byte vectors (double2), four per loop trip so that
// four loads are in flight before the first use; the grid is a whole number
// of waves over the 256 CUs (8 blocks of 256 threads per CU), walking the
// arrays with a grid stride, so every XCD streams its own interleaved share.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mfm.h"

namespace {

constexpr int kThreads = 256;
constexpr int kUnroll = 4;

__global__ __launch_bounds__(kThreads) void k_read(const double2* __restrict__ a, size_t n,
                                                   double* __restrict__ part) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  double s0 = 0.0, s1 = 0.0;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double2 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) { s0 += v[u].x; s1 += v[u].y; }
  }
  for (; i < n; i += stride) { s0 += a[i].x; s1 += a[i].y; }
  // one value per thread keeps every load live; 8 B per thread is noise
  part[(size_t)blockIdx.x * kThreads + threadIdx.x] = s0 + s1;
}

__global__ __launch_bounds__(kThreads) void k_copy(const double2* __restrict__ a,
                                                   double2* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double2 v[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] = a[i + u * stride];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) b[i + u * stride] = v[u];
  }
  for (; i < n; i += stride) b[i] = a[i];
}

__global__ __launch_bounds__(kThreads) void k_triad(double2* __restrict__ a,
                                                    const double2* __restrict__ b,
                                                    const double2* __restrict__ c,
                                                    double s, size_t n) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double2 x[kUnroll], y[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) { x[u] = b[i + u * stride]; y[u] = c[i + u * stride]; }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      double2 r;
      r.x = x[u].x + s * y[u].x;
      r.y = x[u].y + s * y[u].y;
      a[i + u * stride] = r;
    }
  }
  for (; i < n; i += stride) {
    double2 r;
    r.x = b[i].x + s * c[i].x;
    r.y = b[i].y + s * c[i].y;
    a[i] = r;
  }
}

__global__ __launch_bounds__(kThreads) void k_fill(double2* __restrict__ a, size_t n, double v) {
  const size_t stride = (size_t)gridDim.x * kThreads;
  for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride)
    a[i] = make_double2(v, v);
}

}  // namespace

extern "C" int mfm_hbm_ceiling(int kind, size_t bytes, int reps, double* gbps,
                               double* best_ms) {
  if (kind < MFM_CEIL_READ || kind > MFM_CEIL_TRIAD || reps < 1 || !gbps ||
      bytes < (1u << 20) || bytes % 16 != 0)
    return MFM_ERR_ARG;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return MFM_ERR_DEVICE;
  const size_t n = bytes / 16;  // double2 elements per array
  const int blocks = 8 * (ncu > 0 ? ncu : 256);
  const int narr = kind == MFM_CEIL_TRIAD ? 3 : (kind == MFM_CEIL_COPY ? 2 : 1);
  double2* arr[3] = {nullptr, nullptr, nullptr};
  double* part = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MFM_OK;
  for (int q = 0; q < narr && rc == MFM_OK; ++q)
    if (hipMalloc((void**)&arr[q], bytes) != hipSuccess) rc = MFM_ERR_NOMEM;
  if (rc == MFM_OK &&
      hipMalloc((void**)&part, (size_t)blocks * kThreads * sizeof(double)) != hipSuccess)
    rc = MFM_ERR_NOMEM;
  if (rc == MFM_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = MFM_ERR_DEVICE;
  if (rc == MFM_OK) {
    for (int q = 0; q < narr; ++q)
      hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(kThreads), 0, 0, arr[q], n, 1.0 + q);
    double best = 1e300;
    for (int r = 0; r <= reps && rc == MFM_OK; ++r) {  // launch 0 warms up
      if (hipEventRecord(e0, 0) != hipSuccess) { rc = MFM_ERR_DEVICE; break; }
      if (kind == MFM_CEIL_READ)
        hipLaunchKernelGGL(k_read, dim3(blocks), dim3(kThreads), 0, 0, arr[0], n, part);
      else if (kind == MFM_CEIL_COPY)
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(kThreads), 0, 0, arr[0], arr[1], n);
      else
        hipLaunchKernelGGL(k_triad, dim3(blocks), dim3(kThreads), 0, 0, arr[0], arr[1],
                           arr[2], 0.5, n);
      float ms = 0.0f;
      if (hipEventRecord(e1, 0) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = MFM_ERR_DEVICE;
        break;
      }
      if (r > 0 && ms < best) best = ms;
    }
    if (rc == MFM_OK) {
      *gbps = (double)narr * (double)bytes / (best * 1e-3) / 1e9;
      if (best_ms) *best_ms = best;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (int q = 0; q < 3; ++q)
    if (arr[q]) (void)hipFree(arr[q]);
  if (part) (void)hipFree(part);
  return rc;
}
