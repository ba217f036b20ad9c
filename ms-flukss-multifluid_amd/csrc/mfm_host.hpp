// mfm_host.hpp -- host-side internals shared by the C ABI units: boxes,
// the context, constants construction and the batched update driver.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/mfm.h"
#include "mfm_device.hpp"
#include "mfm_kernels.hpp"

struct mfm_ctx {
  mfm_params p;
  mfm::Consts C;
  int device = 0;
  std::string err = "";
  // device arena for scratch + tables
  char* arena = nullptr;
  size_t arena_bytes = 0;
  // pinned host staging for tables and results
  char* hbuf = nullptr;
  size_t hbuf_bytes = 0;
  size_t fsize() const { return p.precision == MFM_PREC_FP32_FLUX ? 4 : 8; }
  // optional per-kernel timing, accumulated per stage and kernel kind
  mfm::KernelTimer* timer = nullptr;
  double kt_ms[2][mfm::KK_COUNT] = {};
  long kt_n[2][mfm::KK_COUNT] = {};
  double kt_items[2][mfm::KK_COUNT] = {};
  // optional hipGraph replay of the launch sequence (not with the timer)
  bool graphs = false;
  hipStream_t cap_stream = nullptr;   // private stream used for capture
  hipGraphExec_t gexec = nullptr;
  std::vector<long long> gkey;        // what the captured sequence depends on
  // work walk of the region kernels (row tile in (y, z), XCD renumbering)
  int walk_tile = 8;
  int walk_xcd = 1;
  // kernel structure of RK2 / first-order stages (MFM_STAGE_*): separate
  // per-direction flux kernels, the fused stage kernel (k_stage), the
  // face-sum tile kernels (k_fsum) before k_update, or the choice per batch
  // between separate and summed (AUTO); the tile kernels walk z-chunks of
  // stage_tz planes per block (8: 12.01 / 11.75 against 12.03 / 12.07 ms per
  // 256^3 step for 16, 12.26 for 12, 12.10 for 4, 12.56 for 32)
  int stage_mode = MFM_STAGE_AUTO;
  int stage_tz = 8;
  // host staging of the allocation-free entry point (tables and results of
  // one patch); part of the context, so that entry point needs no allocation
  static constexpr size_t kFixedHostBytes = 64 * 1024;
  alignas(256) char fixed_host[kFixedHostBytes];
};

namespace mfm {
namespace host {

// Inclusive cell box.
struct HBox {
  int lo[3], hi[3];
  long n(int d) const { return (long)hi[d] - lo[d] + 1; }
  long size() const {
    for (int d = 0; d < 3; ++d)
      if (hi[d] < lo[d]) return 0;
    return n(0) * n(1) * n(2);
  }
};
inline HBox hb(const mfm_box& b) {
  HBox r;
  for (int d = 0; d < 3; ++d) { r.lo[d] = b.lo[d]; r.hi[d] = b.hi[d]; }
  return r;
}
inline HBox grow(HBox b, int g, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] -= g; b.hi[d] += g; }
  return b;
}
inline bool contains(const HBox& a, const HBox& b) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < a.lo[d] || b.hi[d] > a.hi[d]) return false;
  return true;
}
inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
// Cells per FAB component the update kernels take: their 32-bit cell
// offsets, and the 32-bit byte offsets of their 4-byte arrays, cannot wrap.
constexpr long kMaxFabCells = 1L << 30;

int fail(mfm_ctx* c, int code, const std::string& m);
int hipfail(mfm_ctx* c, hipError_t e, const char* what);
int build_consts(const mfm_params* p, Consts& C, std::string& why);

// The plasma solver family of p (Consts::flux_kind): the kernels are
// compiled per family, and one launch serves one family.  0: 8-wave Roe
// (flux 0, 1, 3), 1: Rusanov (2, 4), 2: 7-wave Roe (11, 13, and GLM with
// 1 or 3), 3: HLLC (21, with or without GLM).
inline int flux_kind_of(const mfm_params& p) {
  if (p.flux == MFM_FLUX_HLLC) return 3;
  if (p.dedner || p.flux == 11 || p.flux == 13) return 2;
  return (p.flux == 2 || p.flux == 4) ? 1 : 0;
}

// Caller-provided memory for run_batch instead of the context's growing
// arena and staging (the allocation-free entry point): dev holds the patch
// scratch, tables and reductions; host the staged tables and results.
struct FixedScratch {
  char* dev;
  size_t dev_bytes;
  char* host;
  size_t host_bytes;
};

// A vector with N elements of inline storage: lists of one patch's update
// fit inline, so building them touches no heap.
template <typename T, int N>
struct SmallVec {
  T inl[N];
  std::vector<T> ext;
  int n = 0;
  void push_back(const T& x) {
    if (n < N) {
      inl[n] = x;
    } else {
      if (n == N) ext.assign(inl, inl + N);
      ext.push_back(x);
    }
    ++n;
  }
  void resize(int m) {
    if (m > N) ext.resize(m);
    n = m;
  }
  int size() const { return n; }
  T* data() { return n > N ? ext.data() : inl; }
  const T* data() const { return n > N ? ext.data() : inl; }
  T& operator[](int i) { return data()[i]; }
  const T& operator[](int i) const { return data()[i]; }
};

// One patch of a batched update.
struct PatchIn {
  mfm_fab U;
  mfm_box valid;
  mfm_fab* F;
  mfm_fab* divB;
  int cidx;                // row of the constants table
  double dt, dx, time;
};

int run_batch(mfm_ctx* c, const Consts* Ct, int nct, const PatchIn* in, int n,
              double* new_dt, int* min_cell, double* level_dt, int* min_patch,
              hipStream_t s, const FixedScratch* fixed = nullptr);
// Resolves, once per device, the address of the fixed scratch of the
// allocation-free entry point on the context's device (mfm_noalloc.hip).
int fixed_scratch_setup(mfm_ctx* c);

// computeDt of one box with the constants Cq.
int run_dt(mfm_ctx* c, const Consts& Cq, const mfm_fab& U, const mfm_box& box,
           double dx, double* dt, int* min_cell, hipStream_t s);

}  // namespace host
}  // namespace mfm
