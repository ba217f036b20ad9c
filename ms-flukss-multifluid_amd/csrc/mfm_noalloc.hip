// mfm_noalloc.hip -- the allocation-free patch update.
//
// mfm_patch_update_noalloc_dev runs exactly the launch sequence of
// mfm_patch_update_dev, but every byte of device scratch (first-stage
// state, primitives, face fluxes, tables, CFL partials) lives in one
// file-scope device array, `scratch`, and the staged tables in a buffer
// inside the context.  The call itself therefore makes no device or pinned
// allocation; it is meant for callers that cannot allocate (code under graph
// capture, tight serving loops).
//
// Each GPU holds its own copy of `scratch` (at its own address) for the life
// of the process.  Its address on a device is resolved once, with that device
// current, when the first context is created there; host code passes only
// that resolved address to kernels and copies, never the array's name.
// All calls on one GPU share the array and it keeps the last call's
// contents: calls that overlap on one GPU race.  Nothing in it has to start
// at zero (every kernel writes what it later reads within the call).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/mfm.h"
#include "mfm_host.hpp"

namespace {

__device__ __align__(256) unsigned char scratch[MFM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};  // resolved address per device

}  // namespace

namespace mfm {
namespace host {

int fixed_scratch_setup(mfm_ctx* c) {
  const int dev = c->device;
  if (dev < 0 || dev >= kMaxDevices) return fail(c, MFM_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[dev]) return MFM_OK;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return hipfail(c, e, "hipGetDevice");
  if (cur != dev) return fail(c, MFM_ERR_DEVICE, "context device is not the current device");
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hipfail(c, e, "hipGetSymbolAddress(scratch)");
  g_scratch[dev] = ptr;
  return MFM_OK;
}

}  // namespace host
}  // namespace mfm

extern "C" int mfm_patch_update_noalloc_dev(mfm_ctx* c, mfm_fab U, mfm_box valid,
                                            double dt, double dx, double time,
                                            mfm_fab* F, mfm_fab* divB, double* new_dt,
                                            int min_cell[3], void* stream) {
  using namespace mfm::host;
  if (!c) return MFM_ERR_ARG;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != c->device)
    return fail(c, MFM_ERR_DEVICE, "the current device is not the context's device");
  void* ptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    ptr = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
  }
  if (!ptr) return fail(c, MFM_ERR_DEVICE, "fixed scratch not resolved on this device");
  FixedScratch fx{static_cast<char*>(ptr), (size_t)MFM_NOALLOC_SCRATCH_BYTES, c->fixed_host,
                  mfm_ctx::kFixedHostBytes};
  PatchIn in{U, valid, F, divB, 0, dt, dx, time};
  double d = 0.0;
  int cell[3] = {0, 0, 0};
  const int rc = run_batch(c, &c->C, 1, &in, 1, &d, cell, nullptr, nullptr, (hipStream_t)stream, &fx);
  if (new_dt) *new_dt = d;
  if (min_cell) { min_cell[0] = cell[0]; min_cell[1] = cell[1]; min_cell[2] = cell[2]; }
  return rc;
}
