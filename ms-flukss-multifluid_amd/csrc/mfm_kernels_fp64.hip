// FP64 (parity) build of the stage kernels.  Compiled with
// -ffp-contract=off so that every product and sum rounds as in the
// reference's Fortran (no fused multiply-add).
#define MFM_SHARED_KERNELS 1
#include "mfm_kernels.inc"

namespace mfm {
template hipError_t launch_stage<double>(const Consts&, const Consts*, const DevPatch*,
                                         const StageLists&, const StepArgs&, double*,
                                         long*, hipStream_t, KernelTimer*);
template hipError_t launch_update_summed<double>(const Consts&, const Consts*, const DevPatch*,
                                              const StageLists&, const StepArgs&, double*,
                                              long*, hipStream_t, KernelTimer*);

hipError_t launch_dt_final(const RegionList& L, const double* dt_part,
                           const long* idx_part, double* dt_out, long* idx_out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_dt_final, dim3(L.n), dim3(kBlock), 0, s, L.r, L.n,
                     L.nblk, dt_part, idx_part, dt_out, idx_out);
  return hipGetLastError();
}
hipError_t launch_dt_only(const Consts& C, const Consts* CT, const DevPatch* P,
                          const RegionList& L, double* dt_part, long* idx_part,
                          hipStream_t s) {
#define MFM_DT_ONLY(NF, DED)                                                    \
  hipLaunchKernelGGL((k_dt_only<Sys<NF, DED>>), dim3(L.nblk), dim3(kBlock), 0, \
                     s, CT, P, L.r, L.n, dt_part, idx_part)
  switch (C.nf * 2 + C.dedner) {
    case 2: MFM_DT_ONLY(1, 0); break;
    case 3: MFM_DT_ONLY(1, 1); break;
    case 4: MFM_DT_ONLY(2, 0); break;
    case 6: MFM_DT_ONLY(3, 0); break;
    case 8: MFM_DT_ONLY(4, 0); break;
    case 5: MFM_DT_ONLY(2, 1); break;
    case 7: MFM_DT_ONLY(3, 1); break;
    case 9: MFM_DT_ONLY(4, 1); break;
    default: return hipErrorInvalidValue;
  }
#undef MFM_DT_ONLY
  return hipGetLastError();
}
}  // namespace mfm
