// FP32-flux (fast) build of the stage kernels: reconstruction and Riemann
// fluxes in FP32, cons->prim, sources, floors and accumulation in FP64.
// Contraction into FMA is allowed here.
#include "mfm_kernels.inc"

namespace mfm {
template hipError_t launch_stage<float>(const Consts&, const Consts*, const DevPatch*,
                                         const StageLists&, const StepArgs&, double*,
                                         long*, hipStream_t, KernelTimer*);
template hipError_t launch_update_summed<float>(const Consts&, const Consts*, const DevPatch*,
                                              const StageLists&, const StepArgs&, double*,
                                              long*, hipStream_t, KernelTimer*);
}  // namespace mfm
