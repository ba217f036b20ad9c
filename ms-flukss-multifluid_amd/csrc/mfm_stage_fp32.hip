// FP32-flux build of the fused stage kernel (same floating-point rules as
// mfm_kernels_fp32.hip: FMA contraction, fast FP32 division and square
// root, FP32 denormals flushed; FP64 parts unchanged).
#include "mfm_kernels.inc"
#include "mfm_stage.inc"

namespace mfm {
template hipError_t launch_stage_fused<float>(const Consts&, const Consts*, const DevPatch*,
                                              const StageLists&, const StepArgs&, double*,
                                              long*, hipStream_t, KernelTimer*);
}  // namespace mfm
