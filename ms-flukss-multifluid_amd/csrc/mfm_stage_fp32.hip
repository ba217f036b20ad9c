// FP32-flux build of the tile kernels (fused stage, face sums), with the
// floating-point rules of mfm_kernels_fp32.hip: FMA contraction, fast FP32
// division and square root, FP32 denormals flushed; FP64 parts unchanged.
#include "mfm_kernels.inc"
#include "mfm_stage.inc"

namespace mfm {
template hipError_t launch_stage_fused<float>(const Consts&, const Consts*, const DevPatch*,
                                              const StageLists&, const StepArgs&, double*,
                                              long*, hipStream_t, KernelTimer*);
template hipError_t launch_face_sums<float>(const Consts&, const Consts*, const DevPatch*,
                                            const StageLists&, const StepArgs&, hipStream_t,
                                            KernelTimer*);
}  // namespace mfm
