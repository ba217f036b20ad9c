// FP64 (parity) build of the tile kernels (fused stage, face sums),
// -ffp-contract=off as mfm_kernels_fp64.hip: every product and sum rounds
// as in the reference.
#include "mfm_kernels.inc"
#include "mfm_stage.inc"

namespace mfm {
template hipError_t launch_stage_fused<double>(const Consts&, const Consts*, const DevPatch*,
                                               const StageLists&, const StepArgs&, double*,
                                               long*, hipStream_t, KernelTimer*);
template hipError_t launch_face_sums<double>(const Consts&, const Consts*, const DevPatch*,
                                            const StageLists&, const StepArgs&, hipStream_t,
                                            KernelTimer*);
}  // namespace mfm
