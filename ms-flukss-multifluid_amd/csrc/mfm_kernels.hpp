// mfm_kernels.hpp -- structures shared by the host driver and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfm {

struct Consts;

// One patch as the kernels see it.  U/Up/W share the index space of the
// caller's U FAB (box ulo..ulo+un-1); F/Bn scratch live on the face boxes of
// the first-stage value box.
struct DevPatch {
  double* U;        // caller state (updated in place on the valid box)
  double* Up;       // first-stage state U* (scratch)
  void* W;          // primitive variables (scratch, flux precision)
  void* W2;         // Hancock: primitives of the predicted state
  void* Wn;         // where the stage writes W of its output (the next
                    // stage's W): W itself, or a second buffer when the
                    // fused stage kernel still reads W from other blocks
  void* D;          // summed stages: per cell, the face-flux differences
                    // summed over the directions (NC components, then the
                    // normal-B difference), index space of U
  void* F[3];       // face fluxes per direction (scratch, flux precision)
  void* Bn[3];      // face-normal B per direction (scratch)
  double* Fo[3];    // optional time-averaged fluxes returned to the caller
  double* divB;     // optional div(B) output
  int ulo[3], un[3];
  long ucs;         // component stride of U/Up/W
  int flo[3][3], fn[3][3];
  long fcs[3];
  int folo[3][3], fon[3][3];
  long focs[3];
  int dlo[3], dn[3];
  int reinit;       // patch box may intersect the heliospheric inner sphere
  int cidx;         // row of the constants table
  double dx, dt, time;   // cell size, step and time of this patch
  double psi_decay[2];   // exp(-dt_stage Ch^2/Cp^2) of stage 0 / 1
};

// A box of work items belonging to one patch; blk0 = first block.
struct Region {
  int patch;
  int lo[3];
  int n[3];
  int blk0;
};

struct RegionList {
  const Region* r;  // device pointer
  int n;            // number of regions
  int nblk;         // total blocks
  long items;       // cells (or faces) covered
};

// flux[d]: interior faces of direction d; fbc[d]: faces of direction d on
// a physical (non-periodic) domain boundary.
struct StageLists {
  RegionList prim, flux[3], fbc[3], upd, sum;
  int fused;        // upd is a list of k_stage tiles (interior faces and the
                    // update in one launch); flux[] is then unused
  int summed;       // sum is a list of tiles whose face sums (D) the face-sum
                    // kernels form; k_update then reads D; flux[] unused
};

struct StepArgs {
  double dt_frac;      // this Euler stage advances by dt_frac * patch dt
  int stage;           // 0: first (or only) stage, 1: RK2 corrector
  int in_up;           // stage input: 0 = U, 1 = Up
  int out_up;          // stage output: 0 = U, 1 = Up
  int average;         // RK2 second stage: U += (U^n - U*)/2
  int fo_mode;         // caller fluxes: 0 none, 1 average with stage 1, 2 copy
  int reduce_dt;       // compute the CFL reduction from the stage output
  int write_w;         // store W of the stage output (input of the next stage)
  int row_tile;        // work walk: (y, z) tile of rows (see cell_of)
  int xcd_order;       // work walk: 1 = renumber blocks per XCD
  int hancock;         // 0: Euler / RK2 stage; 1: Hancock predictor;
                       // 2: Hancock corrector
  int tz;              // k_stage: planes of one z-chunk of a tile
};

// Optional kernel timing: a pair of events brackets every launch on the
// launch stream; the host reads them after the step has synchronized.
enum KernelKind { KK_PRIM = 0, KK_FLUX_PLASMA, KK_FLUX_NEUTRAL, KK_FLUX_BC, KK_UPDATE,
                  KK_PREDICT, KK_STAGE, KK_FSUM_PLASMA, KK_FSUM_NEUTRAL, KK_COUNT };
struct KernelTimer {
  static constexpr int kMax = 32;
  hipEvent_t ev[2 * kMax];
  int kind[kMax];
  long items[kMax];
  int n = 0;
  void begin(int k, long it, hipStream_t s) {
    if (n >= kMax) return;
    kind[n] = k;
    items[n] = it;
    (void)hipEventRecord(ev[2 * n], s);
  }
  void end(hipStream_t s) {
    if (n >= kMax) return;
    (void)hipEventRecord(ev[2 * n + 1], s);
    ++n;
  }
};

template <typename TF>
hipError_t launch_stage(const Consts& C, const Consts* CT, const DevPatch* P,
                        const StageLists& L, const StepArgs& A, double* dt_part,
                        long* idx_part, hipStream_t s, KernelTimer* T);

// The fused stage kernel (k_stage) over L.upd; launch_stage has already
// run k_prim and the boundary-face fluxes of the stage.
template <typename TF>
hipError_t launch_stage_fused(const Consts& C, const Consts* CT, const DevPatch* P,
                              const StageLists& L, const StepArgs& A, double* dt_part,
                              long* idx_part, hipStream_t s, KernelTimer* T);

// Column tile of the face-sum kernels (k_fsum): long in x, so that the
// faces a wavefront solves lie along contiguous rows; one thread per column.
// A tile solves (TX+1)·TY + TX·(TY+1) + TX·TY faces per plane in whole
// wavefront passes: 14 passes per 256 columns for 64 x 4, 26 per 512 for
// 64 x 8 (7 % fewer per cell; 11.99 against 12.41 ms per 256^3 step), while
// 64 x 16 (50 per 1024, one block per CU) is slower again (12.31 ms).
constexpr int kSumTX = 64, kSumTY = 8, kSumBlock = kSumTX * kSumTY;

// Summed stages: the face-sum kernels over L.sum (after launch_stage has run
// k_prim and the boundary-face fluxes), then k_update reading their sums.
template <typename TF>
hipError_t launch_face_sums(const Consts& C, const Consts* CT, const DevPatch* P,
                            const StageLists& L, const StepArgs& A, hipStream_t s,
                            KernelTimer* T);
template <typename TF>
hipError_t launch_update_summed(const Consts& C, const Consts* CT, const DevPatch* P,
                                const StageLists& L, const StepArgs& A, double* dt_part,
                                long* idx_part, hipStream_t s, KernelTimer* T);

hipError_t launch_dt_final(const RegionList& L, const double* dt_part,
                           const long* idx_part, double* dt_out, long* idx_out,
                           hipStream_t s);
hipError_t launch_dt_only(const Consts& C, const Consts* CT, const DevPatch* P,
                          const RegionList& L, double* dt_part, long* idx_part,
                          hipStream_t s);
hipError_t launch_init(const Consts& C, double* U, const int ulo[3],
                       const int un[3], long ucs, const int lo[3],
                       const int n[3], double dx, hipStream_t s);
hipError_t launch_fill_periodic(double* U, const int ulo[3], const int un[3],
                                long ucs, int nc, const int vlo[3],
                                const int vn[3], int dims, hipStream_t s);

}  // namespace mfm
