// mfm_fill.hpp -- structures shared by the ghost-fill plan (host) and its
// kernels: rectangle copies between FABs and the coarse-fine interpolation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfm {

// A FArrayBox as the fill kernels see it (component-major, x fastest).
struct FillFab {
  double* p;
  int lo[3], n[3];
  long cs;  // component stride = n0 n1 n2
};

// Copy of a rectangle of cells (all components) from FAB `src` of one table
// to FAB `dst` of another; src cell = dst cell - (dlo - slo).  blk0 = first
// block of the launch that works on it.
struct CopyRect {
  int src, dst;
  int slo[3], dlo[3], n[3];
  int blk0;
};

// One coarse cell gathered into a work FAB: source coarse FAB and cell
// (periodic image already undone), destination work FAB and offset.
struct GatherCell {
  int src, dst, doff;
  int s[3];
};

// One fine box of an interpolation plan: its coarse work FAB (coarsened box
// grown by the slope radius + 1) and the fine cells it may write (box grown
// by the interpolation radius, clipped to the non-periodic domain).
struct FillPatch {
  int wlo[3], wn[3];
  long woff;  // offset of the work FAB in the work arena (doubles)
  int glo[3], ghi[3];
};

// Per coarse cell of the interpolation: 2 bits per direction for the slope
// (0 centred, 1 one-sided low, 2 one-sided high) and the limiter flag (the
// cell lies inside the limiting box).
enum : uint8_t { kSlopeCentred = 0, kSlopeLo = 1, kSlopeHi = 2, kFillLimit = 64 };

struct InterpArgs {
  const int2* cell;     // (fine box, cell offset in its work FAB)
  const uint8_t* flag;  // per cell
  long ncell;
  const FillPatch* P;   // per fine box
  const double* work;   // work arena
  const FillFab* U;     // fine FABs (per call)
  int nc, dims, r, ibx; // ibx < 0: no div(B) slope correction
  double c_centre, c_side, dxc, dxf, coeff, tiny;
};

hipError_t launch_rect_copy(const CopyRect* R, int nr, int nblk,
                            const FillFab* src_new, const FillFab* src_old,
                            const FillFab* dst, int nc, double a, double b,
                            int mode, hipStream_t s);
hipError_t launch_cell_copy(const GatherCell* G, long n, const FillFab* src_new,
                            const FillFab* src_old, const FillFab* dst, int nc,
                            double a, double b, int mode, hipStream_t s);
hipError_t launch_fill_interp(const InterpArgs& A, hipStream_t s);

}  // namespace mfm

namespace mfm {

// One coarse face of a coarse-fine interface (flux register entry): the
// coarse cell outside the fine region that owns it (box `cb`, cell c), the
// direction, sigma = +1 when the face is the cell's high face, and the fine
// box whose boundary faces cover it (first fine face f0; the others follow
// along the transverse directions).
struct RegFace {
  int cb, d, sigma, fb;
  int c[3];
  int f0[3];
};

// The coarse cells of a reflux: entries [e0, e1) of the face list (grouped
// by cell, in direction order, low face first).
struct RegCell {
  int e0, e1;
};

// A coarse cell of an average-down: fine box, coarse box, coarse cell.
struct AvgRect {
  int fb, cb;
  int lo[3], n[3];
  int blk0;
};

hipError_t launch_reg_zero(double* reg, long n, hipStream_t s);
hipError_t launch_reg_coarse(const RegFace* E, long ne, const FillFab* F, int dims,
                             double* reg, int nc, double scale, hipStream_t s);
hipError_t launch_reg_fine(const RegFace* E, long ne, const FillFab* F, int dims, int r,
                           double* reg, int nc, double scale, hipStream_t s);
hipError_t launch_reflux(const RegCell* C, long ncell, const RegFace* E, long ne,
                         const FillFab* U, const double* reg, int nc, double scale,
                         double invvol, hipStream_t s);
hipError_t launch_average_down(const AvgRect* R, int nr, int nblk, const FillFab* Uc,
                               const FillFab* Uf, int nc, int dims, int r, double vol,
                               hipStream_t s);

}  // namespace mfm
