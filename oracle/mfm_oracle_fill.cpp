// mfm_oracle_fill.cpp -- CPU restatement of the ghost fill that precedes
// the patch update in AMRLevelIdealMHD::step
// (srcIdeal/AMRLevelIdealMHD.cpp:701-772):
//   - LevelData<FArrayBox>::exchange (Chombo 3.2.4, not in the reference
//     tree): every ghost cell of a box that lies in another box of the level,
//     or in a periodic image of any box, receives that box's valid value;
//   - PiecewiseLinearFillPatchMHDAM::define + fillInterp
//     (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:118-880) with interplimit_v
//     and interp_divB_0 (srcIdeal/FineInterpMHDAMF.ChF:235-555), Cartesian
//     coordinates (src/CSHCartesian.cpp: constant dx, PLFP_* no-ops);
//   - the coarse-fine synchronisation after the updates
//     (srcIdeal/AMRLevelIdealMHD.cpp:695, 1282-1307, 1636-1647): the flux
//     register (Chombo 3.2.4 LevelFluxRegister, not in the reference tree;
//     restated from its published algorithm: the register holds, per coarse
//     face of the coarse-fine interface, the time-integrated fine flux
//     averaged over the face minus the coarse one, and reflux corrects the
//     coarse cell outside the fine region by it) and the average down
//     (CoarseAverageExt, src/CoarseAverageExt.cpp:126-194).
//
// TEST INFRASTRUCTURE ONLY, like mfm_oracle.cpp.  The IntVectSets of the
// reference are kept as ordered sets of cells and built by the same set
// operations in the same order; FABs are dense arrays with the reference's
// -666.666 fill.
#include <array>
#include <cmath>
#include <set>
#include <vector>

#include "mfm_oracle.h"

namespace {

using IV = std::array<int, 3>;
using IVS = std::set<IV>;

int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a - 1) / r) - 1; }

struct OBox {
  IV lo, hi;
  bool empty() const { return lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]; }
  bool has(const IV& v) const {
    for (int d = 0; d < 3; ++d)
      if (v[d] < lo[d] || v[d] > hi[d]) return false;
    return true;
  }
  bool contains(const OBox& b) const { return has(b.lo) && has(b.hi); }
  long n(int d) const { return (long)hi[d] - lo[d] + 1; }
};
OBox ob(const mfm_box& b) {
  OBox o;
  for (int d = 0; d < 3; ++d) { o.lo[d] = b.lo[d]; o.hi[d] = b.hi[d]; }
  return o;
}
OBox growB(OBox b, int g, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] -= g; b.hi[d] += g; }
  return b;
}
OBox coarsenB(OBox b, int r, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] = fdiv(b.lo[d], r); b.hi[d] = fdiv(b.hi[d], r); }
  return b;
}
OBox shiftB(OBox b, const IV& s) {
  for (int d = 0; d < 3; ++d) { b.lo[d] += s[d]; b.hi[d] += s[d]; }
  return b;
}
OBox andB(const OBox& a, const OBox& b) {
  OBox c;
  for (int d = 0; d < 3; ++d) { c.lo[d] = std::max(a.lo[d], b.lo[d]); c.hi[d] = std::min(a.hi[d], b.hi[d]); }
  return c;
}
IV coarsenIV(IV v, int r, int dims) {
  for (int d = 0; d < dims; ++d) v[d] = fdiv(v[d], r);
  return v;
}

// ProblemDomain: box + periodicity
struct PD {
  OBox dom;
  bool per[3] = {false, false, false};
  int dims = 3;
  bool periodic() const { return per[0] || per[1] || per[2]; }
  // ProblemDomain & Box: no clipping in periodic directions
  OBox clip(OBox b) const {
    for (int d = 0; d < dims; ++d)
      if (!per[d]) { b.lo[d] = std::max(b.lo[d], dom.lo[d]); b.hi[d] = std::min(b.hi[d], dom.hi[d]); }
    return b;
  }
  // ShiftIterator: every non-zero combination of -1, 0, +1 in the periodic
  // directions, times the domain size
  std::vector<IV> shifts() const {
    std::vector<IV> v;
    int lim[3];
    for (int d = 0; d < 3; ++d) lim[d] = (d < dims && per[d]) ? 1 : 0;
    for (int c = -lim[2]; c <= lim[2]; ++c)
      for (int b = -lim[1]; b <= lim[1]; ++b)
        for (int a = -lim[0]; a <= lim[0]; ++a)
          if (a || b || c)
            v.push_back({a * (int)dom.n(0), b * (int)dom.n(1), c * (int)dom.n(2)});
    return v;
  }
};

IVS ivs_of(const OBox& b) {
  IVS s;
  if (b.empty()) return s;
  for (int k = b.lo[2]; k <= b.hi[2]; ++k)
    for (int j = b.lo[1]; j <= b.hi[1]; ++j)
      for (int i = b.lo[0]; i <= b.hi[0]; ++i) s.insert({i, j, k});
  return s;
}
void ivs_sub(IVS& s, const OBox& b) {
  for (auto it = s.begin(); it != s.end();) {
    if (b.has(*it)) it = s.erase(it); else ++it;
  }
}
void ivs_sub(IVS& s, const IVS& o) {
  for (const IV& v : o) s.erase(v);
}
IVS ivs_shift(const IVS& s, const IV& sh) {
  IVS o;
  for (IV v : s) { for (int d = 0; d < 3; ++d) v[d] += sh[d]; o.insert(v); }
  return o;
}
IVS ivs_refine(const IVS& s, int r, int dims) {
  IVS o;
  for (const IV& v : s) {
    IV lo = v, hi = v;
    for (int d = 0; d < dims; ++d) { lo[d] = v[d] * r; hi[d] = v[d] * r + r - 1; }
    for (int k = lo[2]; k <= hi[2]; ++k)
      for (int j = lo[1]; j <= hi[1]; ++j)
        for (int i = lo[0]; i <= hi[0]; ++i) o.insert({i, j, k});
  }
  return o;
}
void ivs_and(IVS& s, const OBox& b) {
  for (auto it = s.begin(); it != s.end();) {
    if (!b.has(*it)) it = s.erase(it); else ++it;
  }
}

struct DFab {
  OBox b;
  int nc = 0;
  std::vector<double> v;
  void define(const OBox& bb, int ncomp, double val) {
    b = bb;
    nc = ncomp;
    v.assign((size_t)b.n(0) * b.n(1) * b.n(2) * nc, val);
  }
  double& at(const IV& x, int c) {
    return v[(size_t)((x[0] - b.lo[0]) + b.n(0) * ((x[1] - b.lo[1]) + b.n(1) * ((long)(x[2] - b.lo[2]) + b.n(2) * c)))];
  }
};

struct HFab {  // view of a caller FAB
  double* p;
  OBox b;
  int nc;
  double& at(const IV& x, int c) const {
    return p[(size_t)((x[0] - b.lo[0]) + b.n(0) * ((x[1] - b.lo[1]) + b.n(1) * ((long)(x[2] - b.lo[2]) + b.n(2) * c)))];
  }
};
HFab hf(const mfm_fab& f) { return HFab{f.data, ob(f.box), f.ncomp}; }

PD level_domain(const mfm_params* p) {
  PD D;
  D.dims = p->spacedim;
  for (int d = 0; d < 3; ++d) {
    D.dom.lo[d] = p->domain_lo[d];
    D.dom.hi[d] = p->domain_hi[d];
    D.per[d] = d < D.dims && p->periodic[d] != 0;
  }
  return D;
}

// LevelData::copyTo of the valid data of a coarse layout into dest (whose
// domain is periodic like the coarse domain): each dest cell lying in a
// coarse box, or in a periodic image of one, takes that box's value.
void copy_to(const PD& CD, const std::vector<OBox>& cb, const mfm_fab* src, DFab& dst) {
  std::vector<IV> sh = CD.shifts();
  sh.insert(sh.begin(), IV{0, 0, 0});
  for (size_t j = 0; j < cb.size(); ++j) {
    HFab S = hf(src[j]);
    for (const IV& s : sh) {
      OBox I = andB(shiftB(cb[j], s), dst.b);
      if (I.empty()) continue;
      for (int k = I.lo[2]; k <= I.hi[2]; ++k)
        for (int jj = I.lo[1]; jj <= I.hi[1]; ++jj)
          for (int i = I.lo[0]; i <= I.hi[0]; ++i)
            for (int c = 0; c < dst.nc; ++c)
              dst.at({i, jj, k}, c) = S.at({i - s[0], jj - s[1], k - s[2]}, c);
    }
  }
}

// minmod of src/reconstructionF.ChF:4-14
double minmod(double A, double B) {
  const double signA = std::copysign(1.0, A);
  return signA * std::max(0.0, std::min(std::fabs(A), signA * B));
}

// interp_divB_0 (srcIdeal/FineInterpMHDAMF.ChF:449-555) for one cell
void divb0(double& sx, double& sy, double& sz, int dims) {
  const double eps_divB = 1.0e-8;
  const double DBXDX = sx, DBYDY = sy, DBZDZ = dims == 3 ? sz : 0.0;
  const double fact = -1.0 / (dims == 3 ? (1.0 + 1.0 + 1.0) : (1.0 + 1.0));
  const double divB = dims == 3 ? DBXDX + DBYDY + DBZDZ : DBXDX + DBYDY;
  double del = fact * divB;
  double adel = std::fabs(del);
  if (adel > eps_divB) {
    double delX = minmod(-DBXDX, del);
    double delY = minmod(-DBYDY, del);
    double delZ = dims == 3 ? minmod(-DBZDZ, del) : 0.0;
    double del2 = dims == 3 ? delX + delY + delZ : delX + delY;
    if (std::fabs(del2) < std::fabs(divB)) {
      if (dims == 2) {
        delX = minmod(DBXDX, -DBYDY);
        delY = -delX;
        delX = delX - DBXDX;
        delY = delY - DBYDY;
      } else {
        if (std::fabs(delX) < adel) {
          del = -0.5 * (delX + divB);
          delY = minmod(-DBYDY, del);
          delZ = minmod(-DBZDZ, del);
          del2 = delX + delY + delZ;
          if (std::fabs(del2) < std::fabs(divB)) {
            if (std::fabs(delY) < adel) delZ = -(divB + delX + delY);
            else delY = -(divB + delX + delZ);
          }
        } else {
          if (std::fabs(delY) < adel) {
            del = -0.5 * (delY + divB);
            delX = minmod(-DBXDX, del);
            delZ = minmod(-DBZDZ, del);
            del2 = delX + delY + delZ;
            if (std::fabs(del2) < std::fabs(divB)) {
              if (std::fabs(delX) < adel) delZ = -(divB + delX + delY);
              else delX = -(divB + delY + delZ);
            }
          } else {
            del = -0.5 * (delZ + divB);
            adel = std::fabs(del);
            delX = minmod(-DBXDX, del);
            delY = minmod(-DBYDY, del);
            del2 = delX + delY + delZ;
            if (std::fabs(del2) < std::fabs(divB)) {
              if (std::fabs(delX) < adel) delY = -(divB + delX + delZ);
              else delX = -(divB + delY + delZ);
            }
          }
        }
      }
    }
    sx = sx + delX;
    sy = sy + delY;
    if (dims == 3) sz = sz + delZ;
  }
}

}  // namespace

extern "C" {

int oracle_level_exchange(const mfm_params* p, const mfm_box* boxes, int n,
                          mfm_fab* U, int ng) {
  const PD D = level_domain(p);
  std::vector<OBox> B(n);
  for (int i = 0; i < n; ++i) B[i] = ob(boxes[i]);
  for (int i = 0; i < n; ++i) {
    HFab Ui = hf(U[i]);
    const OBox G = growB(B[i], ng, D.dims);
    for (int k = G.lo[2]; k <= G.hi[2]; ++k)
      for (int j = G.lo[1]; j <= G.hi[1]; ++j)
        for (int ii = G.lo[0]; ii <= G.hi[0]; ++ii) {
          const IV g = {ii, j, k};
          if (B[i].has(g)) continue;
          IV w = g;
          bool inside = true;
          for (int d = 0; d < D.dims; ++d) {
            const long L = D.dom.n(d);
            if (D.per[d]) {
              long o = ((long)w[d] - D.dom.lo[d]) % L;
              if (o < 0) o += L;
              w[d] = (int)(D.dom.lo[d] + o);
            } else if (w[d] < D.dom.lo[d] || w[d] > D.dom.hi[d]) {
              inside = false;
            }
          }
          if (!inside) continue;
          for (int jb = 0; jb < n; ++jb)
            if (B[jb].has(w)) {
              HFab Uj = hf(U[jb]);
              for (int c = 0; c < Ui.nc; ++c) Ui.at(g, c) = Uj.at(w, c);
              break;
            }
        }
  }
  return MFM_OK;
}

void oracle_interp_divb(double s[3], int dims) { divb0(s[0], s[1], s[2], dims); }

int oracle_fill_interp(const mfm_params* p, const mfm_box* boxes, int n, mfm_fab* U,
                       const mfm_box* cboxes, int ncb, const mfm_fab* cold,
                       const mfm_fab* cnew, int r, int ng, double dxc, double alpha) {
  const PD FD = level_domain(p);
  const int dims = FD.dims;
  PD CD = FD;
  CD.dom = coarsenB(FD.dom, r, dims);
  std::vector<OBox> fb(n), cb(ncb), cfb(n);
  for (int i = 0; i < n; ++i) { fb[i] = ob(boxes[i]); cfb[i] = coarsenB(fb[i], r, dims); }
  for (int j = 0; j < ncb; ++j) cb[j] = ob(cboxes[j]);
  const int nc = U[0].ncomp;
  const double dxf = dxc / r;                 // src/CSHCartesian.cpp:50-52
  const int iBX = p->dedner ? 5 : -1;          // srcIdeal/AMRLevelIdealMHD.cpp:388-394
  const double tiny = p->ce_single_literals ? (double)1.0e-9f : 1.0e-9;

  // ---- define (PiecewiseLinearFillPatchMHDAM.cpp:230-376)
  const int coarse_slope_radius = (ng + r - 1) / r;
  const int coarse_ghost_radius = coarse_slope_radius + 1;  // s_stencil_radius = 1
  OBox periodicTestBox = CD.dom;
  if (CD.periodic())
    for (int d = 0; d < dims; ++d)
      if (CD.per[d]) { periodicTestBox.lo[d] += 1; periodicTestBox.hi[d] -= 1; }
  const std::vector<IV> csh = CD.shifts();
  std::vector<IVS> fine_interp(n);
  std::vector<std::array<IVS, 3>> cen(n), clo(n), chi(n);
  for (int i = 0; i < n; ++i) {
    const OBox coarsened_fine_box = CD.clip(coarsenB(growB(fb[i], ng, dims), r, dims));
    IVS cfi = ivs_of(coarsened_fine_box);
    for (int o = 0; o < n; ++o) {
      const OBox& other = cfb[o];
      ivs_sub(cfi, other);
      if (CD.periodic() && !periodicTestBox.contains(other) &&
          !periodicTestBox.contains(coarsened_fine_box))
        for (const IV& s : csh) ivs_sub(cfi, shiftB(other, s));
    }
    for (int dir = 0; dir < dims; ++dir) {
      IV e = {0, 0, 0};
      e[dir] = 1;
      const IV me = {-e[0], -e[1], -e[2]};
      IVS& centered = cen[i][dir];
      centered = cfi;
      IVS lo = ivs_shift(centered, e);
      IVS hi = ivs_shift(centered, me);
      for (int j = 0; j < ncb; ++j) {
        ivs_sub(lo, cb[j]);
        ivs_sub(hi, cb[j]);
        if (CD.periodic() && !periodicTestBox.contains(cb[j]) &&
            !periodicTestBox.contains(coarsened_fine_box))
          for (const IV& s : csh) { ivs_sub(lo, shiftB(cb[j], s)); ivs_sub(hi, shiftB(cb[j], s)); }
      }
      lo = ivs_shift(lo, me);
      hi = ivs_shift(hi, e);
      ivs_sub(centered, lo);
      ivs_sub(centered, hi);
      clo[i][dir] = lo;
      chi[i][dir] = hi;
    }
    IVS fi = ivs_refine(cfi, r, dims);
    ivs_and(fi, FD.clip(growB(fb[i], ng, dims)));
    fine_interp[i] = fi;
  }

  // ---- fillInterp (:383-432)
  // timeInterp (:435-533) into the coarsened-fine work FABs
  const bool have_old = cold != nullptr, have_new = cnew != nullptr;
  std::vector<DFab> work(n);
  for (int i = 0; i < n; ++i) {
    work[i].define(growB(cfb[i], coarse_ghost_radius, dims), nc, -666.666);
    if (alpha == 1.0 || !have_old) {
      copy_to(CD, cb, cnew, work[i]);
    } else if (alpha == 0.0 || !have_new) {
      copy_to(CD, cb, cold, work[i]);
    } else {
      copy_to(CD, cb, cnew, work[i]);
      DFab tmp;
      tmp.define(work[i].b, nc, -666.666);
      copy_to(CD, cb, cold, tmp);
      for (size_t q = 0; q < work[i].v.size(); ++q) work[i].v[q] *= alpha;
      for (size_t q = 0; q < tmp.v.size(); ++q) tmp.v[q] *= (1. - alpha);
      for (size_t q = 0; q < work[i].v.size(); ++q) work[i].v[q] += tmp.v[q];
    }
  }
  // fillConstantInterp (:537-585)
  for (int i = 0; i < n; ++i) {
    HFab F = hf(U[i]);
    for (const IV& f : fine_interp[i]) {
      const IV c = coarsenIV(f, r, dims);
      for (int q = 0; q < nc; ++q) F.at(f, q) = work[i].at(c, q);
    }
  }
  // computeSlopes (:588-800)
  std::vector<std::array<DFab, 3>> slopes(n);
  for (int i = 0; i < n; ++i) {
    for (int dir = 0; dir < 3; ++dir)
      slopes[i][dir].define(growB(cfb[i], coarse_slope_radius, dims), nc, -666.666);
    DFab& W = work[i];
    for (int dir = 0; dir < dims; ++dir) {
      IV e = {0, 0, 0};
      e[dir] = 1;
      DFab& S = slopes[i][dir];
      double c1 = 0.5 / dxc;
      for (const IV& iv : cen[i][dir]) {
        const IV ivlo = {iv[0] - e[0], iv[1] - e[1], iv[2] - e[2]};
        const IV ivhi = {iv[0] + e[0], iv[1] + e[1], iv[2] + e[2]};
        for (int q = 0; q < nc; ++q) S.at(iv, q) = c1 * (W.at(ivhi, q) - W.at(ivlo, q));
      }
      c1 = 1.0 / dxc;
      for (const IV& iv : clo[i][dir]) {
        const IV ivlo = {iv[0] - e[0], iv[1] - e[1], iv[2] - e[2]};
        for (int q = 0; q < nc; ++q) S.at(iv, q) = c1 * (W.at(iv, q) - W.at(ivlo, q));
      }
      c1 = 1.0 / dxc;
      for (const IV& iv : chi[i][dir]) {
        const IV ivhi = {iv[0] + e[0], iv[1] + e[1], iv[2] + e[2]};
        for (int q = 0; q < nc; ++q) S.at(iv, q) = c1 * (W.at(ivhi, q) - W.at(iv, q));
      }
    }
    // multidimensional limiting: interplimit_v over b_mod
    OBox b_mod = growB(CD.clip(growB(slopes[i][0].b, 1, dims)), -1, dims);
    const int kz = dims == 3 ? 1 : 0;
    for (int var = 0; var < nc; ++var)
      for (int k = b_mod.lo[2]; k <= b_mod.hi[2]; ++k)
        for (int j = b_mod.lo[1]; j <= b_mod.hi[1]; ++j)
          for (int ii = b_mod.lo[0]; ii <= b_mod.hi[0]; ++ii) {
            const IV iv = {ii, j, k};
            double statemax = W.at(iv, var), statemin = W.at(iv, var);
            for (int kk = -kz; kk <= kz; ++kk)
              for (int jj = -1; jj <= 1; ++jj)
                for (int i3 = -1; i3 <= 1; ++i3) {
                  const IV nb = {ii + i3, j + jj, k + kk};
                  if (W.b.has(nb)) {
                    statemax = std::max(statemax, W.at(nb, var));
                    statemin = std::min(statemin, W.at(nb, var));
                  }
                }
            double deltasum = std::fabs(slopes[i][0].at(iv, var) * dxc) +
                              std::fabs(slopes[i][1].at(iv, var) * dxc);
            if (dims == 3) deltasum = deltasum + std::fabs(slopes[i][2].at(iv, var) * dxc);
            deltasum = 0.5 * deltasum;
            double eta = std::min(statemax - W.at(iv, var), W.at(iv, var) - statemin);
            if (eta <= tiny * std::fabs(statemax)) {
              eta = 0.0;
            } else {
              if (deltasum > eta) eta = eta / deltasum;
              else eta = 1.0;
            }
            for (int d = 0; d < dims; ++d) slopes[i][d].at(iv, var) = eta * slopes[i][d].at(iv, var);
          }
    if (iBX >= 0 && iBX + dims - 1 < nc)
      for (int k = b_mod.lo[2]; k <= b_mod.hi[2]; ++k)
        for (int j = b_mod.lo[1]; j <= b_mod.hi[1]; ++j)
          for (int ii = b_mod.lo[0]; ii <= b_mod.hi[0]; ++ii) {
            const IV iv = {ii, j, k};
            double dummy = 0.0;
            double& sz = dims == 3 ? slopes[i][2].at(iv, iBX + 2) : dummy;
            divb0(slopes[i][0].at(iv, iBX), slopes[i][1].at(iv, iBX + 1), sz, dims);
          }
  }
  // incrementLinearInterp (:800-880)
  double refScale = 1.0;
  for (int d = 0; d < dims; ++d) refScale *= r;
  double volc = dxc, volf = dxf;
  for (int d = 1; d < dims; ++d) { volc = volc * dxc; volf = volf * dxf; }
  for (int dir = 0; dir < dims; ++dir)
    for (int i = 0; i < n; ++i) {
      HFab F = hf(U[i]);
      for (const IV& f : fine_interp[i]) {
        const IV c = coarsenIV(f, r, dims);
        const double interp_coef = dxf * (f[dir] + 0.5) - dxc * (c[dir] + 0.5);
        const double coeff = volc / (volf * refScale);
        for (int q = 0; q < nc; ++q) {
          const double slope = coeff * interp_coef * slopes[i][dir].at(c, q);
          F.at(f, q) += slope;
        }
      }
    }
  return MFM_OK;
}

}  // extern "C"

// ---- coarse-fine synchronisation ---------------------------------------------
// The register is a dense array over the coarse problem domain:
//   R[(((d * 2 + h) * ncell) + cell) * nc + comp]
// for the face of side h (0 low, 1 high) of coarse cell `cell` in direction d.
namespace {

struct Sync {
  PD FD, CD;
  int dims, r;
  std::vector<OBox> fb, cfb, cb;
  long ncell;
  Sync(const mfm_params* p, const mfm_box* fine, int nf, const mfm_box* crse, int ncb, int r_)
      : dims(p->spacedim), r(r_) {
    FD = level_domain(p);
    CD = FD;
    CD.dom = coarsenB(FD.dom, r, dims);
    for (int i = 0; i < nf; ++i) { fb.push_back(ob(fine[i])); cfb.push_back(coarsenB(fb.back(), r, dims)); }
    for (int j = 0; j < ncb; ++j) cb.push_back(ob(crse[j]));
    ncell = CD.dom.n(0) * CD.dom.n(1) * CD.dom.n(2);
  }
  long cell(const IV& c) const {
    return (c[0] - CD.dom.lo[0]) + CD.dom.n(0) * ((long)(c[1] - CD.dom.lo[1]) + CD.dom.n(1) * (c[2] - CD.dom.lo[2]));
  }
  // neighbour across face (d, h), wrapped; false outside a non-periodic domain
  bool across(IV c, int d, int h, IV& n) const {
    c[d] += h ? 1 : -1;
    if (c[d] < CD.dom.lo[d] || c[d] > CD.dom.hi[d]) {
      if (!CD.per[d]) return false;
      c[d] += c[d] < CD.dom.lo[d] ? (int)CD.dom.n(d) : -(int)CD.dom.n(d);
    }
    n = c;
    return true;
  }
  int fine_box_over(const IV& c) const {
    for (size_t i = 0; i < cfb.size(); ++i)
      if (cfb[i].has(c)) return (int)i;
    return -1;
  }
  // (C, d, h) is a register face: C outside the fine region, the cell across
  // inside it
  bool is_face(const IV& c, int d, int h, int* f) const {
    if (fine_box_over(c) >= 0) return false;
    IV n;
    if (!across(c, d, h, n)) return false;
    const int i = fine_box_over(n);
    if (i < 0) return false;
    if (f) *f = i;
    return true;
  }
  template <class Fn>
  void each_face(Fn fn) const {
    for (size_t j = 0; j < cb.size(); ++j) {
      const OBox& b = cb[j];
      for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int jj = b.lo[1]; jj <= b.hi[1]; ++jj)
          for (int i = b.lo[0]; i <= b.hi[0]; ++i)
            for (int d = 0; d < dims; ++d)
              for (int h = 0; h < 2; ++h) {
                const IV c = {i, jj, k};
                int f = -1;
                if (is_face(c, d, h, &f)) fn((int)j, c, d, h, f);
              }
    }
  }
};

}  // namespace

extern "C" {

int oracle_fr_increment_coarse(const mfm_params* p, const mfm_box* fine, int nf,
                               const mfm_box* crse, int ncb, int r, const mfm_fab* F,
                               double scale, double* R) {
  const Sync S(p, fine, nf, crse, ncb, r);
  const int nc = F[0].ncomp;
  S.each_face([&](int j, const IV& c, int d, int h, int) {
    HFab f = hf(F[j * S.dims + d]);
    IV face = c;
    face[d] += h;
    for (int q = 0; q < nc; ++q) {
      double& reg = R[((long)(d * 2 + h) * S.ncell + S.cell(c)) * nc + q];
      reg -= scale * f.at(face, q);
    }
  });
  return MFM_OK;
}

int oracle_fr_increment_fine(const mfm_params* p, const mfm_box* fine, int nf,
                             const mfm_box* crse, int ncb, int r, const mfm_fab* F,
                             double scale, double* R) {
  const Sync S(p, fine, nf, crse, ncb, r);
  const int nc = F[0].ncomp;
  double denom = 1.0;
  for (int q = 0; q < S.dims - 1; ++q) denom *= r;
  S.each_face([&](int, const IV& c, int d, int h, int fi) {
    IV n;
    S.across(c, d, h, n);
    HFab f = hf(F[fi * S.dims + d]);
    int t[2] = {-1, -1}, nt = 0;
    for (int e = 0; e < S.dims; ++e)
      if (e != d) t[nt++] = e;
    for (int q = 0; q < nc; ++q) {
      double& reg = R[((long)(d * 2 + h) * S.ncell + S.cell(c)) * nc + q];
      for (int q2 = 0; q2 < (nt > 1 ? r : 1); ++q2)
        for (int q1 = 0; q1 < r; ++q1) {
          IV face = n;
          face[d] = h ? n[d] * r : (n[d] + 1) * r;   // the fine region's boundary face
          face[t[0]] = n[t[0]] * r + q1;
          if (nt > 1) face[t[1]] = n[t[1]] * r + q2;
          reg += (scale / denom) * f.at(face, q);
        }
    }
  });
  return MFM_OK;
}

int oracle_fr_reflux(const mfm_params* p, const mfm_box* fine, int nf, const mfm_box* crse,
                     int ncb, int r, mfm_fab* U, double scale, double dxc, const double* R) {
  const Sync S(p, fine, nf, crse, ncb, r);
  const int nc = U[0].ncomp;
  const double invvol = 1.0 / dxc;
  S.each_face([&](int j, const IV& c, int d, int h, int) {
    HFab u = hf(U[j]);
    const double sigma = h ? 1.0 : -1.0;
    for (int q = 0; q < nc; ++q) {
      const double reg = R[((long)(d * 2 + h) * S.ncell + S.cell(c)) * nc + q];
      u.at(c, q) = u.at(c, q) + (scale * sigma * invvol) * reg;
    }
  });
  return MFM_OK;
}

int oracle_average_down(const mfm_params* p, const mfm_box* fine, int nf, const mfm_box* crse,
                        int ncb, int r, mfm_fab* Uc, const mfm_fab* Uf, double dxf) {
  const Sync S(p, fine, nf, crse, ncb, r);
  const int dims = S.dims;
  double vol = dxf;
  for (int d = 1; d < dims; ++d) vol = vol * dxf;
  for (int i = 0; i < nf; ++i) {
    HFab f = hf(Uf[i]);
    for (int j = 0; j < ncb; ++j) {
      const OBox I = andB(S.cfb[i], S.cb[j]);
      if (I.empty()) continue;
      HFab u = hf(Uc[j]);
      for (int var = 0; var < u.nc; ++var)
        for (int kc = I.lo[2]; kc <= I.hi[2]; ++kc)
          for (int jc = I.lo[1]; jc <= I.hi[1]; ++jc)
            for (int ic = I.lo[0]; ic <= I.hi[0]; ++ic) {
              double coarseSum = 0.0, volc = 0.0;
              const int rz = dims == 3 ? r : 1;
              for (int kk = 0; kk < rz; ++kk)
                for (int jj = 0; jj < r; ++jj)
                  for (int ii = 0; ii < r; ++ii) {
                    const IV fc = {ic * r + ii, jc * r + jj, dims == 3 ? kc * r + kk : kc};
                    coarseSum = coarseSum + vol * f.at(fc, var);
                    volc = volc + vol;
                  }
              u.at({ic, jc, kc}, var) = coarseSum / volc;
            }
    }
  }
  return MFM_OK;
}

}  // extern "C"
