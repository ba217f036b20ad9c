// mfm_fill_plan.cpp -- ghost-fill plans of a level (host) and their entry points.
//
// A plan turns a layout (the level's boxes, its problem domain, the next
// coarser level's boxes) into flat device lists:
//   exchange : rectangles copying valid cells of one box (or of a periodic
//              image) into ghost cells of another -- LevelData::exchange;
//   gather   : the coarse cells the interpolation reads (the neighbourhoods
//              of the coarse parents of the ghost cells to fill), copied
//              into one small work FAB per fine box with the time
//              interpolation folded into the copy -- the copyTo of timeInterp
//              (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:435-533) restricted
//              to the cells that are read;
//   interp   : the coarse cells under the fine ghost cells to interpolate,
//              each with its slope stencil per direction and limiter flag --
//              the IntVectSets of PiecewiseLinearFillPatchMHDAM::define
//              (:118-378).
// The sets are formed on dense per-box masks over the work box; cells are
// listed box by box in x-fastest order so that neighbouring threads write
// neighbouring fine cells.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "mfm_fill.hpp"
#include "mfm_host.hpp"

using mfm::CopyRect;
using mfm::FillFab;
using mfm::FillPatch;
using mfm::GatherCell;
using mfm::host::HBox;
using mfm::host::grow;

namespace {

constexpr int kFB = 256;

int fdiv(int a, int r) { return a >= 0 ? a / r : -((-a - 1) / r) - 1; }

bool empty(const HBox& b) {
  for (int d = 0; d < 3; ++d)
    if (b.hi[d] < b.lo[d]) return true;
  return false;
}
HBox inter(const HBox& a, const HBox& b) {
  HBox c;
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = std::max(a.lo[d], b.lo[d]);
    c.hi[d] = std::min(a.hi[d], b.hi[d]);
  }
  return c;
}
HBox shift(HBox b, const int s[3]) {
  for (int d = 0; d < 3; ++d) { b.lo[d] += s[d]; b.hi[d] += s[d]; }
  return b;
}
HBox coarsen(HBox b, int r, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] = fdiv(b.lo[d], r); b.hi[d] = fdiv(b.hi[d], r); }
  return b;
}
HBox refine(HBox b, int r, int dims) {
  for (int d = 0; d < dims; ++d) { b.lo[d] *= r; b.hi[d] = b.hi[d] * r + r - 1; }
  return b;
}
// Uniform bucket index over a set of boxes, for the "which boxes meet this
// box" queries of plan construction.
struct BoxIndex {
  int bs = 16, dims = 3;
  const std::vector<HBox>* boxes = nullptr;
  std::unordered_map<long long, std::vector<int>> map;
  std::vector<int> stamp;
  int tick = 0;
  static long long key(int a, int b, int c) {
    return ((long long)(a & 0x1fffff) << 42) | ((long long)(b & 0x1fffff) << 21) |
           (long long)(c & 0x1fffff);
  }
  void build(const std::vector<HBox>& v, int dims_) {
    boxes = &v;
    dims = dims_;
    long maxe = 1;
    for (const HBox& b : v)
      for (int d = 0; d < dims; ++d) maxe = std::max(maxe, b.n(d));
    bs = (int)std::max(4L, std::min(maxe, 64L));
    stamp.assign(v.size(), 0);
    for (int id = 0; id < (int)v.size(); ++id) {
      int lo[3], hi[3];
      range(v[id], lo, hi);
      for (int c = lo[2]; c <= hi[2]; ++c)
        for (int b = lo[1]; b <= hi[1]; ++b)
          for (int a = lo[0]; a <= hi[0]; ++a) map[key(a, b, c)].push_back(id);
    }
  }
  void range(const HBox& q, int lo[3], int hi[3]) const {
    for (int d = 0; d < 3; ++d) {
      if (d < dims) { lo[d] = fdiv(q.lo[d], bs); hi[d] = fdiv(q.hi[d], bs); }
      else { lo[d] = hi[d] = 0; }
    }
  }
  // ids of boxes that intersect q
  void query(const HBox& q, std::vector<int>& out) {
    out.clear();
    if (empty(q)) return;
    ++tick;
    int lo[3], hi[3];
    range(q, lo, hi);
    for (int c = lo[2]; c <= hi[2]; ++c)
      for (int b = lo[1]; b <= hi[1]; ++b)
        for (int a = lo[0]; a <= hi[0]; ++a) {
          auto it = map.find(key(a, b, c));
          if (it == map.end()) continue;
          for (int id : it->second) {
            if (stamp[id] == tick) continue;
            stamp[id] = tick;
            if (!empty(inter((*boxes)[id], q))) out.push_back(id);
          }
        }
    std::sort(out.begin(), out.end());
  }
};

// Per-call FAB tables kept on the device, one slot per kind of call.  A
// slot is rewritten only when its content changes (a level's FABs rarely
// move), so repeated calls neither copy nor wait.  When it does change, the
// host first waits for the slot's previous copy (its pinned buffer is
// reused); the new copy is stream-ordered after the kernels that read the
// slot before, which is why the calls on one plan use one stream.
struct TableSlot {
  FillFab* h = nullptr;
  FillFab* d = nullptr;
  size_t cap = 0;
  std::vector<FillFab> last;
  bool valid = false;
  hipEvent_t ev = nullptr;
  bool pending = false;
  void release() {
    if (pending) (void)hipEventSynchronize(ev);
    if (ev) (void)hipEventDestroy(ev);
    if (h) (void)hipHostFree(h);
    if (d) (void)hipFree(d);
    h = nullptr; d = nullptr; ev = nullptr; cap = 0; pending = false; valid = false;
  }
};

int slot_put(mfm_ctx* c, TableSlot& S, const std::vector<FillFab>& tab, hipStream_t st) {
  if (S.valid && tab.size() == S.last.size() &&
      (tab.empty() || std::memcmp(tab.data(), S.last.data(), tab.size() * sizeof(FillFab)) == 0))
    return MFM_OK;
  hipError_t e;
  if (tab.size() > S.cap || !S.d) {
    if (S.pending) (void)hipEventSynchronize(S.ev);
    S.pending = false;
    if (S.d) {
      e = hipDeviceSynchronize();  // kernels may still read the old slot
      if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipDeviceSynchronize");
      (void)hipFree(S.d);
      S.d = nullptr;
    }
    if (S.h) { (void)hipHostFree(S.h); S.h = nullptr; }
    const size_t cap = std::max<size_t>(tab.size(), 8);
    e = hipMalloc((void**)&S.d, cap * sizeof(FillFab));
    if (e != hipSuccess) { S.d = nullptr; return mfm::host::fail(c, MFM_ERR_NOMEM, "hipMalloc of a FAB table failed"); }
    e = hipHostMalloc((void**)&S.h, cap * sizeof(FillFab), hipHostMallocDefault);
    if (e != hipSuccess) { S.h = nullptr; return mfm::host::fail(c, MFM_ERR_NOMEM, "hipHostMalloc of a FAB table failed"); }
    if (!S.ev) {
      e = hipEventCreateWithFlags(&S.ev, hipEventDisableTiming);
      if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipEventCreate");
    }
    S.cap = cap;
  }
  if (S.pending) {
    e = hipEventSynchronize(S.ev);
    if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipEventSynchronize");
    S.pending = false;
  }
  S.valid = false;
  if (!tab.empty()) {
    std::memcpy(S.h, tab.data(), tab.size() * sizeof(FillFab));
    e = hipMemcpyAsync(S.d, S.h, tab.size() * sizeof(FillFab), hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return mfm::host::hipfail(c, e, "FAB table copy");
    e = hipEventRecord(S.ev, st);
    if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipEventRecord");
    S.pending = true;
  }
  S.last = tab;
  S.valid = true;
  return MFM_OK;
}

}  // namespace

struct mfm_fill_plan {
  int dims = 3, nc = 0, r = 1, ng = 0, ibx = -1;
  bool coarse = false;
  bool ok = false;  // built without error
  double dxc = 0, dxf = 0, coeff = 0, tiny = 1e-9;
  std::string err;
  std::vector<HBox> fine, crse;
  std::vector<CopyRect> xr;      // exchange
  int xr_blk = 0;
  std::vector<GatherCell> gc;    // gather
  std::vector<int2> cells;
  std::vector<uint8_t> flags;
  std::vector<FillPatch> fp;
  size_t work_doubles = 0;
  long x_cells = 0, f_cells = 0;
  // device copies, made on first use
  int device = -1;
  char* dmem = nullptr;
  size_t dbytes = 0;
  CopyRect* d_xr = nullptr;
  GatherCell* d_gc = nullptr;
  int2* d_cells = nullptr;
  uint8_t* d_flags = nullptr;
  FillPatch* d_fp = nullptr;
  double* d_work = nullptr;
  FillFab* d_wfab = nullptr;  // work FABs
  TableSlot tab[2];           // exchange: fine; interpolation: fine, coarse new, coarse old
};

namespace {

int pfail(mfm_fill_plan* P, int code, const std::string& m) {
  P->err = m;
  return code;
}

// Periodic shift vectors (in cells of a domain of size L): every combination
// of -1, 0, +1 over the periodic directions, the zero shift first.
std::vector<std::array<int, 3>> shifts(const int per[3], const int L[3], int dims) {
  std::vector<std::array<int, 3>> v;
  int lim[3];
  for (int d = 0; d < 3; ++d) lim[d] = (d < dims && per[d]) ? 1 : 0;
  v.push_back({0, 0, 0});
  for (int c = -lim[2]; c <= lim[2]; ++c)
    for (int b = -lim[1]; b <= lim[1]; ++b)
      for (int a = -lim[0]; a <= lim[0]; ++a)
        if (a || b || c) v.push_back({a * L[0], b * L[1], c * L[2]});
  return v;
}

void add_rect(std::vector<CopyRect>& v, int& nblk, int src, int dst,
              const HBox& dstbox, const int s[3]) {
  const long cells = dstbox.size();
  if (cells <= 0) return;
  CopyRect r;
  r.src = src;
  r.dst = dst;
  for (int d = 0; d < 3; ++d) {
    r.dlo[d] = dstbox.lo[d];
    r.slo[d] = dstbox.lo[d] - s[d];
    r.n[d] = (int)dstbox.n(d);
  }
  r.blk0 = nblk;
  nblk += (int)((cells + kFB - 1) / kFB);
  v.push_back(r);
}

int build(mfm_fill_plan* P, const mfm_params* p, const mfm_box* boxes, int n,
          const mfm_box* cboxes, int ncb) {
  using mfm::host::hb;
  const int dims = P->dims;
  HBox D;
  int per[3] = {0, 0, 0}, L[3] = {1, 1, 1};
  for (int d = 0; d < 3; ++d) {
    D.lo[d] = p->domain_lo[d];
    D.hi[d] = p->domain_hi[d];
    if (d < dims) { per[d] = p->periodic[d] != 0; L[d] = (int)D.n(d); }
  }
  if (empty(D)) return pfail(P, MFM_ERR_ARG, "empty problem domain");
  for (int i = 0; i < n; ++i) {
    HBox b = hb(boxes[i]);
    if (empty(b)) return pfail(P, MFM_ERR_SHAPE, "empty box in the layout");
    for (int d = 0; d < dims; ++d)
      if (b.lo[d] < D.lo[d] || b.hi[d] > D.hi[d])
        return pfail(P, MFM_ERR_SHAPE, "box outside the problem domain");
    P->fine.push_back(b);
  }
  // boxes of a layout are disjoint
  BoxIndex FI;
  FI.build(P->fine, dims);
  std::vector<int> hits;
  for (int i = 0; i < n; ++i) {
    FI.query(P->fine[i], hits);
    if (hits.size() != 1) return pfail(P, MFM_ERR_SHAPE, "boxes of the layout overlap");
  }
  const auto sh = shifts(per, L, dims);

  // ---- exchange: ghost cells of box i inside box j or a periodic image
  for (int i = 0; i < n; ++i) {
    const HBox G = grow(P->fine[i], P->ng, dims);
    for (const auto& s : sh) {
      const int ms[3] = {-s[0], -s[1], -s[2]};
      FI.query(shift(G, ms), hits);
      for (int j : hits) {
        if (j == i && !s[0] && !s[1] && !s[2]) continue;
        HBox I = inter(G, shift(P->fine[j], s.data()));
        if (empty(I)) continue;
        add_rect(P->xr, P->xr_blk, j, i, I, s.data());
        P->x_cells += I.size();
      }
    }
  }
  if (!P->coarse) return MFM_OK;

  // ---- coarse-fine interpolation
  const int r = P->r;
  for (int d = 0; d < dims; ++d)
    if (fdiv(D.lo[d], r) * r != D.lo[d] || (D.hi[d] + 1) % r != 0)
      return pfail(P, MFM_ERR_SHAPE, "problem domain is not a refinement of a coarse domain");
  const HBox Dc = coarsen(D, r, dims);
  int Lc[3] = {1, 1, 1};
  for (int d = 0; d < dims; ++d) Lc[d] = L[d] / r;
  const auto shc = shifts(per, Lc, dims);
  for (int j = 0; j < ncb; ++j) {
    HBox b = hb(cboxes[j]);
    if (empty(b)) return pfail(P, MFM_ERR_SHAPE, "empty coarse box");
    for (int d = 0; d < dims; ++d)
      if (b.lo[d] < Dc.lo[d] || b.hi[d] > Dc.hi[d])
        return pfail(P, MFM_ERR_SHAPE, "coarse box outside the coarse domain");
    P->crse.push_back(b);
  }
  BoxIndex CI;
  CI.build(P->crse, dims);
  for (int j = 0; j < ncb; ++j) {
    CI.query(P->crse[j], hits);
    if (hits.size() != 1) return pfail(P, MFM_ERR_SHAPE, "coarse boxes overlap");
  }
  auto clipc = [&](HBox b) {  // ProblemDomain & Box: non-periodic directions
    for (int d = 0; d < dims; ++d)
      if (!per[d]) { b.lo[d] = std::max(b.lo[d], Dc.lo[d]); b.hi[d] = std::min(b.hi[d], Dc.hi[d]); }
    return b;
  };
  auto clipf = [&](HBox b) {
    for (int d = 0; d < dims; ++d)
      if (!per[d]) { b.lo[d] = std::max(b.lo[d], D.lo[d]); b.hi[d] = std::min(b.hi[d], D.hi[d]); }
    return b;
  };
  const int csr = (P->ng + r - 1) / r;  // coarse slope radius
  std::vector<uint8_t> cov, fcov, need;
  std::vector<int> own, own_sh;
  for (int i = 0; i < n; ++i) {
    const HBox cfb = coarsen(P->fine[i], r, dims);
    const HBox W = grow(cfb, csr + 1, dims);    // work FAB
    const HBox S = grow(cfb, csr, dims);        // slope box
    const HBox bmod = grow(clipc(grow(S, 1, dims)), -1, dims);  // limited cells
    const HBox want = clipc(coarsen(grow(P->fine[i], P->ng, dims), r, dims));
    const HBox G = clipf(grow(P->fine[i], P->ng, dims));
    const int wn[3] = {(int)W.n(0), (int)W.n(1), (int)W.n(2)};
    const long nw = W.size();
    FillPatch f;
    for (int d = 0; d < 3; ++d) {
      f.wlo[d] = W.lo[d]; f.wn[d] = wn[d]; f.glo[d] = G.lo[d]; f.ghi[d] = G.hi[d];
    }
    f.woff = (long)P->work_doubles;
    P->work_doubles += (size_t)nw * P->nc;
    P->fp.push_back(f);
    auto widx = [&](int a, int b, int c) {
      return (long)(a - W.lo[0]) + (long)wn[0] * ((long)(b - W.lo[1]) + (long)wn[1] * (c - W.lo[2]));
    };
    auto paint = [&](std::vector<uint8_t>& m, const HBox& b) {
      const HBox I = inter(b, W);
      if (empty(I)) return;
      for (int c = I.lo[2]; c <= I.hi[2]; ++c)
        for (int bb = I.lo[1]; bb <= I.hi[1]; ++bb)
          for (int a = I.lo[0]; a <= I.hi[0]; ++a) m[widx(a, bb, c)] = 1;
    };
    // coarse coverage (valid data of the coarse level or its periodic
    // images) and which coarse box / image holds each work cell
    cov.assign(nw, 0);
    own.assign(nw, -1);
    own_sh.assign(nw, 0);
    for (size_t si = 0; si < shc.size(); ++si) {
      const auto& s = shc[si];
      const int ms[3] = {-s[0], -s[1], -s[2]};
      CI.query(shift(W, ms), hits);
      for (int j : hits) {
        const HBox B = inter(shift(P->crse[j], s.data()), W);
        if (empty(B)) continue;
        for (int c = B.lo[2]; c <= B.hi[2]; ++c)
          for (int bb = B.lo[1]; bb <= B.hi[1]; ++bb)
            for (int a = B.lo[0]; a <= B.hi[0]; ++a) {
              const long w = widx(a, bb, c);
              cov[w] = 1;
              own[w] = j;
              own_sh[w] = (int)si;
            }
      }
    }
    const size_t first_cell = P->cells.size();
    // cells under other fine boxes (or their images) are filled by exchange
    fcov.assign(nw, 0);
    const HBox Wf = refine(W, r, dims);
    for (const auto& s : sh) {
      const int ms[3] = {-s[0], -s[1], -s[2]};
      FI.query(shift(Wf, ms), hits);
      for (int j : hits) {
        int sc[3];
        for (int d = 0; d < 3; ++d) sc[d] = s[d] / (d < dims ? r : 1);
        paint(fcov, shift(coarsen(P->fine[j], r, dims), sc));
      }
    }
    // interpolation cells, their stencils, and the coverage they need
    const int st[3] = {1, wn[0], wn[0] * wn[1]};
    for (int c = W.lo[2]; c <= W.hi[2]; ++c)
      for (int b = W.lo[1]; b <= W.hi[1]; ++b)
        for (int a = W.lo[0]; a <= W.hi[0]; ++a) {
          const int iv[3] = {a, b, c};
          bool in = true;
          for (int d = 0; d < 3; ++d) in = in && iv[d] >= want.lo[d] && iv[d] <= want.hi[d];
          const long w = widx(a, b, c);
          if (!in || fcov[w]) continue;
          for (int d = 0; d < dims; ++d)  // every stencil and neighbour read stays in W
            if (iv[d] - 1 < W.lo[d] || iv[d] + 1 > W.hi[d])
              return pfail(P, MFM_ERR_SHAPE, "internal: interpolation stencil leaves the work box");
          if (!cov[w])
            return pfail(P, MFM_ERR_UNSUPPORTED,
                         "coarse level does not cover a coarse cell under a fine ghost cell "
                         "(layouts not properly nested)");
          uint8_t fl = 0;
          for (int d = 0; d < dims; ++d) {
            // the stencil stays inside W: want lies in the slope box S
            const bool lo = !cov[w + st[d]], hi = !cov[w - st[d]];
            if (lo && hi)
              return pfail(P, MFM_ERR_UNSUPPORTED,
                           "coarse level is one cell wide under a fine ghost cell");
            fl |= (uint8_t)((hi ? mfm::kSlopeHi : lo ? mfm::kSlopeLo : mfm::kSlopeCentred) << (2 * d));
          }
          bool lim = true;
          for (int d = 0; d < 3; ++d) lim = lim && iv[d] >= bmod.lo[d] && iv[d] <= bmod.hi[d];
          if (lim) {
            fl |= mfm::kFillLimit;
            const int kz = dims == 3 ? 1 : 0;
            for (int kk = -kz; kk <= kz; ++kk)
              for (int jj = -1; jj <= 1; ++jj)
                for (int ii = -1; ii <= 1; ++ii)
                  if (!cov[w + ii + jj * st[1] + kk * st[2]])
                    return pfail(P, MFM_ERR_UNSUPPORTED,
                                 "coarse level does not cover the limiter neighbourhood "
                                 "(layouts not properly nested)");
          }
          long kids = 1;
          for (int d = 0; d < dims; ++d)
            kids *= std::max(0, std::min(iv[d] * r + r - 1, G.hi[d]) - std::max(iv[d] * r, G.lo[d]) + 1);
          if (kids == 0) continue;
          P->f_cells += kids;
          P->cells.push_back(make_int2(i, (int)w));
          P->flags.push_back(fl);
        }
    // gather exactly the covered cells in the neighbourhoods of this box's
    // interpolation cells (what the slopes and the limiter read)
    need.assign(nw, 0);
    const int kz = dims == 3 ? 1 : 0;
    for (size_t q = first_cell; q < P->cells.size(); ++q) {
      const long w = P->cells[q].y;
      for (int kk = -kz; kk <= kz; ++kk)
        for (int jj = -1; jj <= 1; ++jj)
          for (int ii = -1; ii <= 1; ++ii) need[w + ii + jj * st[1] + kk * st[2]] = 1;
    }
    for (int c = W.lo[2]; c <= W.hi[2]; ++c)
      for (int b = W.lo[1]; b <= W.hi[1]; ++b)
        for (int a = W.lo[0]; a <= W.hi[0]; ++a) {
          const long w = widx(a, b, c);
          if (!need[w] || !cov[w]) continue;
          const auto& s = shc[own_sh[w]];
          GatherCell g;
          g.src = own[w];
          g.dst = i;
          g.doff = (int)w;
          g.s[0] = a - s[0];
          g.s[1] = b - s[1];
          g.s[2] = c - s[2];
          P->gc.push_back(g);
        }
  }
  return MFM_OK;
}

int upload(mfm_ctx* c, mfm_fill_plan* P) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipGetDevice");
  if (P->dmem) {
    if (dev != P->device) return mfm::host::fail(c, MFM_ERR_ARG, "fill plan used on another device");
    return MFM_OK;
  }
  using mfm::host::align256;
  const size_t b_xr = align256(P->xr.size() * sizeof(CopyRect));
  const size_t b_gc = align256(P->gc.size() * sizeof(GatherCell));
  const size_t b_cells = align256(P->cells.size() * sizeof(int2));
  const size_t b_flags = align256(P->flags.size());
  const size_t b_fp = align256(P->fp.size() * sizeof(FillPatch));
  const size_t b_wfab = align256(P->fp.size() * sizeof(FillFab));
  const size_t b_work = align256(P->work_doubles * sizeof(double));
  const size_t total = b_xr + b_gc + b_cells + b_flags + b_fp + b_wfab + b_work + 256;
  e = hipMalloc((void**)&P->dmem, total);
  if (e != hipSuccess) { P->dmem = nullptr; return mfm::host::fail(c, MFM_ERR_NOMEM, "hipMalloc of the fill plan failed"); }
  P->device = dev;
  P->dbytes = total;
  char* q = P->dmem;
  P->d_xr = (CopyRect*)q; q += b_xr;
  P->d_gc = (GatherCell*)q; q += b_gc;
  P->d_cells = (int2*)q; q += b_cells;
  P->d_flags = (uint8_t*)q; q += b_flags;
  P->d_fp = (FillPatch*)q; q += b_fp;
  P->d_wfab = (FillFab*)q; q += b_wfab;
  P->d_work = (double*)q;
  std::vector<FillFab> wf(P->fp.size());
  for (size_t i = 0; i < P->fp.size(); ++i) {
    const FillPatch& f = P->fp[i];
    wf[i].p = P->d_work + f.woff;
    for (int d = 0; d < 3; ++d) { wf[i].lo[d] = f.wlo[d]; wf[i].n[d] = f.wn[d]; }
    wf[i].cs = (long)f.wn[0] * f.wn[1] * f.wn[2];
  }
  struct Up { void* dst; const void* src; size_t bytes; } ups[] = {
      {P->d_xr, P->xr.data(), P->xr.size() * sizeof(CopyRect)},
      {P->d_gc, P->gc.data(), P->gc.size() * sizeof(GatherCell)},
      {P->d_cells, P->cells.data(), P->cells.size() * sizeof(int2)},
      {P->d_flags, P->flags.data(), P->flags.size()},
      {P->d_fp, P->fp.data(), P->fp.size() * sizeof(FillPatch)},
      {P->d_wfab, wf.data(), wf.size() * sizeof(FillFab)}};
  for (const Up& u : ups) {
    if (!u.bytes) continue;
    e = hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return mfm::host::hipfail(c, e, "fill plan upload");
  }
  return MFM_OK;
}

int to_fillfab(mfm_ctx* c, const mfm_fab& f, int nc, const HBox& need, int dims,
               FillFab& o, const char* what) {
  using mfm::host::hb;
  if (!f.data) return mfm::host::fail(c, MFM_ERR_ARG, std::string(what) + ": null data");
  if (f.ncomp != nc) return mfm::host::fail(c, MFM_ERR_SHAPE, std::string(what) + ": ncomp mismatch");
  const HBox b = hb(f.box);
  HBox nd = need;
  for (int d = dims; d < 3; ++d) { nd.lo[d] = b.lo[d]; nd.hi[d] = b.hi[d]; }
  if (!mfm::host::contains(b, nd))
    return mfm::host::fail(c, MFM_ERR_SHAPE, std::string(what) + ": FAB does not hold the cells the fill touches");
  if (b.size() >= (1L << 31)) return mfm::host::fail(c, MFM_ERR_SHAPE, std::string(what) + ": FAB too large");
  o.p = f.data;
  for (int d = 0; d < 3; ++d) { o.lo[d] = b.lo[d]; o.n[d] = (int)b.n(d); }
  o.cs = b.size();
  return MFM_OK;
}

// Stage the per-call FAB tables into slot k: fine, then coarse new, then
// coarse old (absent ones as empty entries).
int stage_tables(mfm_ctx* c, mfm_fill_plan* P, int k, const mfm_fab* U,
                 const mfm_fab* cn, const mfm_fab* co, hipStream_t s) {
  const int n = (int)P->fine.size(), m = (int)P->crse.size();
  std::vector<FillFab> t((size_t)n + (k == 1 ? 2 * (size_t)m : 0));
  std::memset(t.data(), 0, t.size() * sizeof(FillFab));
  for (int i = 0; i < n; ++i) {
    int rc = to_fillfab(c, U[i], P->nc, grow(P->fine[i], P->ng, P->dims), P->dims, t[i], "U");
    if (rc) return rc;
  }
  if (k == 1)
    for (int j = 0; j < m; ++j) {
      if (cn) { int rc = to_fillfab(c, cn[j], P->nc, P->crse[j], P->dims, t[n + j], "coarse_new"); if (rc) return rc; }
      if (co) { int rc = to_fillfab(c, co[j], P->nc, P->crse[j], P->dims, t[n + m + j], "coarse_old"); if (rc) return rc; }
    }
  return slot_put(c, P->tab[k], t, s);
}

}  // namespace

extern "C" {

int mfm_fill_plan_create(const mfm_params* p, const mfm_box* boxes, int n,
                         const mfm_box* coarse_boxes, int ncoarse, int ref_ratio,
                         int interp_radius, double dx_coarse, mfm_fill_plan** out) {
  if (!out) return MFM_ERR_ARG;
  *out = nullptr;
  if (!p || (!boxes && n > 0) || n < 0 || ncoarse < 0 || (ncoarse > 0 && !coarse_boxes))
    return MFM_ERR_ARG;
  mfm::Consts C;
  std::string why;
  int rc = mfm::host::build_consts(p, C, why);
  mfm_fill_plan* P = new mfm_fill_plan();
  *out = P;
  if (rc) return pfail(P, rc, why);
  if (interp_radius < 1) return pfail(P, MFM_ERR_ARG, "interp_radius must be >= 1");
  P->dims = p->spacedim;
  P->nc = C.nc;
  P->ng = interp_radius;
  P->coarse = ncoarse > 0;
  if (P->coarse) {
    if (ref_ratio < 2) return pfail(P, MFM_ERR_ARG, "ref_ratio must be >= 2");
    if (!(dx_coarse > 0.0)) return pfail(P, MFM_ERR_ARG, "dx_coarse must be > 0");
    P->r = ref_ratio;
    P->dxc = dx_coarse;
    P->dxf = dx_coarse / ref_ratio;  // CSHCartesian::define (src/CSHCartesian.cpp:50-52)
    // coeff = volc / (volf refScale), src/CSHCartesian.cpp:90-97 and
    // PiecewiseLinearFillPatchMHDAM.cpp:820-852
    double refScale = 1.0, volc = P->dxc, volf = P->dxf;
    for (int d = 0; d < P->dims; ++d) refScale *= ref_ratio;
    for (int d = 1; d < P->dims; ++d) { volc = volc * P->dxc; volf = volf * P->dxf; }
    P->coeff = volc / (volf * refScale);
    // the 1.e-9 of interplimit_v is a default-real literal
    P->tiny = p->ce_single_literals ? (double)1.0e-9f : 1.0e-9;
    // the div(B) slope correction is set up with GLM cleaning
    // (srcIdeal/AMRLevelIdealMHD.cpp:388-394)
    P->ibx = p->dedner ? 5 : -1;
  }
  rc = build(P, p, boxes, n, coarse_boxes, ncoarse);
  if (rc == MFM_OK && P->cells.size() >= (size_t)1 << 31)
    rc = pfail(P, MFM_ERR_UNSUPPORTED, "too many interpolation cells");
  P->ok = rc == MFM_OK;
  return rc;
}

int mfm_fill_plan_destroy(mfm_fill_plan* P) {
  if (!P) return MFM_OK;
  if (P->dmem) {
    (void)hipDeviceSynchronize();
    (void)hipFree(P->dmem);
  }
  for (TableSlot& t : P->tab) t.release();
  delete P;
  return MFM_OK;
}

const char* mfm_fill_plan_last_error(const mfm_fill_plan* P) {
  return P ? P->err.c_str() : "null plan";
}

int mfm_fill_plan_info(const mfm_fill_plan* P, long* xc, long* fc, long* cc,
                       size_t* bytes) {
  if (!P) return MFM_ERR_ARG;
  if (xc) *xc = P->x_cells;
  if (fc) *fc = P->f_cells;
  if (cc) *cc = (long)P->cells.size();
  if (bytes) {
    using mfm::host::align256;
    *bytes = align256(P->xr.size() * sizeof(CopyRect)) + align256(P->gc.size() * sizeof(GatherCell)) +
             align256(P->cells.size() * sizeof(int2)) + align256(P->flags.size()) +
             align256(P->fp.size() * sizeof(FillPatch)) + align256(P->fp.size() * sizeof(FillFab)) +
             align256(P->work_doubles * sizeof(double)) + 256;
  }
  return MFM_OK;
}

int mfm_level_exchange_dev(mfm_ctx* c, mfm_fill_plan* P, const mfm_fab* U,
                           void* stream) {
  if (!c || !P || (!U && !P->fine.empty())) return MFM_ERR_ARG;
  if (!P->ok) return mfm::host::fail(c, MFM_ERR_ARG, "fill plan was not built");
  if (P->nc != c->C.nc) return mfm::host::fail(c, MFM_ERR_SHAPE, "plan and context differ in ncomp");
  if (P->xr.empty()) return MFM_OK;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(c, P);
  if (rc) return rc;
  rc = stage_tables(c, P, 0, U, nullptr, nullptr, s);
  if (rc) return rc;
  const FillFab* T = P->tab[0].d;
  hipError_t e = mfm::launch_rect_copy(P->d_xr, (int)P->xr.size(), P->xr_blk, T, T, T, P->nc,
                                       1.0, 0.0, 0, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "exchange launch");
  return MFM_OK;
}

int mfm_level_fill_interp_dev(mfm_ctx* c, mfm_fill_plan* P, const mfm_fab* U,
                              const mfm_fab* cold, const mfm_fab* cnew,
                              double alpha, void* stream) {
  if (!c || !P || (!U && !P->fine.empty())) return MFM_ERR_ARG;
  if (!P->ok) return mfm::host::fail(c, MFM_ERR_ARG, "fill plan was not built");
  if (!P->coarse) return mfm::host::fail(c, MFM_ERR_ARG, "plan has no coarse level");
  if (P->nc != c->C.nc) return mfm::host::fail(c, MFM_ERR_SHAPE, "plan and context differ in ncomp");
  if (!cold && !cnew)
    return mfm::host::fail(c, MFM_ERR_ARG, "no old coarse data and no new coarse data");
  if (!(alpha >= 0.0 && alpha <= 1.0))
    return mfm::host::fail(c, MFM_ERR_ARG, "time interpolation coefficient outside [0, 1]");
  // timeInterp's cases (srcIdeal/PiecewiseLinearFillPatchMHDAM.cpp:447-477)
  int mode = 2;
  if (alpha == 1.0 || !cold) mode = 0;
  else if (alpha == 0.0 || !cnew) mode = 1;
  if (P->cells.empty()) return MFM_OK;
  hipStream_t s = (hipStream_t)stream;
  int rc = upload(c, P);
  if (rc) return rc;
  rc = stage_tables(c, P, 1, U, mode == 1 ? nullptr : cnew, mode == 0 ? nullptr : cold, s);
  if (rc) return rc;
  const int n = (int)P->fine.size(), m = (int)P->crse.size();
  const FillFab* T = P->tab[1].d;
  hipError_t e = mfm::launch_cell_copy(P->d_gc, (long)P->gc.size(), T + n, T + n + m,
                                       P->d_wfab, P->nc, alpha,
                                       1.0 - alpha, mode, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "coarse gather launch");
  mfm::InterpArgs A;
  A.cell = P->d_cells;
  A.flag = P->d_flags;
  A.ncell = (long)P->cells.size();
  A.P = P->d_fp;
  A.work = P->d_work;
  A.U = T;
  A.nc = P->nc;
  A.dims = P->dims;
  A.r = P->r;
  A.ibx = P->ibx;
  A.c_centre = 0.5 / P->dxc;
  A.c_side = 1.0 / P->dxc;
  A.dxc = P->dxc;
  A.dxf = P->dxf;
  A.coeff = P->coeff;
  A.tiny = P->tiny;
  e = mfm::launch_fill_interp(A, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "interpolation launch");
  return MFM_OK;
}

int mfm_fill_time_coef(double time, double tnew, double dtc, double dtf,
                       int ref_ratio, double* alpha) {
  if (!alpha || ref_ratio < 1 || !(dtc > 0.0)) return MFM_ERR_ARG;
  const double told = tnew - dtc;
  double a = (time - told) / (tnew - told);
  const double eps = 0.04 * dtf / ref_ratio;
  if (std::fabs(a) < eps) a = 0.0;
  if (std::fabs(1.0 - a) < eps) a = 1.0;
  if (!(a >= 0.0) || a > 1.0) return MFM_ERR_ARG;
  *alpha = a;
  return MFM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Coarse-fine synchronisation of a level pair: the flux register of the
// interface (LevelFluxRegister of Chombo 3.2.4 as MHDAM drives it:
// setToZero at the coarse step, incrementCoarse / incrementFine with dt,
// reflux with scale -1 and 1/dx, srcIdeal/AMRLevelIdealMHD.cpp:695,
// 1282-1307, 1641) and the average down (CoarseAverageExt,
// src/CoarseAverageExt.cpp:126-194).
//
// Register faces are the coarse faces on the boundary of the coarsened fine
// region that do not lie on a non-periodic domain boundary; each is kept by
// the coarse cell outside the fine region (after periodic wrapping), so
// reflux touches only cells of the coarse level.  After the increments a
// face holds dF = sum over fine steps of dt_f <F_f> - dt_c F_c, <F_f> the
// mean of the r^(d-1) fine faces covering it; reflux with scale -1 adds
// -sigma dF / dx to its cell (sigma = +1 on the cell's high face), which
// replaces the coarse flux through the interface by the fine fluxes.

using mfm::AvgRect;
using mfm::RegCell;
using mfm::RegFace;

struct mfm_flux_register {
  int dims = 3, nc = 0, r = 2;
  bool ok = false;
  std::string err;
  std::vector<HBox> fine, crse;
  std::vector<RegFace> faces;
  std::vector<RegCell> cells;
  std::vector<AvgRect> avg;
  int avg_blk = 0;
  long avg_cells = 0;
  int device = -1;
  char* dmem = nullptr;
  RegFace* d_faces = nullptr;
  RegCell* d_cells = nullptr;
  AvgRect* d_avg = nullptr;
  double* d_reg = nullptr;
  TableSlot tab[4];  // coarse fluxes, fine fluxes, coarse U, coarse + fine U
};

namespace {

int rfail(mfm_flux_register* R, int code, const std::string& m) {
  R->err = m;
  return code;
}

int build_register(mfm_flux_register* R, const mfm_params* p, const mfm_box* fb, int nf,
                   const mfm_box* cbx, int ncb) {
  using mfm::host::hb;
  const int dims = R->dims, r = R->r;
  HBox D;
  int per[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    D.lo[d] = p->domain_lo[d];
    D.hi[d] = p->domain_hi[d];
    if (d < dims) per[d] = p->periodic[d] != 0;
  }
  if (empty(D)) return rfail(R, MFM_ERR_ARG, "empty problem domain");
  for (int d = 0; d < dims; ++d)
    if (fdiv(D.lo[d], r) * r != D.lo[d] || (D.hi[d] + 1) % r != 0)
      return rfail(R, MFM_ERR_SHAPE, "problem domain is not a refinement of a coarse domain");
  const HBox Dc = coarsen(D, r, dims);
  int Lc[3] = {1, 1, 1};
  for (int d = 0; d < dims; ++d) Lc[d] = (int)Dc.n(d);
  std::vector<HBox> cf;
  for (int i = 0; i < nf; ++i) {
    const HBox b = hb(fb[i]);
    if (empty(b)) return rfail(R, MFM_ERR_SHAPE, "empty fine box");
    for (int d = 0; d < dims; ++d) {
      if (b.lo[d] < D.lo[d] || b.hi[d] > D.hi[d])
        return rfail(R, MFM_ERR_SHAPE, "fine box outside the problem domain");
      if (fdiv(b.lo[d], r) * r != b.lo[d] || (b.hi[d] + 1) % r != 0)
        return rfail(R, MFM_ERR_SHAPE, "fine box is not coarsenable by the refinement ratio");
    }
    R->fine.push_back(b);
    cf.push_back(coarsen(b, r, dims));
  }
  for (int j = 0; j < ncb; ++j) {
    const HBox b = hb(cbx[j]);
    if (empty(b)) return rfail(R, MFM_ERR_SHAPE, "empty coarse box");
    for (int d = 0; d < dims; ++d)
      if (b.lo[d] < Dc.lo[d] || b.hi[d] > Dc.hi[d])
        return rfail(R, MFM_ERR_SHAPE, "coarse box outside the coarse domain");
    R->crse.push_back(b);
  }
  BoxIndex FI, CI;
  FI.build(cf, dims);
  CI.build(R->crse, dims);
  std::vector<int> hits;
  for (int i = 0; i < nf; ++i) {
    FI.query(cf[i], hits);
    if (hits.size() != 1) return rfail(R, MFM_ERR_SHAPE, "fine boxes overlap");
  }
  for (int j = 0; j < ncb; ++j) {
    CI.query(R->crse[j], hits);
    if (hits.size() != 1) return rfail(R, MFM_ERR_SHAPE, "coarse boxes overlap");
  }
  auto unit = [](const int c[3]) {
    HBox b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = c[d]; b.hi[d] = c[d]; }
    return b;
  };
  // ---- register faces, one pass over the faces of every coarsened fine box
  for (int f = 0; f < nf; ++f)
    for (int d = 0; d < dims; ++d)
      for (int side = 0; side < 2; ++side) {
        HBox plane = cf[f];
        const int cd = side == 0 ? cf[f].lo[d] - 1 : cf[f].hi[d] + 1;  // cell across
        plane.lo[d] = plane.hi[d] = cd;
        for (int k = plane.lo[2]; k <= plane.hi[2]; ++k)
          for (int j = plane.lo[1]; j <= plane.hi[1]; ++j)
            for (int i = plane.lo[0]; i <= plane.hi[0]; ++i) {
              int C[3] = {i, j, k};
              if (C[d] < Dc.lo[d] || C[d] > Dc.hi[d]) {
                if (!per[d]) continue;  // physical boundary: no register
                C[d] += C[d] < Dc.lo[d] ? Lc[d] : -Lc[d];
              }
              FI.query(unit(C), hits);
              if (!hits.empty()) continue;  // interior face of the fine region
              CI.query(unit(C), hits);
              if (hits.empty())
                return rfail(R, MFM_ERR_UNSUPPORTED,
                             "no coarse box next to the fine region (layouts not properly nested)");
              RegFace e;
              e.cb = hits[0];
              e.d = d;
              e.sigma = side == 0 ? +1 : -1;
              e.fb = f;
              const int t[3] = {i, j, k};
              for (int q = 0; q < 3; ++q) {
                e.c[q] = C[q];
                e.f0[q] = q < dims ? t[q] * r : t[q];
              }
              e.f0[d] = side == 0 ? R->fine[f].lo[d] : R->fine[f].hi[d] + 1;
              R->faces.push_back(e);
            }
      }
  // group by coarse cell: box, cell (x fastest), then direction, low face first
  std::sort(R->faces.begin(), R->faces.end(), [](const RegFace& a, const RegFace& b) {
    if (a.cb != b.cb) return a.cb < b.cb;
    for (int q = 2; q >= 0; --q)
      if (a.c[q] != b.c[q]) return a.c[q] < b.c[q];
    if (a.d != b.d) return a.d < b.d;
    return a.sigma < b.sigma;
  });
  for (size_t e = 0; e < R->faces.size();) {
    size_t e1 = e + 1;
    while (e1 < R->faces.size() && R->faces[e1].cb == R->faces[e].cb &&
           R->faces[e1].c[0] == R->faces[e].c[0] && R->faces[e1].c[1] == R->faces[e].c[1] &&
           R->faces[e1].c[2] == R->faces[e].c[2])
      ++e1;
    R->cells.push_back(RegCell{(int)e, (int)e1});
    e = e1;
  }
  // ---- average down: coarsened fine boxes onto the coarse boxes
  for (int f = 0; f < nf; ++f) {
    CI.query(cf[f], hits);
    for (int cb : hits) {
      const HBox I = inter(cf[f], R->crse[cb]);
      if (empty(I)) continue;
      AvgRect a;
      a.fb = f;
      a.cb = cb;
      for (int d = 0; d < 3; ++d) { a.lo[d] = I.lo[d]; a.n[d] = (int)I.n(d); }
      a.blk0 = R->avg_blk;
      R->avg_blk += (int)((I.size() + kFB - 1) / kFB);
      R->avg_cells += I.size();
      R->avg.push_back(a);
    }
  }
  return MFM_OK;
}

int reg_upload(mfm_ctx* c, mfm_flux_register* R) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "hipGetDevice");
  if (R->dmem) {
    if (dev != R->device) return mfm::host::fail(c, MFM_ERR_ARG, "flux register used on another device");
    return MFM_OK;
  }
  using mfm::host::align256;
  const size_t b_f = align256(R->faces.size() * sizeof(RegFace));
  const size_t b_c = align256(R->cells.size() * sizeof(RegCell));
  const size_t b_a = align256(R->avg.size() * sizeof(AvgRect));
  const size_t b_r = align256(R->faces.size() * (size_t)R->nc * sizeof(double));
  e = hipMalloc((void**)&R->dmem, b_f + b_c + b_a + b_r + 256);
  if (e != hipSuccess) { R->dmem = nullptr; return mfm::host::fail(c, MFM_ERR_NOMEM, "hipMalloc of the flux register failed"); }
  R->device = dev;
  char* q = R->dmem;
  R->d_faces = (RegFace*)q; q += b_f;
  R->d_cells = (RegCell*)q; q += b_c;
  R->d_avg = (AvgRect*)q; q += b_a;
  R->d_reg = (double*)q;
  struct Up { void* dst; const void* src; size_t bytes; } ups[] = {
      {R->d_faces, R->faces.data(), R->faces.size() * sizeof(RegFace)},
      {R->d_cells, R->cells.data(), R->cells.size() * sizeof(RegCell)},
      {R->d_avg, R->avg.data(), R->avg.size() * sizeof(AvgRect)}};
  for (const Up& u : ups) {
    if (!u.bytes) continue;
    e = hipMemcpy(u.dst, u.src, u.bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return mfm::host::hipfail(c, e, "flux register upload");
  }
  e = mfm::launch_reg_zero(R->d_reg, (long)R->faces.size() * R->nc, nullptr);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "flux register clear");
  return MFM_OK;
}

// Stage n FABs (each checked to hold `need[i]`) into table slot k.
int reg_stage(mfm_ctx* c, mfm_flux_register* R, int k, const mfm_fab* F,
              const std::vector<HBox>& need, const char* what, hipStream_t s) {
  std::vector<FillFab> t(need.size());
  for (size_t i = 0; i < need.size(); ++i) {
    int rc = to_fillfab(c, F[i], R->nc, need[i], R->dims, t[i], what);
    if (rc) return rc;
  }
  return slot_put(c, R->tab[k], t, s);
}

std::vector<HBox> face_boxes(const std::vector<HBox>& boxes, int dims) {
  std::vector<HBox> v;
  for (const HBox& b : boxes)
    for (int d = 0; d < dims; ++d) {
      HBox f = b;
      f.hi[d] += 1;  // surroundingNodes(d)
      v.push_back(f);
    }
  return v;
}

int reg_check(mfm_ctx* c, mfm_flux_register* R) {
  if (!c || !R) return MFM_ERR_ARG;
  if (!R->ok) return mfm::host::fail(c, MFM_ERR_ARG, "flux register was not built");
  if (R->nc != c->C.nc) return mfm::host::fail(c, MFM_ERR_SHAPE, "flux register and context differ in ncomp");
  return reg_upload(c, R);
}

}  // namespace

extern "C" {

int mfm_flux_register_create(const mfm_params* p, const mfm_box* fine, int nf,
                             const mfm_box* coarse, int ncb, int ref_ratio,
                             mfm_flux_register** out) {
  if (!out) return MFM_ERR_ARG;
  *out = nullptr;
  if (!p || nf < 0 || ncb < 0 || (nf > 0 && !fine) || (ncb > 0 && !coarse)) return MFM_ERR_ARG;
  mfm_flux_register* R = new mfm_flux_register();
  *out = R;
  mfm::Consts C;
  std::string why;
  int rc = mfm::host::build_consts(p, C, why);
  if (rc) return rfail(R, rc, why);
  if (ref_ratio < 2) return rfail(R, MFM_ERR_ARG, "ref_ratio must be >= 2");
  R->dims = p->spacedim;
  R->nc = C.nc;
  R->r = ref_ratio;
  rc = build_register(R, p, fine, nf, coarse, ncb);
  R->ok = rc == MFM_OK;
  return rc;
}

int mfm_flux_register_destroy(mfm_flux_register* R) {
  if (!R) return MFM_OK;
  if (R->dmem) {
    (void)hipDeviceSynchronize();
    (void)hipFree(R->dmem);
  }
  for (TableSlot& t : R->tab) t.release();
  delete R;
  return MFM_OK;
}

const char* mfm_flux_register_last_error(const mfm_flux_register* R) {
  return R ? R->err.c_str() : "null flux register";
}

int mfm_flux_register_info(const mfm_flux_register* R, long* faces, long* cells,
                           long* average_cells) {
  if (!R) return MFM_ERR_ARG;
  if (faces) *faces = (long)R->faces.size();
  if (cells) *cells = (long)R->cells.size();
  if (average_cells) *average_cells = R->avg_cells;
  return MFM_OK;
}

int mfm_flux_register_set_zero(mfm_ctx* c, mfm_flux_register* R, void* stream) {
  int rc = reg_check(c, R);
  if (rc) return rc;
  hipError_t e = mfm::launch_reg_zero(R->d_reg, (long)R->faces.size() * R->nc, (hipStream_t)stream);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "flux register clear");
  return MFM_OK;
}

int mfm_flux_register_increment_coarse(mfm_ctx* c, mfm_flux_register* R, const mfm_fab* F,
                                       double scale, void* stream) {
  int rc = reg_check(c, R);
  if (rc) return rc;
  if (!F && !R->crse.empty()) return mfm::host::fail(c, MFM_ERR_ARG, "null flux table");
  hipStream_t s = (hipStream_t)stream;
  rc = reg_stage(c, R, 0, F, face_boxes(R->crse, R->dims), "coarse flux", s);
  if (rc) return rc;
  hipError_t e = mfm::launch_reg_coarse(R->d_faces, (long)R->faces.size(), R->tab[0].d, R->dims,
                                        R->d_reg, R->nc, scale, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "flux register increment");
  return MFM_OK;
}

int mfm_flux_register_increment_fine(mfm_ctx* c, mfm_flux_register* R, const mfm_fab* F,
                                     double scale, void* stream) {
  int rc = reg_check(c, R);
  if (rc) return rc;
  if (!F && !R->fine.empty()) return mfm::host::fail(c, MFM_ERR_ARG, "null flux table");
  hipStream_t s = (hipStream_t)stream;
  rc = reg_stage(c, R, 1, F, face_boxes(R->fine, R->dims), "fine flux", s);
  if (rc) return rc;
  hipError_t e = mfm::launch_reg_fine(R->d_faces, (long)R->faces.size(), R->tab[1].d, R->dims, R->r,
                                      R->d_reg, R->nc, scale, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "flux register increment");
  return MFM_OK;
}

int mfm_flux_register_reflux(mfm_ctx* c, mfm_flux_register* R, const mfm_fab* U,
                             double scale, double dx_coarse, void* stream) {
  int rc = reg_check(c, R);
  if (rc) return rc;
  if (!U && !R->crse.empty()) return mfm::host::fail(c, MFM_ERR_ARG, "null state table");
  if (!(dx_coarse > 0.0)) return mfm::host::fail(c, MFM_ERR_ARG, "dx_coarse must be > 0");
  hipStream_t s = (hipStream_t)stream;
  rc = reg_stage(c, R, 2, U, R->crse, "coarse U", s);
  if (rc) return rc;
  // the inverse cell volume of the reference's Cartesian scalingFactor
  // (src/CSHCartesian.cpp:290-305) is 1/dx
  hipError_t e = mfm::launch_reflux(R->d_cells, (long)R->cells.size(), R->d_faces,
                                    (long)R->faces.size(), R->tab[2].d, R->d_reg, R->nc, scale,
                                    1.0 / dx_coarse, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "reflux launch");
  return MFM_OK;
}

int mfm_level_average_down_dev(mfm_ctx* c, mfm_flux_register* R, const mfm_fab* Uc,
                               const mfm_fab* Uf, double dx_fine, void* stream) {
  int rc = reg_check(c, R);
  if (rc) return rc;
  if ((!Uc && !R->crse.empty()) || (!Uf && !R->fine.empty()))
    return mfm::host::fail(c, MFM_ERR_ARG, "null state table");
  if (!(dx_fine > 0.0)) return mfm::host::fail(c, MFM_ERR_ARG, "dx_fine must be > 0");
  hipStream_t s = (hipStream_t)stream;
  std::vector<mfm_fab> both(Uc, Uc + R->crse.size());
  both.insert(both.end(), Uf, Uf + R->fine.size());
  std::vector<HBox> need = R->crse;
  need.insert(need.end(), R->fine.begin(), R->fine.end());
  rc = reg_stage(c, R, 3, both.data(), need, "U", s);
  if (rc) return rc;
  // getCellVolumes of the fine level (src/CSHCartesian.cpp:90-97)
  double vol = dx_fine;
  for (int d = 1; d < R->dims; ++d) vol = vol * dx_fine;
  hipError_t e = mfm::launch_average_down(R->d_avg, (int)R->avg.size(), R->avg_blk, R->tab[3].d,
                                          R->tab[3].d + R->crse.size(), R->nc, R->dims, R->r, vol, s);
  if (e != hipSuccess) return mfm::host::hipfail(c, e, "average-down launch");
  return MFM_OK;
}

}  // extern "C"
