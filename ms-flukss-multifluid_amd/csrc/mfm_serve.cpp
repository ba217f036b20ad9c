// mfm_serve.cpp -- ensemble serving: independent solves of one equation
// class (dimension, fluids, GLM, Riemann solver family, time scheme,
// precision, characteristic or component-wise slopes) admitted by an HBM
// budget and advanced together, one RK2 step per iteration, in the same
// batched launches (continuous batching).  Each request keeps its state on
// the device from admission until the caller fetches and releases it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/mfm.h"
#include "mfm_host.hpp"

using namespace mfm;
using namespace mfm::host;

namespace {

enum { REQ_PENDING = 0, REQ_ACTIVE = 1, REQ_DONE = 2, REQ_FAILED = 3 };

struct Request {
  long id = 0;
  mfm_params p;
  Consts C;
  HBox valid, fab;
  double dx = 0.0, cfl = 0.0, cfl0 = 0.0, growth = 0.0, t = 0.0;
  double dt_raw = 0.0, dt_prev = 0.0;  // last CFL-free dt, last step taken
  int nsteps = 0, done = 0, state = REQ_PENDING;
  double* U = nullptr;
  size_t bytes = 0;  // state + the update scratch charged to the budget
  std::string err;
};

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct mfm_server {
  mfm_ctx* ctx = nullptr;
  size_t budget = 0, in_use = 0;
  int max_active = 0;
  long next_id = 1;
  std::deque<Request*> pending;
  std::vector<Request*> active;
  std::map<long, Request*> all;
  std::vector<double> iter_ms;
  double cell_updates = 0.0, busy_ms = 0.0;
  std::string err;
};

namespace {

int sfail(mfm_server* s, int code, const std::string& m) {
  s->err = m;
  return code;
}

// Device bytes of one request: its FAB plus the scratch run_batch allocates
// for it (U*, W, face fluxes and normal B of the first-stage value box).
size_t request_bytes(const mfm_params& cls, int mode, const Consts& C, const HBox& fab,
                     const HBox& valid) {
  const int dims = C.dims;
  const size_t fs = cls.precision == MFM_PREC_FP32_FLUX ? 4 : 8;
  const size_t cells = (size_t)fab.size();
  const bool tiles = cls.time_approx != MFM_TIME_HANCOCK && !cls.char_reconstruction;
  size_t b = align256(cells * C.nc * 8);          // U
  b += align256(cells * C.nc * 8);                // U*
  b += align256(cells * C.nc * fs);               // W
  if (cls.time_approx == MFM_TIME_HANCOCK) b += align256(cells * C.nc * fs);  // W of U_pred
  if (tiles && mode == MFM_STAGE_FUSED && cls.time_approx == MFM_TIME_RK2)
    b += align256(cells * C.nc * fs);             // W of U*
  if (tiles && mode == MFM_STAGE_SUMMED && cls.precision == MFM_PREC_FP32_FLUX)
    b += align256(cells * (C.nc + 1) * fs);       // face sums
  HBox v1 = cls.time_approx == MFM_TIME_RK2 ? grow(valid, 2, dims)
            : (cls.time_approx == MFM_TIME_HANCOCK ? grow(valid, 1, dims) : valid);
  for (int d = 0; d < dims; ++d) {
    HBox f = v1;
    f.hi[d] += 1;
    b += align256((size_t)f.size() * C.nc * fs) + align256((size_t)f.size() * fs);
  }
  return b + 4096;  // tables
}

// Riemann solver family (the plasma flux kernel is compiled per family)
int solver(const mfm_params& p) { return flux_kind_of(p); }

bool same_class(const mfm_params& a, const mfm_params& b) {
  return a.spacedim == b.spacedim && a.coord_sys == b.coord_sys &&
         a.nfluids == b.nfluids && a.dedner == b.dedner &&
         solver(a) == solver(b) &&
         a.time_approx == b.time_approx && a.precision == b.precision &&
         (a.char_reconstruction != 0) == (b.char_reconstruction != 0);
}

// Admit pending requests in arrival order while the budget allows.
int admit(mfm_server* s) {
  while (!s->pending.empty()) {
    Request* r = s->pending.front();
    if ((int)s->active.size() >= s->max_active) break;
    if (s->in_use + r->bytes > s->budget) {
      if (s->active.empty()) {  // can never fit
        s->pending.pop_front();
        r->state = REQ_FAILED;
        r->err = "request exceeds the HBM budget";
        continue;
      }
      break;
    }
    s->pending.pop_front();
    const size_t ub = (size_t)r->fab.size() * r->C.nc * sizeof(double);
    hipError_t e = hipMalloc((void**)&r->U, ub);
    if (e != hipSuccess) {
      r->U = nullptr;
      (void)hipGetLastError();
      s->pending.push_front(r);
      if (s->active.empty()) {
        s->pending.pop_front();
        r->state = REQ_FAILED;
        r->err = "hipMalloc of the state failed";
        continue;
      }
      break;  // wait for memory to free up
    }
    // initial state (problem init on the whole FAB) and the first dt
    int ulo[3], un[3], lo[3], n[3];
    for (int d = 0; d < 3; ++d) {
      ulo[d] = r->fab.lo[d]; un[d] = (int)r->fab.n(d);
      lo[d] = r->fab.lo[d]; n[d] = (int)r->fab.n(d);
    }
    e = launch_init(r->C, r->U, ulo, un, r->fab.size(), lo, n, r->dx, nullptr);
    if (e != hipSuccess) return sfail(s, MFM_ERR_DEVICE, std::string("init launch: ") + hipGetErrorString(e));
    mfm_fab fab{r->U, {{r->fab.lo[0], r->fab.lo[1], r->fab.lo[2]}, {r->fab.hi[0], r->fab.hi[1], r->fab.hi[2]}}, r->C.nc};
    mfm_box vb{{r->valid.lo[0], r->valid.lo[1], r->valid.lo[2]}, {r->valid.hi[0], r->valid.hi[1], r->valid.hi[2]}};
    double dt0 = 0.0;
    int cell[3];
    int rc = run_dt(s->ctx, r->C, fab, vb, r->dx, &dt0, cell, nullptr);
    if (rc != MFM_OK) {
      r->state = REQ_FAILED;
      r->err = s->ctx->err;
      (void)hipFree(r->U);
      r->U = nullptr;
      continue;
    }
    r->dt_raw = dt0;
    r->dt_prev = 0.0;
    r->state = REQ_ACTIVE;
    s->in_use += r->bytes;
    s->active.push_back(r);
  }
  return MFM_OK;
}

}  // namespace

extern "C" {

int mfm_server_create(const mfm_params* cls, double hbm_budget_bytes, int max_active,
                      mfm_server** out) {
  if (!out || !cls || !(hbm_budget_bytes > 0.0) || max_active < 1) return MFM_ERR_ARG;
  *out = nullptr;
  mfm_ctx* c = nullptr;
  int rc = mfm_ctx_create(cls, &c);
  if (rc != MFM_OK) return rc;
  // a request's result may not depend on the requests batched with it:
  // the kernel structure is fixed instead of chosen per batch (AUTO)
  c->stage_mode = MFM_STAGE_SEPARATE;
  mfm_server* s = new mfm_server();
  s->ctx = c;
  s->budget = (size_t)hbm_budget_bytes;
  s->max_active = max_active;
  *out = s;
  return MFM_OK;
}

int mfm_server_destroy(mfm_server* s) {
  if (!s) return MFM_OK;
  (void)hipDeviceSynchronize();
  for (auto& kv : s->all) {
    if (kv.second->U) (void)hipFree(kv.second->U);
    delete kv.second;
  }
  mfm_ctx_destroy(s->ctx);
  delete s;
  return MFM_OK;
}

int mfm_server_set_stage_kernel(mfm_server* s, int mode) {
  if (!s) return MFM_ERR_ARG;
  if (mode != MFM_STAGE_SEPARATE && mode != MFM_STAGE_SUMMED && mode != MFM_STAGE_FUSED)
    return sfail(s, MFM_ERR_ARG, "the server takes a fixed stage structure (not AUTO)");
  if (!s->all.empty())
    return sfail(s, MFM_ERR_ARG, "set the stage structure before the first request");
  s->ctx->stage_mode = mode;
  return MFM_OK;
}

const char* mfm_server_last_error(const mfm_server* s) {
  if (!s) return "null server";
  return s->err.c_str();
}

int mfm_server_submit(mfm_server* s, const mfm_params* p, double dx, int nsteps, double cfl,
                      double initial_cfl, double max_dt_growth, long* id) {
  if (!s || !p || !id || nsteps < 1 || !(dx > 0.0) || !(cfl > 0.0) || !(initial_cfl > 0.0) ||
      !(max_dt_growth >= 1.0))
    return MFM_ERR_ARG;
  if (!same_class(*p, s->ctx->p))
    return sfail(s, MFM_ERR_ARG, "request parameters differ from the server's equation class");
  Request* r = new Request();
  std::string why;
  int rc = build_consts(p, r->C, why);
  if (rc != MFM_OK) {
    delete r;
    return sfail(s, rc, why);
  }
  r->p = *p;
  const int dims = r->C.dims;
  const int ng = mfm_num_ghost_cells(p);
  for (int d = 0; d < 3; ++d) {
    r->valid.lo[d] = d < dims ? p->domain_lo[d] : 0;
    r->valid.hi[d] = d < dims ? p->domain_hi[d] : 0;
  }
  if (r->valid.size() <= 0) {
    delete r;
    return sfail(s, MFM_ERR_SHAPE, "empty problem domain");
  }
  r->fab = grow(r->valid, ng, dims);
  if (r->fab.size() >= kMaxFabCells) {
    delete r;
    return sfail(s, MFM_ERR_SHAPE, "a FAB component must hold fewer than 2^30 cells");
  }
  r->dx = dx;
  r->cfl = cfl;
  r->cfl0 = initial_cfl;
  r->growth = max_dt_growth;
  r->nsteps = nsteps;
  r->bytes = request_bytes(s->ctx->p, s->ctx->stage_mode, r->C, r->fab, r->valid);
  r->id = s->next_id++;
  s->all[r->id] = r;
  s->pending.push_back(r);
  *id = r->id;
  return MFM_OK;
}

int mfm_server_step(mfm_server* s, int* n_active, int* n_pending, double* iter_ms) {
  if (!s) return MFM_ERR_ARG;
  int rc = admit(s);
  if (rc != MFM_OK) return rc;
  double ms = 0.0;
  if (!s->active.empty()) {
    const int n = (int)s->active.size();
    std::vector<Consts> Ct(n);
    std::vector<PatchIn> in(n);
    std::vector<double> nd(n);
    std::vector<int> cell(3 * n);
    for (int q = 0; q < n; ++q) {
      Request* r = s->active[q];
      Ct[q] = r->C;
      in[q].U = mfm_fab{r->U, {{r->fab.lo[0], r->fab.lo[1], r->fab.lo[2]}, {r->fab.hi[0], r->fab.hi[1], r->fab.hi[2]}}, r->C.nc};
      in[q].valid = mfm_box{{r->valid.lo[0], r->valid.lo[1], r->valid.lo[2]}, {r->valid.hi[0], r->valid.hi[1], r->valid.hi[2]}};
      in[q].F = nullptr;
      in[q].divB = nullptr;
      in[q].cidx = q;
      // AMRLevelIdealMHD::computeNewDt / computeInitialDt: cfl * dt_raw,
      // growth limited; the first step uses initial_cfl
      in[q].dt = r->dt_prev > 0.0 ? std::min(r->cfl * r->dt_raw, r->growth * r->dt_prev)
                                  : r->cfl0 * r->dt_raw;
      in[q].dx = r->dx;
      in[q].time = r->t;
      // periodic directions of a one-patch domain wrap onto themselves
      bool periodic = false;
      for (int d = 0; d < r->C.dims; ++d) periodic = periodic || !r->C.phys[d];
      if (periodic) {
        int ulo[3], un[3], vlo[3], vn[3];
        for (int d = 0; d < 3; ++d) {
          ulo[d] = r->fab.lo[d]; un[d] = (int)r->fab.n(d);
          vlo[d] = r->valid.lo[d]; vn[d] = (int)r->valid.n(d);
        }
        hipError_t e = launch_fill_periodic(r->U, ulo, un, r->fab.size(), r->C.nc, vlo, vn,
                                            r->C.dims, nullptr);
        if (e != hipSuccess) return sfail(s, MFM_ERR_DEVICE, "periodic fill launch");
      }
    }
    const double t0 = now_ms();
    rc = run_batch(s->ctx, Ct.data(), n, in.data(), n, nd.data(), cell.data(), nullptr, nullptr,
                   nullptr);
    ms = now_ms() - t0;
    if (rc != MFM_OK && rc != MFM_ERR_NUMERIC) return sfail(s, rc, s->ctx->err);
    std::vector<Request*> keep;
    for (int q = 0; q < n; ++q) {
      Request* r = s->active[q];
      r->t += in[q].dt;
      r->done += 1;
      s->cell_updates += (double)r->valid.size();
      const double d = nd[q];
      if (!(d > 0.0) || std::isnan(d)) {
        r->state = REQ_FAILED;
        r->err = "non-positive or NaN time step";
      } else {
        r->dt_raw = d;
        r->dt_prev = in[q].dt;
        if (r->done >= r->nsteps) r->state = REQ_DONE;
      }
      if (r->state == REQ_ACTIVE) keep.push_back(r);
    }
    s->active.swap(keep);
    s->iter_ms.push_back(ms);
    s->busy_ms += ms;
  }
  if (n_active) *n_active = (int)s->active.size();
  if (n_pending) *n_pending = (int)s->pending.size();
  if (iter_ms) *iter_ms = ms;
  return MFM_OK;
}

int mfm_server_request_info(mfm_server* s, long id, int* state, int* steps_done, double* t,
                            mfm_box* fab_box, int* ncomp) {
  if (!s) return MFM_ERR_ARG;
  auto it = s->all.find(id);
  if (it == s->all.end()) return sfail(s, MFM_ERR_ARG, "unknown request");
  const Request* r = it->second;
  if (state) *state = r->state;
  if (steps_done) *steps_done = r->done;
  if (t) *t = r->t;
  if (fab_box)
    for (int d = 0; d < 3; ++d) { fab_box->lo[d] = r->fab.lo[d]; fab_box->hi[d] = r->fab.hi[d]; }
  if (ncomp) *ncomp = r->C.nc;
  return MFM_OK;
}

const char* mfm_server_request_error(mfm_server* s, long id) {
  if (!s) return "null server";
  auto it = s->all.find(id);
  if (it == s->all.end()) return "unknown request";
  return it->second->err.c_str();
}

int mfm_server_fetch(mfm_server* s, long id, double* host, size_t bytes) {
  if (!s || !host) return MFM_ERR_ARG;
  auto it = s->all.find(id);
  if (it == s->all.end()) return sfail(s, MFM_ERR_ARG, "unknown request");
  Request* r = it->second;
  if (!r->U) return sfail(s, MFM_ERR_ARG, "request holds no state");
  const size_t need = (size_t)r->fab.size() * r->C.nc * sizeof(double);
  if (bytes < need) return sfail(s, MFM_ERR_SHAPE, "host buffer too small");
  hipError_t e = hipMemcpy(host, r->U, need, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return sfail(s, MFM_ERR_DEVICE, std::string("fetch: ") + hipGetErrorString(e));
  return MFM_OK;
}

int mfm_server_release(mfm_server* s, long id) {
  if (!s) return MFM_ERR_ARG;
  auto it = s->all.find(id);
  if (it == s->all.end()) return sfail(s, MFM_ERR_ARG, "unknown request");
  Request* r = it->second;
  if (r->state == REQ_ACTIVE || r->state == REQ_PENDING)
    return sfail(s, MFM_ERR_ARG, "request still running");
  if (r->U) {
    (void)hipDeviceSynchronize();
    (void)hipFree(r->U);
    r->U = nullptr;
    s->in_use -= r->bytes;
  }
  s->all.erase(it);
  delete r;
  return MFM_OK;
}

int mfm_server_stats(mfm_server* s, double* p50_ms, double* p99_ms, long* iterations,
                     double* cell_updates, double* busy_ms) {
  if (!s) return MFM_ERR_ARG;
  std::vector<double> v = s->iter_ms;
  std::sort(v.begin(), v.end());
  auto pct = [&](double q) {
    if (v.empty()) return 0.0;
    // nearest-rank percentile
    size_t k = (size_t)std::ceil(q * (double)v.size());
    if (k < 1) k = 1;
    return v[k - 1];
  };
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (iterations) *iterations = (long)s->iter_ms.size();
  if (cell_updates) *cell_updates = s->cell_updates;
  if (busy_ms) *busy_ms = s->busy_ms;
  return MFM_OK;
}

}  // extern "C"
