// Continuous-batching solve server (no reference counterpart).
//
// Per scheduling step:
//   1. admit queued requests (FIFO) while a slot is free and the HBM budget
//      holds their device block; upload points and right-hand sides, X = 0;
//   2. residual phase, for every request that needs one: Z = A X in FP64
//      (ragged product; skipped while X is still 0), R = B - Z, per-column
//      true relative residuals; the host reads them once for the whole batch,
//      retires converged requests and starts an inner solve for the others
//      (per-column tolerances, converged columns frozen);
//   3. inner phase: `iterations_per_step` batched CG iterations for every
//      request in its inner solve, all requests in the same ragged launches;
//      requests whose inner solve finished get X += E and go back to 2;
//   4. retire: copy solutions to the host, release slots and memory.
// Device memory comes from a caching pool, so admission after warm-up does
// not call hipMalloc.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../../include/mpla_kernel.h"
#include "../../../include/mpla_server.h"
#include "../../../include/mplak.h"
#include "runtime.hpp"

using mpla_rt::fail;
using mpla_rt::hip_ck;

namespace {

inline void ck(int status, const char* what) {
  if (status != MPLAK_OK) fail(what, mplak_status_string(status));
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// FP64 residual evaluations after which a request retires as
// MPLA_REQUEST_ITER_LIMIT whatever its iteration count (as the mixed solver's
// outer-step limit)
constexpr int kMaxRefinementSteps = 40;

// Caching device allocator: blocks are rounded to 1 MiB and kept on free
// lists after use.
class DevicePool {
 public:
  ~DevicePool() {
    for (auto& kv : free_) (void)hipFree(kv.second);
  }
  static size_t round(size_t b) { return (b + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1); }
  void* get(size_t bytes) {
    bytes = round(bytes);
    auto it = free_.find(bytes);
    if (it != free_.end()) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    hip_ck(hipMalloc(&p, bytes), "server pool hipMalloc");
    return p;
  }
  void put(void* p, size_t bytes) { free_.emplace(round(bytes), p); }

 private:
  std::multimap<size_t, void*> free_;
};

enum Phase { kQueued, kResidual, kInner, kFinished };

struct Layout {
  size_t points, vec, part, cgpart, perm, total;
};

Layout layout_for(int64_t n, int dim, int nrhs) {
  Layout L;
  L.points = align256(sizeof(double) * (size_t)dim * (size_t)n);
  L.vec = align256(sizeof(double) * (size_t)n * (size_t)nrhs);
  mplak_ragged_system s{};
  s.n = n;
  s.nrhs = nrhs;
  s.dim = dim;
  s.kind = MPLAK_KERNEL_GAUSSIAN;
  s.length_scale = 1.0;
  s.precision = MPLAK_PREC_F64;
  L.part = align256(mplak_ragged_plan(&s));
  L.cgpart = align256(mplak_ragged_cg_part_bytes(n, nrhs));
  L.perm = align256(sizeof(int64_t) * (size_t)n);
  L.total = L.points + 6 * L.vec + L.part + L.cgpart + L.perm;  // points, B, X, R, D, Z, E, spatial order
  return L;
}

struct Request {
  int id = 0;
  mpla_solve_request req{};
  mpla_solve_result res{};
  Phase phase = kQueued;
  bool x_zero = true;
  bool escalated = false;
  int slot = -1;
  void* block = nullptr;
  size_t block_bytes = 0;
  double *points = nullptr, *B = nullptr, *X = nullptr, *R = nullptr, *D = nullptr, *Z = nullptr, *E = nullptr;
  double *part = nullptr, *cgpart = nullptr;
  int64_t* perm = nullptr;  // spatial order (mixed mode): stored row k is request row perm[k]
  bool ordered = false;
  double residual_drop = 0.0;  // screen of the FP64 products (mixed mode; 0 = exact)
  double prev_max = INFINITY;
};

// Staging for one upload: system records and item tables, copied to the
// device in one transfer.
struct Upload {
  std::vector<mplak_ragged_system> psys;
  std::vector<mplak_ragged_cg> csys;
  std::vector<mplak_ragged_item> items;
  struct Launch {
    int kind;  // 0 product, 1 finish, 2 cg op
    int op;    // cg op; for a product, the systems' dimension
    int rt, precision, parity;
    double value;
    size_t first;
    int64_t count;
    int scheme = MPLAK_RAGGED_SLABS;  // product work split
  };
  std::vector<Launch> launches;
  void clear() {
    psys.clear();
    csys.clear();
    items.clear();
    launches.clear();
  }
};

}  // namespace

struct mpla_server {
  mpla_server_config cfg{};
  hipStream_t stream = nullptr;
  DevicePool pool;
  std::mutex mu;
  std::deque<int> queue;
  std::vector<Request*> requests;  // by id
  std::vector<int> active;         // ids
  std::vector<int> free_slots;
  size_t resident = 0, peak = 0;
  // state / group pools, one region per slot
  mplak_cg_state* states = nullptr;  // [slot][2][MAX_RHS]
  mplak_bcg_group* groups = nullptr; // [slot]
  mplak_cg_state* host_states = nullptr;
  mplak_bcg_group* host_groups = nullptr;
  // upload staging
  char* pinned = nullptr;
  size_t pinned_bytes = 0;
  char* dev_stage = nullptr;
  size_t dev_stage_bytes = 0;
  std::chrono::steady_clock::time_point t0;
  // drop level of the FP32 inner products (mixed mode; 0 = exact); FP64
  // products are screened per request when residual_drop > 0
  double inner_drop = 0.0, residual_drop = 0.0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  mplak_cg_state* slot_states(int slot) { return states + (size_t)slot * 2 * MPLA_SERVER_MAX_RHS; }
  mplak_cg_state* host_slot_states(int slot) { return host_states + (size_t)slot * 2 * MPLA_SERVER_MAX_RHS; }
};

namespace {

void ensure_stage(mpla_server* s, size_t bytes) {
  if (bytes <= s->pinned_bytes && bytes <= s->dev_stage_bytes) return;
  hip_ck(hipStreamSynchronize(s->stream), "server stage");
  const size_t want = std::max(bytes, 2 * s->pinned_bytes);
  if (s->pinned) (void)hipHostFree(s->pinned);
  if (s->dev_stage) (void)hipFree(s->dev_stage);
  hip_ck(hipHostMalloc(reinterpret_cast<void**>(&s->pinned), want, hipHostMallocDefault), "server stage");
  hip_ck(hipMalloc(reinterpret_cast<void**>(&s->dev_stage), want), "server stage");
  s->pinned_bytes = s->dev_stage_bytes = want;
}

// Copies the upload to the device and enqueues its launches.
void submit_upload(mpla_server* s, Upload& u) {
  if (u.launches.empty()) return;
  const size_t pb = align256(sizeof(mplak_ragged_system) * u.psys.size());
  const size_t cb = align256(sizeof(mplak_ragged_cg) * u.csys.size());
  const size_t ib = align256(sizeof(mplak_ragged_item) * u.items.size());
  ensure_stage(s, pb + cb + ib + 256);
  if (!u.psys.empty()) memcpy(s->pinned, u.psys.data(), sizeof(mplak_ragged_system) * u.psys.size());
  if (!u.csys.empty()) memcpy(s->pinned + pb, u.csys.data(), sizeof(mplak_ragged_cg) * u.csys.size());
  if (!u.items.empty()) memcpy(s->pinned + pb + cb, u.items.data(), sizeof(mplak_ragged_item) * u.items.size());
  hip_ck(hipMemcpyAsync(s->dev_stage, s->pinned, pb + cb + ib, hipMemcpyHostToDevice, s->stream), "server upload");
  auto* dps = reinterpret_cast<const mplak_ragged_system*>(s->dev_stage);
  auto* dcs = reinterpret_cast<const mplak_ragged_cg*>(s->dev_stage + pb);
  auto* dit = reinterpret_cast<const mplak_ragged_item*>(s->dev_stage + pb + cb);
  for (const auto& L : u.launches) {
    const mplak_ragged_item* it = dit + L.first;
    if (L.kind == 0)
      ck(mplak_ragged_product(dps, it, L.count, L.rt, L.op, L.precision, L.scheme, s->stream), "server product");
    else if (L.kind == 1)
      ck(mplak_ragged_finish(dps, it, L.count, s->stream), "server finish");
    else
      ck(mplak_ragged_cg_op(L.op, dcs, it, L.count, L.parity, L.value, s->stream), "server cg");
  }
  // the staging buffer is reused by the next upload: wait for this one
  hip_ck(hipStreamSynchronize(s->stream), "server step");
}

mplak_ragged_system product_system(const Request& r, int precision, const double* V, double* Y, const int32_t* skip,
                                   double drop_log2) {
  mplak_ragged_system sy{};
  sy.drop_log2 = r.ordered ? drop_log2 : 0.0;
  sy.points = r.points;
  sy.V = V;
  sy.Y = Y;
  sy.part = r.part;
  sy.skip = skip;
  sy.n = r.req.n;
  sy.length_scale = r.req.length_scale;
  sy.nugget = r.req.nugget;
  sy.dim = r.req.dim;
  sy.kind = r.req.kind;
  sy.nrhs = r.req.nrhs;
  sy.precision = precision;
  mplak_ragged_plan(&sy);
  return sy;
}

mplak_ragged_cg cg_system(mpla_server* s, const Request& r, double* X, double* Xout) {
  mplak_ragged_cg c{};
  c.X = X;
  c.R = r.R;
  c.D = r.D;
  c.Z = r.Z;
  c.B = r.B;
  c.Xout = Xout;
  c.states = s->slot_states(r.slot);
  c.group = s->groups + r.slot;
  c.part = r.cgpart;
  c.n = r.req.n;
  c.nrhs = r.req.nrhs;
  c.G = mplak_ragged_cg_blocks(r.req.n);
  c.eps = r.req.tolerance;
  return c;
}

// Adds product launches (grouped by rt) and the finish launch for `systems`.
void add_products(Upload& u, const std::vector<mplak_ragged_system>& systems, int precision) {
  if (systems.empty()) return;
  const size_t base = u.psys.size();
  for (const auto& sy : systems) u.psys.push_back(sy);
  // one launch per (scheme, rt, dim) class: kernels specialised for each;
  // the pair scheme (large systems, heavier items) first
  for (int scheme : {MPLAK_RAGGED_PAIRS, MPLAK_RAGGED_SLABS})
    for (int rt : {1, 2, 4, 8})
      for (int dim = 1; dim <= MPLA_SERVER_MAX_DIM; ++dim) {
        const size_t first = u.items.size();
        for (size_t k = 0; k < systems.size(); ++k) {
          if (systems[k].scheme != scheme || systems[k].rt != rt || systems[k].dim != dim) continue;
          const int64_t c = mplak_ragged_product_items(&systems[k], nullptr, 0);
          const size_t at = u.items.size();
          u.items.resize(at + (size_t)c);
          mplak_ragged_product_items(&systems[k], u.items.data() + at, (int32_t)(base + k));
        }
        const int64_t count = (int64_t)(u.items.size() - first);
        if (count > 0) u.launches.push_back({0, dim, rt, precision, 0, 0.0, first, count, scheme});
      }
  const size_t first = u.items.size();
  for (size_t k = 0; k < systems.size(); ++k) {
    const int64_t c = mplak_ragged_finish_items(&systems[k], nullptr, 0);
    const size_t at = u.items.size();
    u.items.resize(at + (size_t)c);
    mplak_ragged_finish_items(&systems[k], u.items.data() + at, (int32_t)(base + k));
  }
  u.launches.push_back({1, 0, 0, precision, 0, 0.0, first, (int64_t)(u.items.size() - first)});
}

// Adds one CG op over `systems` (appended to csys once per call site via base).
enum ItemShape { kBlocksColumns, kColumns, kSystems };

void add_cg_op(Upload& u, size_t base, size_t count_sys, int op, ItemShape shape, int parity, double value = 0.0) {
  const size_t first = u.items.size();
  for (size_t k = 0; k < count_sys; ++k) {
    const mplak_ragged_cg& c = u.csys[base + k];
    const int32_t si = (int32_t)(base + k);
    if (shape == kBlocksColumns) {
      const int64_t n = mplak_ragged_cg_items(&c, nullptr, 0);
      const size_t at = u.items.size();
      u.items.resize(at + (size_t)n);
      mplak_ragged_cg_items(&c, u.items.data() + at, si);
    } else if (shape == kColumns) {
      for (int32_t col = 0; col < c.nrhs; ++col) u.items.push_back({si, 0, col, 0});
    } else {
      u.items.push_back({si, 0, 0, 0});
    }
  }
  u.launches.push_back({2, op, 0, 0, parity, value, first, (int64_t)(u.items.size() - first)});
}

void release(mpla_server* s, Request* r) {
  if (r->block) s->pool.put(r->block, r->block_bytes);
  s->resident -= r->block_bytes;
  r->block = nullptr;
  if (r->slot >= 0) s->free_slots.push_back(r->slot);
  r->slot = -1;
}

bool admit(mpla_server* s, Request* r) {
  const Layout L = layout_for(r->req.n, r->req.dim, r->req.nrhs);
  const size_t need = DevicePool::round(L.total);
  if (s->free_slots.empty() || s->resident + need > s->cfg.hbm_budget_bytes) return false;
  r->slot = s->free_slots.back();
  s->free_slots.pop_back();
  r->block = s->pool.get(need);
  r->block_bytes = need;
  s->resident += need;
  s->peak = std::max(s->peak, s->resident);
  char* p = static_cast<char*>(r->block);
  r->points = reinterpret_cast<double*>(p);
  p += L.points;
  double** vecs[6] = {&r->B, &r->X, &r->R, &r->D, &r->Z, &r->E};
  for (double** v : vecs) {
    *v = reinterpret_cast<double*>(p);
    p += L.vec;
  }
  r->part = reinterpret_cast<double*>(p);
  p += L.part;
  r->cgpart = reinterpret_cast<double*>(p);
  p += L.cgpart;
  r->perm = reinterpret_cast<int64_t*>(p);
  const int64_t n = r->req.n;
  const int dim = r->req.dim;
  r->ordered = s->cfg.mode == MPLA_SERVE_MIXED_BATCHED && (s->inner_drop > 0.0 || s->residual_drop > 0.0);
  r->residual_drop =
      s->residual_drop > 0.0 ? mpla_rt::residual_screen_log2(n, r->req.nugget, r->req.tolerance) : 0.0;
  if (!r->ordered) {
    hip_ck(hipMemcpyAsync(r->points, r->req.points, sizeof(double) * (size_t)dim * (size_t)n, hipMemcpyHostToDevice,
                          s->stream),
           "server admit");
    hip_ck(hipMemcpyAsync(r->B, r->req.rhs, sizeof(double) * (size_t)n * (size_t)r->req.nrhs,
                          hipMemcpyHostToDevice, s->stream),
           "server admit");
  } else {
    // spatial (Hilbert) order of the request's points, so that the screened
    // products can skip far blocks; the right-hand sides follow the points
    const size_t ob = mplak_spatial_order_workspace_bytes(n, dim);
    if (ob == 0) fail("server admit", "spatial order workspace query failed");
    const size_t raw = align256(sizeof(double) * (size_t)dim * (size_t)n);
    const size_t tmp_bytes = raw + 256 + align256(ob);
    char* tmp = static_cast<char*>(s->pool.get(tmp_bytes));
    double* raw_points = reinterpret_cast<double*>(tmp);
    double** table = reinterpret_cast<double**>(tmp + raw);
    double* host_table[MPLA_SERVER_MAX_DIM];
    for (int d = 0; d < dim; ++d) host_table[d] = raw_points + (size_t)d * n;
    hip_ck(hipMemcpyAsync(raw_points, r->req.points, sizeof(double) * (size_t)dim * (size_t)n, hipMemcpyHostToDevice,
                          s->stream),
           "server admit");
    hip_ck(hipMemcpyAsync(table, host_table, sizeof(double*) * dim, hipMemcpyHostToDevice, s->stream),
           "server admit");
    ck(mplak_spatial_order(table, dim, 0, n, r->perm, tmp + raw + 256, ob, s->stream), "server admit order");
    ck(mplak_gather_points(table, dim, 0, r->perm, n, r->points, n, s->stream), "server admit order");
    hip_ck(hipMemcpyAsync(r->E, r->req.rhs, sizeof(double) * (size_t)n * (size_t)r->req.nrhs, hipMemcpyHostToDevice,
                          s->stream),
           "server admit");
    for (int c = 0; c < r->req.nrhs; ++c)
      ck(mplak_gather_f64(r->B + (size_t)c * n, r->E + (size_t)c * n, r->perm, n, s->stream), "server admit order");
    // host_table and the temporary block must outlive the copies and kernels
    hip_ck(hipStreamSynchronize(s->stream), "server admit");
    s->pool.put(tmp, tmp_bytes);
  }
  ck(mplak_fill_f64(r->X, n * r->req.nrhs, 0.0, s->stream), "server admit");
  r->phase = kResidual;
  r->x_zero = true;
  r->res.t_admit = s->now();
  return true;
}

double env_or(const char* name, double dflt) {
  const char* v = getenv(name);
  return v ? atof(v) : dflt;
}

}  // namespace

extern "C" {

void mpla_server_default_config(mpla_server_config* cfg) {
  cfg->hbm_budget_bytes = (size_t)200 * 1000 * 1000 * 1000;
  cfg->max_active = 64;
  cfg->mode = MPLA_SERVE_MIXED_BATCHED;
  cfg->iterations_per_step = 4;
  cfg->inner_tol_floor = 1e-4;
  cfg->max_iterations = 20000;
}

size_t mpla_server_request_bytes(int64_t n, int dim, int nrhs) {
  if (n <= 0 || dim < 1 || nrhs < 1) return 0;
  return DevicePool::round(layout_for(n, dim, nrhs).total);
}

mpla_server* mpla_server_create(const mpla_server_config* cfg) {
  mpla_server* s = new mpla_server();
  if (cfg)
    s->cfg = *cfg;
  else
    mpla_server_default_config(&s->cfg);
  if (s->cfg.mode == MPLA_SERVE_FP64_SERIAL) s->cfg.max_active = 1;
  s->cfg.max_active = std::max(1, s->cfg.max_active);
  s->cfg.iterations_per_step = std::max(2, (s->cfg.iterations_per_step + 1) & ~1);
  hip_ck(hipStreamCreate(&s->stream), "mpla_server_create");
  const size_t sb = sizeof(mplak_cg_state) * (size_t)s->cfg.max_active * 2 * MPLA_SERVER_MAX_RHS;
  const size_t gb = sizeof(mplak_bcg_group) * (size_t)s->cfg.max_active;
  hip_ck(hipMalloc(reinterpret_cast<void**>(&s->states), sb), "mpla_server_create");
  hip_ck(hipMalloc(reinterpret_cast<void**>(&s->groups), gb), "mpla_server_create");
  hip_ck(hipHostMalloc(reinterpret_cast<void**>(&s->host_states), sb, hipHostMallocDefault), "mpla_server_create");
  hip_ck(hipHostMalloc(reinterpret_cast<void**>(&s->host_groups), gb, hipHostMallocDefault), "mpla_server_create");
  for (int k = s->cfg.max_active - 1; k >= 0; --k) s->free_slots.push_back(k);
  ensure_stage(s, size_t(4) << 20);
  if (s->cfg.mode == MPLA_SERVE_MIXED_BATCHED) {
    // FP32 inner products drop entries 6 bits below the inner tolerance floor
    // (MPLA_IR_DROP_LOG2 overrides; <= 0 keeps every entry): a wider margin
    // than the single-system solver's 3 bits, since the served requests get
    // no screen-tightening safeguard.  FP64 products drop, per request, the
    // entries whose sum moves a residual by at most 2^-12 of the request's
    // tolerance (mpla_rt::residual_screen_log2; the FP64-exact level at most)
    const double floor = s->cfg.inner_tol_floor > 0.0 ? s->cfg.inner_tol_floor : 1e-4;
    s->inner_drop = env_or("MPLA_IR_DROP_LOG2", std::ceil(-std::log2(floor)) + 6.0);
    if (s->inner_drop > 0.0) s->residual_drop = 80.0;
    if (s->inner_drop < 0.0) s->inner_drop = 0.0;
  }
  s->t0 = std::chrono::steady_clock::now();
  return s;
}

void mpla_server_destroy(mpla_server* s) {
  if (!s) return;
  (void)hipStreamSynchronize(s->stream);
  for (Request* r : s->requests) {
    if (r && r->block) s->pool.put(r->block, r->block_bytes);
    delete r;
  }
  (void)hipFree(s->states);
  (void)hipFree(s->groups);
  (void)hipHostFree(s->host_states);
  (void)hipHostFree(s->host_groups);
  if (s->pinned) (void)hipHostFree(s->pinned);
  if (s->dev_stage) (void)hipFree(s->dev_stage);
  (void)hipStreamDestroy(s->stream);
  delete s;
}

double mpla_server_clock(mpla_server* s) { return s->now(); }

size_t mpla_server_peak_bytes(mpla_server* s) { return s->peak; }

int mpla_server_submit(mpla_server* s, const mpla_solve_request* q) {
  std::lock_guard<std::mutex> lock(s->mu);
  Request* r = new Request();
  r->id = (int)s->requests.size();
  r->req = *q;
  r->res.t_arrival = s->now();
  s->requests.push_back(r);
  const bool bad = q->n <= 0 || q->dim < 1 || q->dim > MPLA_SERVER_MAX_DIM || q->nrhs < 1 ||
                   q->nrhs > MPLA_SERVER_MAX_RHS || !(q->length_scale > 0.0) || !q->points || !q->rhs ||
                   (q->kind != MPLA_KERNEL_GAUSSIAN && q->kind != MPLA_KERNEL_MATERN32) ||
                   mpla_server_request_bytes(q->n, q->dim, q->nrhs) > s->cfg.hbm_budget_bytes;
  if (bad) {
    r->phase = kFinished;
    r->res.status = MPLA_REQUEST_REJECTED;
    r->res.t_done = r->res.t_arrival;
    return MPLA_REQUEST_REJECTED;
  }
  r->res.status = MPLA_REQUEST_PENDING;
  s->queue.push_back(r->id);
  return r->id;
}

int mpla_server_result(mpla_server* s, int id, mpla_solve_result* out) {
  std::lock_guard<std::mutex> lock(s->mu);
  if (id < 0 || id >= (int)s->requests.size()) return MPLA_REQUEST_UNKNOWN;
  if (out) *out = s->requests[id]->res;
  return s->requests[id]->res.status;
}

int mpla_server_step(mpla_server* s) {
  std::lock_guard<std::mutex> lock(s->mu);
  const bool serial = s->cfg.mode == MPLA_SERVE_FP64_SERIAL;
  const int inner_default = serial ? MPLAK_PREC_F64 : MPLAK_PREC_F32;

  // 1. admission, first come first served
  while (!s->queue.empty()) {
    Request* r = s->requests[s->queue.front()];
    if (!admit(s, r)) break;
    s->queue.pop_front();
    s->active.push_back(r->id);
  }

  // 2. residual phase
  std::vector<Request*> resid, inner;
  for (int id : s->active) {
    Request* r = s->requests[id];
    if (r->phase == kResidual) resid.push_back(r);
  }
  std::vector<Request*> finished;
  if (!resid.empty()) {
    Upload u;
    std::vector<mplak_ragged_system> prod;
    for (Request* r : resid)
      if (!r->x_zero) prod.push_back(product_system(*r, MPLAK_PREC_F64, r->X, r->Z, nullptr, r->residual_drop));
    add_products(u, prod, MPLAK_PREC_F64);
    // Z = 0 for requests whose X is still zero (the fill op writes the X field)
    const size_t zb = u.csys.size();
    size_t nz = 0;
    for (Request* r : resid)
      if (r->x_zero) {
        u.csys.push_back(cg_system(s, *r, r->Z, nullptr));
        ++nz;
      }
    if (nz) add_cg_op(u, zb, nz, MPLAK_RAGGED_FILL, kBlocksColumns, 0, 0.0);
    const size_t base = u.csys.size();
    for (Request* r : resid) u.csys.push_back(cg_system(s, *r, r->E, r->X));
    add_cg_op(u, base, resid.size(), MPLAK_RAGGED_START, kBlocksColumns, 0);
    add_cg_op(u, base, resid.size(), MPLAK_RAGGED_START_FINAL, kColumns, 0);
    submit_upload(s, u);
    const size_t sb = sizeof(mplak_cg_state) * (size_t)s->cfg.max_active * 2 * MPLA_SERVER_MAX_RHS;
    hip_ck(hipMemcpyAsync(s->host_states, s->states, sb, hipMemcpyDeviceToHost, s->stream), "server states");
    hip_ck(hipStreamSynchronize(s->stream), "server states");

    Upload v;
    std::vector<Request*> starting;
    for (Request* r : resid) {
      mplak_cg_state* st = s->host_slot_states(r->slot);
      double max_rel = 0.0;
      bool all = true, finite = true;
      for (int c = 0; c < r->req.nrhs; ++c) {
        max_rel = std::max(max_rel, st[c].rel);
        if (!(st[c].rel < r->req.tolerance)) all = false;
        if (!std::isfinite(st[c].rel)) finite = false;
      }
      if (!finite) max_rel = NAN;
      r->res.refinement_steps += 1;
      r->res.max_rel_residual = max_rel;
      // a non-finite residual cannot improve, and refinement steps are capped
      // so that a request whose inner solves gain nothing still retires
      if (all || !finite || r->res.iterations >= s->cfg.max_iterations ||
          r->res.refinement_steps > kMaxRefinementSteps) {
        r->res.status = all ? MPLA_REQUEST_CONVERGED : MPLA_REQUEST_ITER_LIMIT;
        finished.push_back(r);
        continue;
      }
      if (!serial && r->res.refinement_steps > 1 && max_rel > 0.5 * r->prev_max) r->escalated = true;
      r->prev_max = max_rel;
      const bool f64 = serial || r->escalated;
      const double floor_eta = f64 ? 0.0 : s->cfg.inner_tol_floor;
      for (int c = 0; c < r->req.nrhs; ++c) {
        mplak_cg_state& h = st[c];
        if (h.rel < r->req.tolerance) {
          h.done = MPLAK_CONVERGED;
          continue;
        }
        h.bnorm2 = h.rr;
        h.eps = std::min(0.5, std::max(floor_eta, 0.5 * r->req.tolerance / h.rel));
        h.rel = 1.0;
        h.k = 0;
        h.iter_max = std::max(2, s->cfg.max_iterations - r->res.iterations + 1);
        h.rule = MPLAK_STOP_RELATIVE;
        h.done = MPLAK_RUNNING;
      }
      hip_ck(hipMemcpyAsync(s->slot_states(r->slot), st, sizeof(mplak_cg_state) * (size_t)r->req.nrhs,
                            hipMemcpyHostToDevice, s->stream),
             "server states");
      r->phase = kInner;
      starting.push_back(r);
    }
    if (!starting.empty()) {
      const size_t b = v.csys.size();
      for (Request* r : starting) v.csys.push_back(cg_system(s, *r, r->E, r->X));
      add_cg_op(v, b, starting.size(), MPLAK_RAGGED_FILL, kBlocksColumns, 0, 0.0);
      add_cg_op(v, b, starting.size(), MPLAK_RAGGED_GROUP, kSystems, 0);
      submit_upload(s, v);
    }
  }

  // 3. inner phase: L batched iterations for every request in an inner solve
  for (int id : s->active) {
    Request* r = s->requests[id];
    if (r->phase == kInner) inner.push_back(r);
  }
  if (!inner.empty()) {
    Upload u;
    const int L = s->cfg.iterations_per_step;
    std::vector<mplak_ragged_system> p32, p64;
    for (Request* r : inner) {
      const bool f64 = serial || r->escalated;
      (f64 ? p64 : p32).push_back(product_system(*r, f64 ? MPLAK_PREC_F64 : MPLAK_PREC_F32, r->D, r->Z,
                                                 &s->groups[r->slot].all_done,
                                                 f64 ? r->residual_drop : s->inner_drop));
    }
    const size_t base = u.csys.size();
    for (Request* r : inner) u.csys.push_back(cg_system(s, *r, r->E, r->X));
    // one iteration's launches; the other L-1 reuse its item tables with the
    // state buffers of their own parity
    const size_t l0 = u.launches.size();
    add_products(u, p32, MPLAK_PREC_F32);
    add_products(u, p64, MPLAK_PREC_F64);
    add_cg_op(u, base, inner.size(), MPLAK_RAGGED_DZ, kBlocksColumns, 0);
    add_cg_op(u, base, inner.size(), MPLAK_RAGGED_UPDATE, kBlocksColumns, 0);
    add_cg_op(u, base, inner.size(), MPLAK_RAGGED_DIRECT, kBlocksColumns, 0);
    add_cg_op(u, base, inner.size(), MPLAK_RAGGED_GROUP, kSystems, 1);
    const size_t l1 = u.launches.size();
    for (int j = 1; j < L; ++j)
      for (size_t k = l0; k < l1; ++k) {
        Upload::Launch c = u.launches[k];
        if (c.kind == 2) c.parity = c.op == MPLAK_RAGGED_GROUP ? (j + 1) & 1 : j & 1;
        u.launches.push_back(c);
      }
    (void)inner_default;
    submit_upload(s, u);
    const size_t gbytes = sizeof(mplak_bcg_group) * (size_t)s->cfg.max_active;
    const size_t sbytes = sizeof(mplak_cg_state) * (size_t)s->cfg.max_active * 2 * MPLA_SERVER_MAX_RHS;
    hip_ck(hipMemcpyAsync(s->host_groups, s->groups, gbytes, hipMemcpyDeviceToHost, s->stream), "server groups");
    hip_ck(hipMemcpyAsync(s->host_states, s->states, sbytes, hipMemcpyDeviceToHost, s->stream), "server groups");
    hip_ck(hipStreamSynchronize(s->stream), "server groups");
    Upload w;
    std::vector<Request*> done_inner;
    for (Request* r : inner)
      if (s->host_groups[r->slot].all_done) done_inner.push_back(r);
    if (!done_inner.empty()) {
      const size_t b = w.csys.size();
      for (Request* r : done_inner) {
        const mplak_cg_state* st = s->host_slot_states(r->slot);  // buffer 0 after an even number of iterations
        int kmax = 0;
        for (int c = 0; c < r->req.nrhs; ++c) kmax = std::max(kmax, st[c].k);
        r->res.iterations += kmax;
        r->phase = kResidual;
        r->x_zero = false;
        w.csys.push_back(cg_system(s, *r, r->E, r->X));
      }
      add_cg_op(w, b, done_inner.size(), MPLAK_RAGGED_AXPY, kBlocksColumns, 0);
      submit_upload(s, w);
    }
  }

  // 4. retire
  if (!finished.empty()) {
    for (Request* r : finished) {
      if (!r->req.solution) continue;
      const double* out = r->X;
      if (r->ordered) {  // back to the request's row order
        const int64_t n = r->req.n;
        for (int c = 0; c < r->req.nrhs; ++c)
          ck(mplak_scatter_f64(r->Z + (size_t)c * n, r->X + (size_t)c * n, r->perm, n, s->stream), "server retire");
        out = r->Z;
      }
      hip_ck(hipMemcpyAsync(r->req.solution, out, sizeof(double) * (size_t)r->req.n * (size_t)r->req.nrhs,
                            hipMemcpyDeviceToHost, s->stream),
             "server retire");
    }
    hip_ck(hipStreamSynchronize(s->stream), "server retire");
    const double t = s->now();
    for (Request* r : finished) {
      r->res.t_done = t;
      r->phase = kFinished;
      release(s, r);
      s->active.erase(std::find(s->active.begin(), s->active.end(), r->id));
    }
  }
  return (int)(s->queue.size() + s->active.size());
}

int mpla_server_run_trace(mpla_server* s, const mpla_solve_request* requests, int count,
                          mpla_solve_result* results) {
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return requests[a].arrival < requests[b].arrival; });
  std::vector<int> ids(count, -1);
  const double start = s->now();
  int next = 0;
  for (;;) {
    const double t = s->now() - start;
    while (next < count && requests[order[next]].arrival <= t) {
      ids[order[next]] = mpla_server_submit(s, &requests[order[next]]);
      // latency is measured from the trace arrival time
      if (ids[order[next]] >= 0) s->requests[ids[order[next]]]->res.t_arrival = start + requests[order[next]].arrival;
      ++next;
    }
    const int busy = mpla_server_step(s);
    if (busy == 0) {
      if (next >= count) break;
      const double wait = requests[order[next]].arrival - (s->now() - start);
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
    }
  }
  int failed = 0;
  for (int i = 0; i < count; ++i) {
    mpla_solve_result res{};
    if (ids[i] >= 0)
      mpla_server_result(s, ids[i], &res);
    else
      res.status = MPLA_REQUEST_REJECTED;
    if (results) results[i] = res;
    if (res.status != MPLA_REQUEST_CONVERGED) ++failed;
  }
  return failed;
}

}  // extern "C"
