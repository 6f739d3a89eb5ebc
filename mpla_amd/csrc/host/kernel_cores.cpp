// Matrix-free kernel-matrix plugin cores (src/kernel.cu of the reference) and
// the mixed-precision kernel CG.
//
// The reference cores materialise row panels of the distance matrix, apply
// exp in place and call cublasDgemv (src/kernel.cu:87-171), paying 32 bytes of
// HBM traffic per entry.  Here every core is one fused matrix-free product
// (mplak_kernel_matvec) whose entries live only in registers.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../../include/kernel.h"
#include "../../../include/mpla_kernel.h"
#include "chunk.hpp"
#include "runtime.hpp"

using namespace mpla_rt;

namespace {

mplak_kernel_desc make_desc(double** points, int dim, int kind, double length_scale, double nugget,
                            double shift, const mpla_generic_matrix* A) {
  mplak_kernel_desc d;
  d.points = points;
  d.dim = dim;
  d.kind = kind;
  d.length_scale = length_scale;
  d.nugget = nugget;
  d.coincident_shift = shift;
  d.row_offset = A->cur_proc_row_offset;
  d.row_count = A->cur_proc_row_count;
  d.col_offset = A->cur_proc_col_offset;
  d.col_count = A->cur_proc_col_count;
  return d;
}

// The reference plugin's operator: exp(-r*r) (length scale 1/sqrt(2)) plus
// 1e-8 where r < 1e-14 (src/kernel.cu:12-15, 39-50).
mplak_kernel_desc reference_desc(const mpla_generic_matrix* A) {
  const kernel_matrix_data* kd = static_cast<const kernel_matrix_data*>(A->data);
  return make_desc(kd->points, kd->dim, MPLAK_KERNEL_GAUSSIAN, 1.0 / std::sqrt(2.0), 0.0, 1.0e-8, A);
}

mplak_kernel_desc fused_desc(const mpla_generic_matrix* A) {
  const mpla_kernel_matrix_data* kd = static_cast<const mpla_kernel_matrix_data*>(A->data);
  return make_desc(kd->points, kd->dim, kd->kernel, kd->length_scale, kd->nugget, kd->coincident_shift, A);
}

void apply_kernel(const mplak_kernel_desc& d, int precision, const double* x, double* y,
                  const int32_t* skip, mpla_instance* inst, const char* what) {
  const size_t wb = mplak_kernel_matvec_workspace_bytes(&d);
  void* w = arena(inst, kCore, wb);
  ck(mplak_kernel_matvec(&d, precision, x, y, w, wb, skip, stream_of(inst)), what);
}

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Loosest relaxed screen of the mixed inner products (entries below 2^-6 =
// 1/64 are never dropped) and tightest screen the safeguard goes to.
constexpr double kMinRelaxDrop = 6.0;
constexpr double kMaxDrop = 60.0;

// Plain FP64 CG (relative rule) through the fused core, printing nothing.
void plain_fp64_cg(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, int iter_max, double eps,
                   mpla_instance* inst) {
  mpla_kernel_matrix_data f64 = *static_cast<const mpla_kernel_matrix_data*>(A->data);
  f64.precision = MPLA_PREC_F64;
  mpla_generic_matrix A64 = *A;
  A64.data = &f64;
  const int v = inst->verbosity;
  inst->verbosity = v >= 2 ? v : 0;
  mpla_generic_conjugate_gradient_with_checkpoint_restart(b, &A64, x, iter_max, eps, mpla_dgemv_core_kernel_fused,
                                                          0, false, inst);
  inst->verbosity = v;
}

// Operator of the ordered plain FP64 solve: the planned FP64 product on
// points in Hilbert order with the FP64-exact screen.
struct PlannedPayload {
  mplak_kernel_desc desc;
  void* plan;
  size_t plan_bytes;
  double drop;
};

void planned_fp64_core(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, mpla_instance* inst) {
  const PlannedPayload* p = static_cast<const PlannedPayload*>(A->data);
  ck(mplak_kernel_matvec_planned(&p->desc, MPLAK_PREC_F64, p->drop, p->plan, p->plan_bytes, x->data, b->data,
                                 inst->skip_flag, stream_of(inst)),
     "ordered FP64 CG product");
}

// Plain FP64 CG (the reference recurrences and relative rule) on the points
// in Hilbert order, with the FP64 product screened at 2^-(53 + log2 n): the at
// most n entries it leaves out of a row change that row by at most an ulp of
// its largest term.  Square blocks of up to 3 dimensions without a
// coincidence shift; anything else runs plain_fp64_cg.
void ordered_fp64_cg(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, int iter_max, double eps,
                     mpla_instance* inst) {
  const mplak_kernel_desc desc0 = fused_desc(A);
  const int64_t n = x->cur_proc_row_count;
  mplak_kernel_desc shape = desc0;
  shape.row_offset = shape.col_offset = 0;
  const size_t pb = mplak_screen_plan_bytes(&shape);
  const bool ok = desc0.row_offset == desc0.col_offset && desc0.row_count == desc0.col_count && pb > 0 && n > 0 &&
                  desc0.coincident_shift == 0.0;
  if (!ok) {
    plain_fp64_cg(b, A, x, iter_max, eps, inst);
    return;
  }
  const int dim = desc0.dim;
  const size_t vb = align256(sizeof(double) * (size_t)n);
  const size_t ob = mplak_spatial_order_workspace_bytes(n, dim);
  if (ob == 0) fail("mpla_kernel_conjugate_gradient_mixed", "spatial order workspace query failed");
  char* base = static_cast<char*>(arena(inst, kMixed, 3 * vb + (size_t)dim * vb + 256 + align256(ob) + align256(pb)));
  int64_t* perm = reinterpret_cast<int64_t*>(base);
  double* bp = reinterpret_cast<double*>(base + vb);
  double* xp = reinterpret_cast<double*>(base + 2 * vb);
  double* pts = reinterpret_cast<double*>(base + 3 * vb);
  double** table = reinterpret_cast<double**>(base + 3 * vb + (size_t)dim * vb);
  void* owork = base + 3 * vb + (size_t)dim * vb + 256;
  void* plan = base + 3 * vb + (size_t)dim * vb + 256 + align256(ob);
  mplak_stream_t s = stream_of(inst);
  Runtime* rt = runtime(inst);
  ck(mplak_spatial_order(desc0.points, dim, desc0.row_offset, n, perm, owork, ob, s), "ordered FP64 CG order");
  const int64_t ld = (int64_t)(vb / sizeof(double));
  ck(mplak_gather_points(desc0.points, dim, desc0.row_offset, perm, n, pts, ld, s), "ordered FP64 CG points");
  double** host_table = static_cast<double**>(rt->pinned);
  for (int k = 0; k < dim; ++k) host_table[k] = pts + (size_t)k * ld;
  hip_ck(hipMemcpyAsync(table, host_table, sizeof(double*) * dim, hipMemcpyHostToDevice, rt->stream),
         "hipMemcpyAsync");
  ck(mplak_gather_f64(bp, b->data, perm, n, s), "ordered FP64 CG permute");
  ck(mplak_gather_f64(xp, x->data, perm, n, s), "ordered FP64 CG permute");
  PlannedPayload pay;
  pay.desc = desc0;
  pay.desc.points = table;
  pay.desc.row_offset = pay.desc.col_offset = 0;
  pay.plan = plan;
  pay.plan_bytes = pb;
  pay.drop = residual_drop_log2(n);
  ck(mplak_screen_plan_build(&pay.desc, MPLAK_PREC_F64, pay.drop, plan, pb, s), "ordered FP64 CG plan");
  mpla_generic_matrix Ap = *A;
  Ap.data = &pay;
  mpla_vector bv = vector_view(b, bp);
  mpla_vector xv = vector_view(x, xp);
  const int v = inst->verbosity;
  inst->verbosity = v >= 2 ? v : 0;
  mpla_generic_conjugate_gradient_with_checkpoint_restart(&bv, &Ap, &xv, iter_max, eps, planned_fp64_core, 0, false,
                                                          inst);
  inst->verbosity = v;
  ck(mplak_scatter_f64(x->data, xp, perm, n, s), "ordered FP64 CG permute back");
  hip_ck(hipStreamSynchronize(rt->stream), "ordered FP64 CG");
}

double env_double(const char* name, double dflt) {
  const char* v = getenv(name);
  return v ? atof(v) : dflt;
}

}  // namespace

extern "C" {

void mpla_set_gaussian_kernel_rhs(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_instance* instance) {
  const kernel_matrix_data* kd = static_cast<const kernel_matrix_data*>(A->data);
  ck(mplak_gaussian_kernel_rhs(kd->points, kd->dim, A->cur_proc_row_offset, A->cur_proc_row_count, b->data,
                               stream_of(instance)),
     "mpla_set_gaussian_kernel_rhs");
}

void mpla_dgemv_core_kernel_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist,
                                   struct mpla_instance* instance) {
  apply_kernel(reference_desc(A), MPLAK_PREC_F64, x_redist->data, b->data, instance->skip_flag, instance,
               "mpla_dgemv_core_kernel_cublas");
}

void mpla_dgemv_core_kernel_streamed_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A,
                                            struct mpla_vector* x_redist, struct mpla_instance* instance) {
  apply_kernel(reference_desc(A), MPLAK_PREC_F64, x_redist->data, b->data, instance->skip_flag, instance,
               "mpla_dgemv_core_kernel_streamed_cublas");
}

void mpla_dgemv_core_kernel_fused(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist,
                                  struct mpla_instance* instance) {
  const mpla_kernel_matrix_data* kd = static_cast<const mpla_kernel_matrix_data*>(A->data);
  const int prec = kd->precision == MPLA_PREC_F32 ? MPLAK_PREC_F32 : MPLAK_PREC_F64;
  apply_kernel(fused_desc(A), prec, x_redist->data, b->data, instance->skip_flag, instance,
               "mpla_dgemv_core_kernel_fused");
}

void mpla_kernel_conjugate_gradient_mixed(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x,
                                          int iter_max, double epsilon, int inner_precision,
                                          struct mpla_instance* inst) {
  const mpla_kernel_matrix_data* kd = static_cast<const mpla_kernel_matrix_data*>(A->data);
  if (inner_precision == MPLA_PREC_F64) {
    ordered_fp64_cg(b, A, x, iter_max, epsilon, inst);
    return;
  }
  if (inner_precision != MPLA_PREC_F32)
    fail("mpla_kernel_conjugate_gradient_mixed", "inner precision must be MPLA_PREC_F32 or MPLA_PREC_F64");
  (void)kd;

  const mplak_kernel_desc desc0 = fused_desc(A);
  const int64_t n = x->cur_proc_row_count;
  const double n_global = (double)x->vec_row_count;
  // Inner solves stop at a relative residual eta >= MPLA_IR_INNER_TOL
  // (default 1e-4).  The solve runs on the points sorted along a Hilbert curve
  // (b, x permuted in, x permuted back out), so that the FP32 inner products
  // can skip blocks of entries below 2^-drop: MPLA_IR_DROP_LOG2, by default
  // one bit below the inner tolerance (15 for 1e-4); <= 0 keeps every entry.
  // A refinement step that gains much less than eta (the screen costs
  // contraction) tightens the screen by 6 bits for the next inner solve.
  // Within an inner solve the products are relaxed as its residual falls
  // (inexact Krylov: Simoncini & Szyld, SIAM J. Sci. Comput. 25, 2003; van
  // den Eshof & Sleijpen, SIAM J. Matrix Anal. Appl. 26, 2004, where the
  // product error may grow like ||r_0|| / ||r_k||): the screen drops 5 bits
  // more once ||r_k|| <= 2^-5 ||r_0|| and 9 bits more once ||r_k|| <=
  // 2^-9 ||r_0|| (MPLA_IR_RELAX=0 keeps one level).  The FP64 residual
  // products keep every entry above 2^-residual_screen_log2 (what they leave
  // out moves the residual by at most 2^-12 epsilon ||b||; screened per
  // 16 x 16 tile with a plan of their own), so the stopping test and the
  // final residual are the true ones whatever the inner operator dropped.
  const double eta_floor = env_double("MPLA_IR_INNER_TOL", 1.0e-4);
  double drop = env_double("MPLA_IR_DROP_LOG2", std::ceil(-std::log2(eta_floor)) + 1.0);
  const bool relax = env_double("MPLA_IR_RELAX", 1.0) != 0.0;
  // a non-negative coincidence shift adds a positive semi-definite block
  const double res_drop =
      residual_screen_log2(n, desc0.coincident_shift >= 0.0 ? desc0.nugget : 0.0, epsilon);
  mplak_kernel_desc shape = desc0;
  shape.row_offset = shape.col_offset = 0;
  const size_t pb = mplak_screen_plan_bytes(&shape);
  // the planned relaxed inner product has no coincidence-shift form: shifted
  // operators take the unordered path (exact FP32 inner products)
  const bool ordered = drop > 0.0 && desc0.row_offset == desc0.col_offset && desc0.row_count == desc0.col_count &&
                       pb > 0 && n > 0 && desc0.coincident_shift == 0.0;
  const int dim = desc0.dim;
  const size_t vb = align256(sizeof(double) * (size_t)n);
  const size_t sb = align256(2 * sizeof(mplak_cg_state));
  const size_t wb = mplak_cg_workspace_bytes(n);
  const size_t dwb = mplak_dot_workspace_bytes(n);
  const size_t ob = ordered ? mplak_spatial_order_workspace_bytes(n, dim) : 0;
  if (ordered && ob == 0) fail("mpla_kernel_conjugate_gradient_mixed", "spatial order workspace query failed");
  // ordered: inner CG in single-reduction form, whose inner products come
  // from the planned product's finish (direction s = A p and the partials)
  const int64_t nd = ordered ? mplak_planned_dots_count(n) : 0;
  const size_t cg1_bytes = ordered ? vb + align256(2 * sizeof(double) * (size_t)nd) : 0;
  // relaxed levels of the inner products: drop, drop - 5, drop - 9 (each at
  // least kMinRelaxDrop), one screening plan each
  constexpr double kRelaxBits[3] = {0.0, 5.0, 9.0};
  const double relax_below[2] = {std::ldexp(1.0, -5), std::ldexp(1.0, -9)};
  int nlev = 1;
  if (ordered && relax)
    while (nlev < 3 && drop - kRelaxBits[nlev] >= kMinRelaxDrop) ++nlev;
  const size_t order_bytes =
      ordered ? 3 * vb + (size_t)dim * vb + 256 + align256(ob) + (size_t)(nlev + 1) * align256(pb) : 0;
  char* base =
      static_cast<char*>(arena(inst, kMixed, 4 * vb + sb + wb + align256(dwb) + 256 + order_bytes + cg1_bytes));
  double* r = reinterpret_cast<double*>(base);
  double* d = reinterpret_cast<double*>(base + vb);
  double* z = reinterpret_cast<double*>(base + 2 * vb);
  double* e = reinterpret_cast<double*>(base + 3 * vb);
  mplak_cg_state* st = reinterpret_cast<mplak_cg_state*>(base + 4 * vb);
  void* work = base + 4 * vb + sb;
  char* dotw = base + 4 * vb + sb + wb;
  double* dscalar = reinterpret_cast<double*>(dotw + align256(dwb));
  mplak_stream_t s = stream_of(inst);
  Runtime* rt = runtime(inst);

  mplak_kernel_desc desc = desc0;
  double* xs = x->data;        // solution in solver order
  const double* bs = b->data;  // right-hand side in solver order
  int64_t* perm = nullptr;
  void* plans[3] = {nullptr, nullptr, nullptr};
  void* plan64 = nullptr;  // FP64 tile-screened plan of the residual products
  double drops[3] = {drop, drop, drop};
  auto build_plans = [&]() {
    for (int l = 0; l < nlev; ++l) {
      drops[l] = drop - kRelaxBits[l];
      ck(mplak_screen_plan_build(&desc, MPLAK_PREC_F32, drops[l], plans[l], pb, s), "mixed CG screening plan");
    }
  };
  double* sv = nullptr;
  double* dots = nullptr;
  if (ordered) {
    char* ob0 = dotw + align256(dwb) + 256;
    sv = reinterpret_cast<double*>(ob0 + order_bytes);
    dots = reinterpret_cast<double*>(ob0 + order_bytes + vb);
    perm = reinterpret_cast<int64_t*>(ob0);
    double* bp = reinterpret_cast<double*>(ob0 + vb);
    double* xp = reinterpret_cast<double*>(ob0 + 2 * vb);
    double* pts = reinterpret_cast<double*>(ob0 + 3 * vb);
    double** table = reinterpret_cast<double**>(ob0 + 3 * vb + (size_t)dim * vb);
    void* owork = ob0 + 3 * vb + (size_t)dim * vb + 256;
    for (int l = 0; l < nlev; ++l) plans[l] = ob0 + 3 * vb + (size_t)dim * vb + 256 + align256(ob) + l * align256(pb);
    plan64 = ob0 + 3 * vb + (size_t)dim * vb + 256 + align256(ob) + nlev * align256(pb);
    ck(mplak_spatial_order(desc0.points, dim, desc0.row_offset, n, perm, owork, ob, s), "mixed CG spatial order");
    const int64_t ld = (int64_t)(vb / sizeof(double));
    ck(mplak_gather_points(desc0.points, dim, desc0.row_offset, perm, n, pts, ld, s), "mixed CG point gather");
    double** host_table = static_cast<double**>(rt->pinned);
    for (int k = 0; k < dim; ++k) host_table[k] = pts + (size_t)k * ld;
    hip_ck(hipMemcpyAsync(table, host_table, sizeof(double*) * dim, hipMemcpyHostToDevice, rt->stream),
           "hipMemcpyAsync");
    ck(mplak_gather_f64(bp, b->data, perm, n, s), "mixed CG permute");
    ck(mplak_gather_f64(xp, x->data, perm, n, s), "mixed CG permute");
    desc.points = table;
    desc.row_offset = desc.col_offset = 0;
    xs = xp;
    bs = bp;
    build_plans();
    ck(mplak_screen_plan_build(&desc, MPLAK_PREC_F64, res_drop, plan64, pb, s),
       "mixed CG residual plan");
  }
  auto unpermute = [&]() {
    if (ordered) ck(mplak_scatter_f64(x->data, xs, perm, n, s), "mixed CG permute back");
  };
  // size the product workspaces now: they must not grow during graph capture
  (void)arena(inst, kCore,
              std::max(mplak_kernel_matvec_workspace_bytes(&desc), mplak_kernel_matvec_screened_workspace_bytes(&desc)));

  // an all-zero initial guess needs no FP64 product for its residual
  ck(mplak_ddot_f64(n, xs, xs, dscalar, dotw, dwb, s), "mixed CG");
  bool x_zero = *static_cast<double*>(read_back(inst, dscalar, sizeof(double))) == 0.0;

  const int max_outer = 40;
  int total_inner = 0, inner_products = 0, outer = 0, matvecs64 = 0;
  double prev_rel = INFINITY, prev_eta = 1.0;
  mplak_cg_state cur{};

  // inner CG iteration on A_f32 e = r (vectors and scalars in FP64).  On
  // ordered points: single-reduction CG, two launches per iteration (the
  // planned product on r with its finish leaving <r,r> and <r,Ar>, then one
  // vector kernel); otherwise classic CG around the screened/exact product.
  auto iteration = [&](int parity) {
    mplak_cg_state* in = st + parity;
    mplak_cg_state* out = st + (1 - parity);
    if (ordered) {
      ck(mplak_kernel_matvec_planned_relaxed(&desc, nlev, drops, plans, pb, &in->rel, relax_below, r, z, dots,
                                             &in->done, s),
         "mixed CG inner product");
      ck(mplak_cg1_iterate(n, e, r, d, sv, z, dots, nd, in, out, s), "mixed CG inner iteration");
      return;
    }
    apply_kernel(desc, MPLAK_PREC_F32, d, z, &in->done, inst, "mixed CG inner product");
    ck(mplak_cg_iterate(n, e, r, d, z, in, out, work, wb, s), "mixed CG inner iteration");
  };
  auto normalize = [&]() { ck(mplak_cg_copy_state(st + 1, st, s), "mixed CG state copy"); };
  // On ordered points the captured chunk is kept in the instance and
  // replayed by later solves whose kernels would get the same arguments (the
  // same operator shape and workspace), which saves its capture.
  const int chunk = choose_chunk(0, verbosity(inst));
  std::vector<unsigned char> key;
  auto put = [&key](const void* v, size_t bytes) {
    const unsigned char* c = static_cast<const unsigned char*>(v);
    key.insert(key.end(), c, c + bytes);
  };
  if (ordered) {
    const void* ptrs[] = {desc.points, plans[0], plans[1], plans[2], st, r, z, d, e, sv, dots, s};
    put(ptrs, sizeof(ptrs));
    const int64_t ints[] = {n, nd, (int64_t)pb, (int64_t)nlev, (int64_t)chunk, (int64_t)desc.dim,
                            (int64_t)desc.kind, desc.row_offset, desc.row_count, desc.col_offset, desc.col_count};
    put(ints, sizeof(ints));
    const double reals[] = {desc.length_scale, desc.nugget, desc.coincident_shift, drops[0], drops[1], drops[2],
                            relax_below[0], relax_below[1]};
    put(reals, sizeof(reals));
  }
  ChunkRunner runner(inst, true, chunk, iteration, normalize, ordered ? &rt->mixed_chunk : nullptr, key);

  bool stagnated = false;
  std::vector<int> step_iters;
  for (;;) {
    // r = b - A_f64 x, relative to ||b||: the true FP64 residual
    if (x_zero) {
      ck(mplak_fill_f64(z, n, 0.0, s), "mixed CG");
    } else {
      if (ordered) {
        ck(mplak_kernel_matvec_planned(&desc, MPLAK_PREC_F64, res_drop, plan64, pb, xs, z, nullptr, s),
           "mixed CG residual product");
      } else {
        apply_kernel(desc, MPLAK_PREC_F64, xs, z, nullptr, inst, "mixed CG residual product");
      }
      ++matvecs64;
    }
    ck(mplak_cg_start(n, bs, z, r, d, st, MPLAK_STOP_RELATIVE, epsilon, 2, n_global, work, wb, s),
       "mixed CG residual");
    cur = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    if (verbosity(inst) >= 2 && inst->is_parent)
      printf("Refinement step %d: true relative residual %le after %d inner iterations\n", outer, cur.rel,
             total_inner);
    if (cur.rel < epsilon) break;
    // a non-finite residual (non-finite input) or a step that gained less
    // than 2x: refinement cannot go further with this inner precision
    if (!std::isfinite(cur.rel) || (outer > 0 && cur.rel > 0.5 * prev_rel)) {
      stagnated = true;
      break;
    }
    if (total_inner >= iter_max || outer >= max_outer) break;
    if (ordered && outer > 0 && cur.rel > 16.0 * prev_eta * prev_rel && drop + 6.0 <= kMaxDrop) {
      // the screen, not eta, limited the last step: keep more entries
      drop += 6.0;
      build_plans();
    }
    prev_rel = cur.rel;

    // inner solve to ||r_inner|| <= eta ||r||, e_0 = 0 (so r_inner,0 = r, d_0 = r)
    const double eta = std::min(0.5, std::max(eta_floor, 0.5 * epsilon / cur.rel));
    prev_eta = eta;
    mplak_cg_state inner = cur;
    inner.bnorm2 = cur.rr;
    inner.rel = 1.0;
    inner.eps = eta;
    inner.k = 0;
    inner.iter_max = std::max(2, iter_max - total_inner + 1);
    inner.done = MPLAK_RUNNING;
    inner.rule = MPLAK_STOP_RELATIVE;
    *static_cast<mplak_cg_state*>(rt->pinned) = inner;
    hip_ck(hipMemcpyAsync(st, rt->pinned, sizeof(inner), hipMemcpyHostToDevice, rt->stream), "hipMemcpyAsync");
    ck(mplak_fill_f64(e, n, 0.0, s), "mixed CG");
    mplak_cg_state ist = inner;
    // the iterations this step took in the last solve of this size are
    // enqueued before the state is first read back: whole chunks, then the
    // rest one by one (not captured), so that a solve like the last one reads
    // the state once and launches nothing past its convergence; then one
    // chunk per read
    int hint = 0;
    if (rt->mixed_hint_n == n && outer < (int)rt->mixed_hint_iters.size())
      hint = std::max(0, rt->mixed_hint_iters[outer]);
    bool first_read = true;
    while (ist.done == MPLAK_RUNNING) {
      if (first_read && hint > 0) {
        for (int c = 0; c < hint / chunk; ++c) runner.run();
        if (hint % chunk != 0) {
          ChunkRunner rest(inst, false, hint % chunk, iteration, normalize);
          rest.run();
        }
      } else {
        runner.run();
      }
      first_read = false;
      ist = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    }
    // iterations this step needed: its updates, +1 for the converging
    // product of the single-reduction form
    step_iters.push_back(ist.k + 1);
    total_inner += ist.k;
    // the single-reduction form applies the operator once more than it
    // updates when it converges (the product whose <r,r> meets the test)
    inner_products += ist.k + (ordered && ist.done != MPLAK_ITER_LIMIT ? 1 : 0);
    // An inner solve that broke down (the screened FP32 operator lost
    // definiteness: a step length that is not finite and positive) stopped
    // before that update, so e is its last finite iterate.  A non-finite
    // inner residual from the classic form leaves e unusable: x keeps its
    // last finite value and the solve finishes in FP64.
    if (!std::isfinite(ist.rel) || (ist.done == MPLAK_BREAKDOWN && ist.k == 0)) {
      stagnated = true;
      break;
    }
    ck(mplak_daxpy_f64(n, 1.0, e, xs, s), "mixed CG update");
    x_zero = false;
    ++outer;
  }

  unpermute();
  if (ordered) {
    rt->mixed_hint_n = n;
    rt->mixed_hint_iters = step_iters;
  }
  int iterations = total_inner;
  double rel = cur.rel;
  bool converged = cur.rel < epsilon;
  if (stagnated && total_inner < iter_max) {
    // refinement cannot gain more with this inner precision: finish in plain
    // FP64 CG, to half the tolerance so that the drift between its recursive
    // residual and the true one cannot decide the outcome, then report the
    // true residual b - A x of the exact FP64 product
    plain_fp64_cg(b, A, x, iter_max - total_inner, 0.5 * epsilon, inst);
    iterations += inst->last_matvecs - 1;
    matvecs64 += inst->last_matvecs + 1;
    apply_kernel(desc0, MPLAK_PREC_F64, x->data, z, nullptr, inst, "mixed CG final residual");
    ck(mplak_cg_start(n, b->data, z, r, d, st, MPLAK_STOP_RELATIVE, epsilon, 2, n_global, work, wb, s),
       "mixed CG final residual");
    cur = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    rel = cur.rel;
    converged = rel < epsilon;
  }
  inst->last_iterations = iterations;
  inst->last_converged = converged;
  inst->last_rel_res = rel;
  inst->last_matvecs = inner_products + matvecs64;
  if (inst->is_parent && verbosity(inst) >= 1)
    printf("MPLA mixed CG stopped after %d inner iterations (%d refinement steps) with a final true relative "
           "residual norm of %le.\n",
           iterations, outer, rel);
}

}  // extern "C"
