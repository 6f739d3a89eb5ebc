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
    plain_fp64_cg(b, A, x, iter_max, epsilon, inst);
    return;
  }
  if (inner_precision != MPLA_PREC_F32)
    fail("mpla_kernel_conjugate_gradient_mixed", "inner precision must be MPLA_PREC_F32 or MPLA_PREC_F64");
  (void)kd;

  const mplak_kernel_desc desc = fused_desc(A);
  const int64_t n = x->cur_proc_row_count;
  const double n_global = (double)x->vec_row_count;
  const size_t vb = align256(sizeof(double) * (size_t)n);
  const size_t sb = align256(2 * sizeof(mplak_cg_state));
  const size_t wb = mplak_cg_workspace_bytes(n);
  const size_t dwb = mplak_dot_workspace_bytes(n);
  char* base = static_cast<char*>(arena(inst, kMixed, 4 * vb + sb + wb + align256(dwb) + 256));
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
  // size the product workspace now: it must not grow during graph capture
  (void)arena(inst, kCore, mplak_kernel_matvec_workspace_bytes(&desc));

  // an all-zero initial guess needs no FP64 product for its residual
  ck(mplak_ddot_f64(n, x->data, x->data, dscalar, dotw, dwb, s), "mixed CG");
  bool x_zero = *static_cast<double*>(read_back(inst, dscalar, sizeof(double))) == 0.0;

  const double eta_floor = env_double("MPLA_IR_INNER_TOL", 1.0e-5);
  const int max_outer = 40;
  int total_inner = 0, outer = 0, matvecs64 = 0;
  double prev_rel = INFINITY;
  mplak_cg_state cur{};

  // inner CG iteration on A_f32 e = r (vectors and scalars in FP64)
  auto iteration = [&](int parity) {
    mplak_cg_state* in = st + parity;
    mplak_cg_state* out = st + (1 - parity);
    apply_kernel(desc, MPLAK_PREC_F32, d, z, &in->done, inst, "mixed CG inner product");
    ck(mplak_cg_iterate(n, e, r, d, z, in, out, work, wb, s), "mixed CG inner iteration");
  };
  auto normalize = [&]() { ck(mplak_cg_copy_state(st + 1, st, s), "mixed CG state copy"); };
  ChunkRunner runner(inst, true, choose_chunk(0, verbosity(inst)), iteration, normalize);

  bool stagnated = false;
  for (;;) {
    // r = b - A_f64 x, relative to ||b||: the true FP64 residual
    if (x_zero) {
      ck(mplak_fill_f64(z, n, 0.0, s), "mixed CG");
    } else {
      apply_kernel(desc, MPLAK_PREC_F64, x->data, z, nullptr, inst, "mixed CG residual product");
      ++matvecs64;
    }
    ck(mplak_cg_start(n, b->data, z, r, d, st, MPLAK_STOP_RELATIVE, epsilon, 2, n_global, work, wb, s),
       "mixed CG residual");
    cur = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    if (verbosity(inst) >= 2 && inst->is_parent)
      printf("Refinement step %d: true relative residual %le after %d inner iterations\n", outer, cur.rel,
             total_inner);
    if (!(cur.rel >= epsilon)) break;
    if (outer > 0 && cur.rel > 0.5 * prev_rel) {
      stagnated = true;
      break;
    }
    if (total_inner >= iter_max || outer >= max_outer) break;
    prev_rel = cur.rel;

    // inner solve to ||r_inner|| <= eta ||r||, e_0 = 0 (so r_inner,0 = r, d_0 = r)
    const double eta = std::min(0.5, std::max(eta_floor, 0.5 * epsilon / cur.rel));
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
    while (ist.done == MPLAK_RUNNING) {
      runner.run();
      ist = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    }
    total_inner += ist.k;
    ck(mplak_daxpy_f64(n, 1.0, e, x->data, s), "mixed CG update");
    x_zero = false;
    ++outer;
  }

  int iterations = total_inner;
  double rel = cur.rel;
  bool converged = !(cur.rel >= epsilon);
  if (stagnated && total_inner < iter_max) {
    // refinement cannot gain more with this inner precision: finish in FP64
    plain_fp64_cg(b, A, x, iter_max - total_inner, epsilon, inst);
    iterations += inst->last_matvecs - 1;
    matvecs64 += inst->last_matvecs;
    rel = inst->last_rel_res;
    converged = inst->last_converged != 0;
  }
  inst->last_iterations = iterations;
  inst->last_converged = converged;
  inst->last_rel_res = rel;
  inst->last_matvecs = total_inner + matvecs64;
  if (inst->is_parent && verbosity(inst) >= 1)
    printf("MPLA mixed CG stopped after %d inner iterations (%d refinement steps) with a final true relative "
           "residual norm of %le.\n",
           iterations, outer, rel);
}

}  // extern "C"
