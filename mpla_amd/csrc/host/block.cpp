// Multi-right-hand-side kernel CG with FP64 iterative refinement (no
// reference counterpart; the reference solves one right-hand side at a time,
// src/mpla.cpp:887-1001).
//
// Outer loop (FP64): R = B - K X with the FP64 multi-RHS product, per-column
// true relative residuals.  Inner solve: batched CG on K_low E = R, one shared
// reduced-precision product (BF16 / BF16x3 / FP32 MFMA) per iteration, each
// column with its own scalars and its own tolerance; columns that already
// meet epsilon are frozen.  X += E.  If the residual stops falling by at least
// half per refinement step, the inner precision is raised (BF16 -> BF16x3 ->
// FP32 -> FP64), so the loop always reaches epsilon in FP64 terms.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "../../../include/mpla_kernel.h"
#include "chunk.hpp"
#include "runtime.hpp"

using namespace mpla_rt;

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int to_mplak(int prec) {
  switch (prec) {
    case MPLA_PREC_BF16: return MPLAK_PREC_BF16;
    case MPLA_PREC_BF16X3: return MPLAK_PREC_BF16X3;
    case MPLA_PREC_F32: return MPLAK_PREC_F32;
    default: return MPLAK_PREC_F64;
  }
}

int escalate(int prec) {
  switch (prec) {
    case MPLA_PREC_BF16: return MPLA_PREC_BF16X3;
    case MPLA_PREC_BF16X3: return MPLA_PREC_F32;
    default: return MPLA_PREC_F64;
  }
}

double env_double(const char* name, double dflt) {
  const char* v = getenv(name);
  return v ? atof(v) : dflt;
}

// inner relative tolerance floor per precision: below the product's own
// accuracy further inner iterations do not improve the refinement step
double eta_floor(int prec) {
  switch (prec) {
    case MPLA_PREC_BF16: return env_double("MPLA_IR_TOL_BF16", 1e-2);
    case MPLA_PREC_BF16X3: return env_double("MPLA_IR_TOL_BF16X3", 1e-4);
    case MPLA_PREC_F32: return env_double("MPLA_IR_TOL_F32", 1e-5);
    default: return 0.0;
  }
}

}  // namespace

extern "C" {

void mpla_kernel_block_conjugate_gradient_mixed(double* B, double* X, int nrhs, struct mpla_generic_matrix* A,
                                                int iter_max, double epsilon, int inner_precision,
                                                double* rel_res, struct mpla_instance* inst) {
  if (nrhs <= 0) return;
  const mpla_kernel_matrix_data* kd = static_cast<const mpla_kernel_matrix_data*>(A->data);
  mplak_kernel_desc desc;
  desc.points = kd->points;
  desc.dim = kd->dim;
  desc.kind = kd->kernel;
  desc.length_scale = kd->length_scale;
  desc.nugget = kd->nugget;
  desc.coincident_shift = kd->coincident_shift;
  desc.row_offset = A->cur_proc_row_offset;
  desc.row_count = A->cur_proc_row_count;
  desc.col_offset = A->cur_proc_col_offset;
  desc.col_count = A->cur_proc_col_count;
  if (desc.row_count != desc.col_count)
    fail("mpla_kernel_block_conjugate_gradient_mixed", "needs the square local block of a 1 x 1 grid");

  const int64_t n = desc.row_count;
  const int64_t nn = n * (int64_t)nrhs;
  const size_t vb = align256(sizeof(double) * (size_t)nn);
  const size_t sb = align256(2 * sizeof(mplak_cg_state) * (size_t)nrhs);
  const size_t gb = 256;
  const size_t wb = mplak_bcg_workspace_bytes(n, nrhs);
  const size_t dwb = mplak_dot_workspace_bytes(nn);
  char* base = static_cast<char*>(arena(inst, kMixed, 4 * vb + sb + gb + align256(wb) + align256(dwb) + 256));
  double* R = reinterpret_cast<double*>(base);
  double* D = reinterpret_cast<double*>(base + vb);
  double* Z = reinterpret_cast<double*>(base + 2 * vb);
  double* E = reinterpret_cast<double*>(base + 3 * vb);
  mplak_cg_state* st = reinterpret_cast<mplak_cg_state*>(base + 4 * vb);
  mplak_bcg_group* group = reinterpret_cast<mplak_bcg_group*>(base + 4 * vb + sb);
  void* work = base + 4 * vb + sb + gb;
  char* dotw = base + 4 * vb + sb + gb + align256(wb);
  double* dscalar = reinterpret_cast<double*>(dotw + align256(dwb));
  mplak_stream_t s = stream_of(inst);
  Runtime* rt = runtime(inst);

  // product workspace: the largest any precision of this solve needs, sized
  // now because it must not grow while a graph is being captured
  size_t pw = 0;
  for (int p = MPLAK_PREC_F64; p <= MPLAK_PREC_BF16X3; ++p)
    pw = std::max(pw, mplak_kernel_mrhs_workspace_bytes(&desc, p, nrhs));
  void* pwork = arena(inst, kCore, pw);
  auto product = [&](int prec, const double* V, double* Y, const int32_t* skip) {
    ck(mplak_kernel_mrhs(&desc, prec, nrhs, V, n, Y, n, pwork, pw, skip, s), "block CG product");
  };

  ck(mplak_ddot_f64(nn, X, X, dscalar, dotw, dwb, s), "block CG");
  bool x_zero = *static_cast<double*>(read_back(inst, dscalar, sizeof(double))) == 0.0;

  std::vector<mplak_cg_state> host(nrhs);
  auto read_states = [&]() {
    const size_t bytes = sizeof(mplak_cg_state) * (size_t)nrhs;
    hip_ck(hipMemcpyAsync(host.data(), st, bytes, hipMemcpyDeviceToHost, rt->stream), "hipMemcpyAsync");
    hip_ck(hipStreamSynchronize(rt->stream), "hipStreamSynchronize");
  };

  int prec = inner_precision;
  std::unique_ptr<ChunkRunner> runner;
  int runner_prec = -1;
  int total_inner = 0, outer = 0, products64 = 0;
  double prev_max = INFINITY, max_rel = INFINITY;
  bool all_conv = false;
  const int max_outer = 60;
  const int max_bf16_steps = (int)env_double("MPLA_IR_BF16_STEPS", 1.0);
  int bf16_steps = 0;

  for (;;) {
    // true FP64 residuals R = B - K X
    if (x_zero) {
      ck(mplak_fill_f64(Z, nn, 0.0, s), "block CG");
    } else {
      product(MPLAK_PREC_F64, X, Z, nullptr);
      ++products64;
    }
    ck(mplak_bcg_start(n, nrhs, B, Z, R, D, st, group, MPLAK_STOP_RELATIVE, epsilon, 2, (double)n, work, wb, s),
       "block CG residual");
    read_states();
    max_rel = 0.0;
    all_conv = true;
    for (int c = 0; c < nrhs; ++c) {
      max_rel = std::max(max_rel, host[c].rel);
      if (host[c].rel >= epsilon || host[c].rel != host[c].rel) all_conv = false;
    }
    if (verbosity(inst) >= 2 && inst->is_parent)
      printf("Block refinement step %d: max true relative residual %le (inner precision %d, %d inner iterations)\n",
             outer, max_rel, prec, total_inner);
    if (all_conv || total_inner >= iter_max || outer >= max_outer) break;
    if (outer > 0 && max_rel > 0.5 * prev_max) prec = escalate(prec);
    // BF16 entries bound each refinement step to a gain of about its inner
    // tolerance floor (1e-2), and every step restarts the Krylov space: after
    // its first step a BF16 solve costs more per decade than BF16x3 at 1.4x the
    // product time (measured on the 64-RHS Matern system), so it continues
    // in BF16x3 (MPLA_IR_BF16_STEPS sets how many steps stay in BF16).
    if (prec == MPLA_PREC_BF16 && bf16_steps >= max_bf16_steps) prec = MPLA_PREC_BF16X3;
    prev_max = max_rel;

    // inner batched CG on K_prec E = R
    const double floor_eta = eta_floor(prec);
    for (int c = 0; c < nrhs; ++c) {
      mplak_cg_state& h = host[c];
      if (!(h.rel >= epsilon)) {
        h.done = MPLAK_CONVERGED;  // frozen: E stays 0 in this column
        continue;
      }
      h.bnorm2 = h.rr;
      h.eps = std::min(0.5, std::max(floor_eta, 0.5 * epsilon / h.rel));
      h.rel = 1.0;
      h.k = 0;
      h.iter_max = std::max(2, iter_max - total_inner + 1);
      h.rule = MPLAK_STOP_RELATIVE;
      h.done = MPLAK_RUNNING;
    }
    hip_ck(hipMemcpyAsync(st, host.data(), sizeof(mplak_cg_state) * (size_t)nrhs, hipMemcpyHostToDevice, rt->stream),
           "hipMemcpyAsync");
    ck(mplak_bcg_copy_states(st, st, nrhs, group, s), "block CG");
    ck(mplak_fill_f64(E, nn, 0.0, s), "block CG");
    if (!runner || runner_prec != prec) {
      const int mp = to_mplak(prec);
      auto iteration = [&, mp](int parity) {
        mplak_cg_state* in = st + (size_t)parity * nrhs;
        mplak_cg_state* out = st + (size_t)(1 - parity) * nrhs;
        product(mp, D, Z, &group->all_done);
        ck(mplak_bcg_iterate(n, nrhs, E, R, D, Z, in, out, group, work, wb, s), "block CG iteration");
      };
      auto normalize = [&]() { ck(mplak_bcg_copy_states(st + nrhs, st, nrhs, group, s), "block CG"); };
      runner.reset(new ChunkRunner(inst, true, choose_chunk(0, verbosity(inst)), iteration, normalize));
      runner_prec = prec;
    }
    for (;;) {
      runner->run();
      const mplak_bcg_group g = *static_cast<mplak_bcg_group*>(read_back(inst, group, sizeof(mplak_bcg_group)));
      if (g.all_done) break;
    }
    read_states();
    int kmax = 0;
    for (int c = 0; c < nrhs; ++c) kmax = std::max(kmax, host[c].k);
    total_inner += kmax;
    ck(mplak_daxpy_f64(nn, 1.0, E, X, s), "block CG update");
    x_zero = false;
    if (prec == MPLA_PREC_BF16) ++bf16_steps;
    ++outer;
  }
  if (rel_res) {
    for (int c = 0; c < nrhs; ++c) rel_res[c] = host[c].rel;
  }
  inst->last_iterations = total_inner;
  inst->last_converged = all_conv;
  inst->last_rel_res = max_rel;
  inst->last_matvecs = total_inner + products64;
  if (inst->is_parent && verbosity(inst) >= 1)
    printf("MPLA block CG stopped after %d shared products (%d refinement steps, final inner precision %d) with a "
           "largest true relative residual norm of %le.\n",
           total_inner, outer, prec, max_rel);
}

}  // extern "C"
