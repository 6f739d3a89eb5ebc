// Krylov solvers of the drop-in API (src/mpla.cpp:817-1183 of the reference).
//
// Each iteration is the plugin core plus the fused kernels of mplak.h with
// all scalars on the device.  The host only looks at the solver state once
// per chunk of iterations.  When the core is capture-safe, the chunk is
// recorded once into a HIP graph and replayed; kernels of iterations after
// convergence see `done` and return at once, so replaying a whole chunk never
// changes the result.  Chunk boundaries fall on checkpoint iterations, so the
// checkpoint files hold exactly the vectors the reference writes.

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>

#include "chunk.hpp"
#include "runtime.hpp"

using namespace mpla_rt;

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

int reported_k(int done, int k, int iter_max) {
  // the reference prints its loop counter: the converging k, or the value it
  // has after running to completion
  if (done == MPLAK_ITER_LIMIT) return std::max(1, iter_max);
  return k;
}

void save_as(mpla_vector* v, const char* name, mpla_instance* inst) {
  char buf[64];
  snprintf(buf, sizeof buf, "%s", name);
  mpla_save_vector(v, buf, inst);
}

void load_as(mpla_vector* v, const char* name, mpla_instance* inst) {
  char buf[64];
  snprintf(buf, sizeof buf, "%s", name);
  mpla_load_vector(v, buf, inst);
}

// ---------------------------------------------------------------------------
// CG: both stopping rules, optional checkpoint/restart.

void run_cg(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, int iter_max, double eps,
            mpla_dgemv_core_t core, int rule, int ckpt_every, bool restart, bool print_final,
            mpla_instance* inst) {
  const int64_t n = x->cur_proc_row_count;
  const size_t vb = align256(sizeof(double) * (size_t)n);
  const size_t sb = align256(2 * sizeof(mplak_cg_state));
  const size_t wb = mplak_cg_workspace_bytes(n);
  char* base = static_cast<char*>(arena(inst, kSolver, 3 * vb + sb + wb));
  mpla_vector r = vector_view(x, reinterpret_cast<double*>(base));
  mpla_vector d = vector_view(x, reinterpret_cast<double*>(base + vb));
  mpla_vector z = vector_view(x, reinterpret_cast<double*>(base + 2 * vb));
  mplak_cg_state* st = reinterpret_cast<mplak_cg_state*>(base + 3 * vb);
  void* work = base + 3 * vb + sb;
  mplak_stream_t s = stream_of(inst);
  const int verbose = verbosity(inst);

  // r_0 = b - A x_0, d_0 = r_0 (src/mpla.cpp:829-840, 904-915)
  inst->skip_flag = nullptr;
  generic_apply(&z, A, x, core, inst);
  ck(mplak_cg_start(n, b->data, z.data, r.data, d.data, st, rule, eps, iter_max, (double)x->vec_row_count,
                    work, wb, s),
     "CG start");
  if (restart) {
    // loaded after the initial residual is formed, as src/mpla.cpp:922-933
    load_as(x, "x.dat", inst);
    load_as(&r, "r.dat", inst);
    load_as(&d, "d.dat", inst);
    load_as(&z, "z.dat", inst);
    ck(mplak_cg_refresh_rr(n, r.data, st, work, wb, s), "CG restart");
  }
  mplak_cg_state cur = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));

  auto iteration = [&](int parity) {
    mplak_cg_state* in = st + parity;
    mplak_cg_state* out = st + (1 - parity);
    inst->skip_flag = &in->done;
    generic_apply(&z, A, &d, core, inst);
    inst->skip_flag = nullptr;
    ck(mplak_cg_iterate(n, x->data, r.data, d.data, z.data, in, out, work, wb, s), "CG iteration");
  };
  auto normalize = [&]() { ck(mplak_cg_copy_state(st + 1, st, s), "CG state copy"); };
  ChunkRunner runner(inst, is_graph_safe(core), choose_chunk(ckpt_every, verbose), iteration, normalize);

  while (cur.done == MPLAK_RUNNING) {
    runner.run();
    cur = *static_cast<mplak_cg_state*>(read_back(inst, st, sizeof(mplak_cg_state)));
    if (verbose >= 2 && inst->is_parent) {
      if (rule == MPLAK_STOP_RMS && cur.k % 10 == 0) printf("%d: %e\n", cur.k, cur.rel);
      if (rule == MPLAK_STOP_RELATIVE && cur.k % 2 == 0)
        printf("Iteration: %d\t Relative residual norm: %le\n", cur.k, cur.rel);
    }
    if (cur.done == MPLAK_CONVERGED) break;
    if (ckpt_every > 0 && cur.k % ckpt_every == 0) {
      // src/mpla.cpp:977-991; after the direction update z holds d
      if (inst->is_parent && verbose >= 1) printf("Checkpointing...\n");
      save_as(x, "x.dat", inst);
      save_as(&r, "r.dat", inst);
      save_as(&d, "d.dat", inst);
      save_as(&d, "z.dat", inst);
    }
  }
  const int k = reported_k(cur.done, cur.k, iter_max);
  inst->last_iterations = k;
  inst->last_converged = cur.done == MPLAK_CONVERGED;
  inst->last_rel_res = cur.rel;
  inst->last_matvecs = 1 + cur.k;
  if (print_final && inst->is_parent && verbose >= 1)
    printf("MPLA CG stopped after %d iterations with a final relative residual norm of %le.\n", k, cur.rel);
}

// ---------------------------------------------------------------------------
// BiCGSTAB with the reference's true-residual test (3 products per iteration).

void run_bicgstab(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, int iter_max, double eps,
                  mpla_dgemv_core_t core, int ckpt_every, bool restart, mpla_instance* inst) {
  const int64_t n = x->cur_proc_row_count;
  const size_t vb = align256(sizeof(double) * (size_t)n);
  const size_t sb = align256(2 * sizeof(mplak_bicgstab_state));
  const size_t wb = mplak_bicgstab_workspace_bytes(n);
  char* base = static_cast<char*>(arena(inst, kSolver, 6 * vb + sb + wb));
  mpla_vector r = vector_view(x, reinterpret_cast<double*>(base));
  mpla_vector r_hat = vector_view(x, reinterpret_cast<double*>(base + vb));
  mpla_vector p = vector_view(x, reinterpret_cast<double*>(base + 2 * vb));
  mpla_vector z = vector_view(x, reinterpret_cast<double*>(base + 3 * vb));
  mpla_vector zz = vector_view(x, reinterpret_cast<double*>(base + 4 * vb));
  mpla_vector sv = vector_view(x, reinterpret_cast<double*>(base + 5 * vb));
  mplak_bicgstab_state* st = reinterpret_cast<mplak_bicgstab_state*>(base + 6 * vb);
  void* work = base + 6 * vb + sb;
  mplak_stream_t s = stream_of(inst);
  const int verbose = verbosity(inst);

  // src/mpla.cpp:1022-1060
  inst->skip_flag = nullptr;
  generic_apply(&z, A, x, core, inst);
  ck(mplak_fill_f64(zz.data, n, 0.0, s), "BiCGSTAB start");
  ck(mplak_fill_f64(sv.data, n, 0.0, s), "BiCGSTAB start");
  ck(mplak_bicgstab_start(n, b->data, z.data, r.data, r_hat.data, p.data, st, eps, iter_max, 1, work, wb, s),
     "BiCGSTAB start");
  if (restart) {
    load_as(x, "x.dat", inst);
    load_as(&r, "r.dat", inst);
    load_as(&r_hat, "r_hat.dat", inst);
    load_as(&z, "z.dat", inst);
    load_as(&zz, "zz.dat", inst);
    load_as(&sv, "s.dat", inst);
    load_as(&p, "p.dat", inst);
    ck(mplak_bicgstab_refresh_rho(n, r.data, r_hat.data, st, work, wb, s), "BiCGSTAB restart");
  }
  mplak_bicgstab_state cur =
      *static_cast<mplak_bicgstab_state*>(read_back(inst, st, sizeof(mplak_bicgstab_state)));

  auto iteration = [&](int parity) {
    mplak_bicgstab_state* in = st + parity;
    mplak_bicgstab_state* out = st + (1 - parity);
    inst->skip_flag = &in->done;
    generic_apply(&z, A, &p, core, inst);  // z = A p      (src/mpla.cpp:1093)
    ck(mplak_bicgstab_step_s(n, r.data, z.data, sv.data, r_hat.data, in, work, wb, s), "BiCGSTAB");
    generic_apply(&zz, A, &sv, core, inst);  // zz = A s   (:1106)
    ck(mplak_bicgstab_step_x(n, x->data, r.data, r_hat.data, p.data, sv.data, z.data, zz.data, in, out, work,
                             wb, s),
       "BiCGSTAB");
    generic_apply(&z, A, x, core, inst);  // z = A x       (:1133)
    inst->skip_flag = nullptr;
    ck(mplak_bicgstab_step_true(n, b->data, z.data, zz.data, in, out, work, wb, s), "BiCGSTAB");
  };
  auto normalize = [&]() { ck(mplak_bicgstab_copy_state(st + 1, st, s), "BiCGSTAB state copy"); };
  ChunkRunner runner(inst, is_graph_safe(core), choose_chunk(ckpt_every, verbose), iteration, normalize);

  while (cur.done == MPLAK_RUNNING) {
    runner.run();
    cur = *static_cast<mplak_bicgstab_state*>(read_back(inst, st, sizeof(mplak_bicgstab_state)));
    if (verbose >= 2 && inst->is_parent)
      printf("Iteration: %d\t Relative residual norm: %le\n", cur.k, cur.rel);
    if (cur.done == MPLAK_CONVERGED) break;
    if (ckpt_every > 0 && cur.k % ckpt_every == 0) {
      // src/mpla.cpp:1152-1170
      if (inst->is_parent && verbose >= 1) printf("Checkpointing...\n");
      save_as(x, "x.dat", inst);
      save_as(&r, "r.dat", inst);
      save_as(&r_hat, "r_hat.dat", inst);
      save_as(&z, "z.dat", inst);
      save_as(&zz, "zz.dat", inst);
      save_as(&sv, "s.dat", inst);
      save_as(&p, "p.dat", inst);
    }
  }
  const int k = reported_k(cur.done, cur.k, iter_max);
  inst->last_iterations = k;
  inst->last_converged = cur.done == MPLAK_CONVERGED;
  inst->last_rel_res = cur.rel;
  inst->last_matvecs = 1 + 3 * cur.k;
  if (inst->is_parent && verbose >= 1)
    printf("MPLA CG stopped after %d iterations with a final relative residual norm of %le.\n", k, cur.rel);
}

}  // namespace

extern "C" {

void mpla_generic_conjugate_gradient(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x,
                                     int iter_max, double epsilon,
                                     void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*,
                                                             struct mpla_vector*, struct mpla_instance*),
                                     struct mpla_instance* instance) {
  run_cg(b, A, x, iter_max, epsilon, mpla_dgemv_core, MPLAK_STOP_RMS, 0, false, false, instance);
}

void mpla_generic_conjugate_gradient_with_checkpoint_restart(
    struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon,
    void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*,
                            struct mpla_instance*),
    int iterations_until_checkpoint, bool restart, struct mpla_instance* instance) {
  run_cg(b, A, x, iter_max, epsilon, mpla_dgemv_core, MPLAK_STOP_RELATIVE, iterations_until_checkpoint, restart,
         true, instance);
}

void mpla_generic_BiCGSTAB_with_checkpoint_restart(
    struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon,
    void (*mpla_dgemv_core)(struct mpla_vector*, struct mpla_generic_matrix*, struct mpla_vector*,
                            struct mpla_instance*),
    int iterations_until_checkpoint, bool restart, struct mpla_instance* instance) {
  run_bicgstab(b, A, x, iter_max, epsilon, mpla_dgemv_core, iterations_until_checkpoint, restart, instance);
}

}  // extern "C"
