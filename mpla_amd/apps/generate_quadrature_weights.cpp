// Kernel-quadrature weights on Halton nodes: the application driver of
// src/generate_quadrature_weights.cpp, on this library.
//
//   generate_quadrature_weights [n=1000] [dim=3] [tol=1e-7] [restart=0] [iterations_until_checkpoint=100]
//
// Nodes p_1..p_n are the first n Halton points in [0,1]^dim.  The weights w
// solve K w = b with K_ij = exp(-|p_i - p_j|^2) (+1e-8 on coincident points,
// the reference kernel core) and b_i = integral over [0,1]^dim of
// exp(-|y - p_i|^2) dy, so that sum_i w_i f(p_i) integrates every f in the
// span of the kernel translates exactly.  The solve is the checkpointing CG
// (relative residual rule) started from uniform random weights; with
// restart=1 it resumes from the checkpoint files x.dat, r.dat, d.dat, z.dat of
// an earlier run.  Prints the solve time and writes weights.dat (int32 n, then
// n float64), as the reference does (:139-160).

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "kernel.h"
#include "mpla.h"
#include "mplak.h"
#include "quadrature.h"

namespace {

void hip_or_die(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    exit(EXIT_FAILURE);
  }
}

// Row panel height from free device memory, as the reference computes it
// (:31-39); the fused cores never materialise a panel, so it is informational.
int max_row_count(int cols) {
  size_t free_bytes = 0, total_bytes = 0;
  hip_or_die(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
  const double rows = (double)(free_bytes / sizeof(double) / (size_t)cols) * 0.95;
  return rows > 2147483647.0 ? 2147483647 : (int)rows;
}

}  // namespace

int main(int argc, char* argv[]) {
  int n = argc > 1 ? atoi(argv[1]) : 1000;
  int dim = argc > 2 ? atoi(argv[2]) : 3;
  double tol = argc > 3 ? atof(argv[3]) : 1.0e-7;
  int restart = argc > 4 ? (int)atof(argv[4]) : 0;
  int iterations_until_checkpoint = argc > 5 ? atoi(argv[5]) : 100;
  if (n < 1 || dim < 1 || dim > 16) {
    fprintf(stderr, "usage: %s [n >= 1] [1 <= dim <= 16] [tol] [restart] [iterations_until_checkpoint]\n", argv[0]);
    return EXIT_FAILURE;
  }

  MPI_Init(&argc, &argv);
  mpla_instance instance;
  mpla_init_instance(&instance, MPI_COMM_WORLD);
  mpla_generic_matrix A;
  mpla_init_generic_matrix(&A, &instance, n, n);
  mpla_vector x, b;
  mpla_init_vector(&x, &instance, n);
  mpla_init_vector(&b, &instance, n);

  // Halton nodes, SoA on the host, then one device array per coordinate and a
  // device table of those pointers (the layout kernel_matrix_data expects)
  std::vector<double> nodes((size_t)dim * n);
  if (mpla_halton_points(nodes.data(), dim, n, 0, n) != 0) {
    fprintf(stderr, "mpla_halton_points failed\n");
    return EXIT_FAILURE;
  }
  std::vector<double*> coord(dim, nullptr);
  for (int d = 0; d < dim; ++d) {
    hip_or_die(hipMalloc(&coord[d], sizeof(double) * n), "hipMalloc");
    hip_or_die(hipMemcpy(coord[d], nodes.data() + (size_t)d * n, sizeof(double) * n, hipMemcpyHostToDevice),
               "hipMemcpy");
  }
  double** points = nullptr;
  hip_or_die(hipMalloc(&points, sizeof(double*) * dim), "hipMalloc");
  hip_or_die(hipMemcpy(points, coord.data(), sizeof(double*) * dim, hipMemcpyHostToDevice), "hipMemcpy");

  kernel_matrix_data matrix_data;
  matrix_data.points = points;
  matrix_data.dim = dim;
  matrix_data.max_row_count_per_dgemv = max_row_count(A.cur_proc_col_count);
  A.data = &matrix_data;

  mpla_set_gaussian_kernel_rhs(&b, &A, &instance);
  // initial weights ~ U[0,1) (the reference draws them with cuRAND, seed 1235)
  if (mplak_fill_uniform_f64(x.data, x.cur_proc_row_count, 1235, 0, 0, 0.0, 1.0, mpla_instance_stream(&instance)) !=
      MPLAK_OK) {
    fprintf(stderr, "mplak_fill_uniform_f64 failed\n");
    return EXIT_FAILURE;
  }

  hipEvent_t start, stop;
  hip_or_die(hipEventCreate(&start), "hipEventCreate");
  hip_or_die(hipEventCreate(&stop), "hipEventCreate");
  hip_or_die(hipEventRecord(start, 0), "hipEventRecord");
  mpla_generic_conjugate_gradient_with_checkpoint_restart(&b, &A, &x, 100000, tol,
                                                          mpla_dgemv_core_kernel_streamed_cublas,
                                                          iterations_until_checkpoint, restart != 0, &instance);
  hip_or_die(hipEventRecord(stop, 0), "hipEventRecord");
  hip_or_die(hipEventSynchronize(stop), "hipEventSynchronize");
  float ms = 0.0f;
  hip_or_die(hipEventElapsedTime(&ms, start, stop), "hipEventElapsedTime");
  printf("Time spent: %.10f\n", ms / 1000.0);

  char filename[] = "weights.dat";
  mpla_save_vector(&x, filename, &instance);

  mpla_free_generic_matrix(&A, &instance);
  mpla_free_vector(&x, &instance);
  mpla_free_vector(&b, &instance);
  for (int d = 0; d < dim; ++d) hip_or_die(hipFree(coord[d]), "hipFree");
  hip_or_die(hipFree(points), "hipFree");
  hip_or_die(hipEventDestroy(start), "hipEventDestroy");
  hip_or_die(hipEventDestroy(stop), "hipEventDestroy");
  mpla_destroy_instance(&instance);
  MPI_Finalize();
  return 0;
}
