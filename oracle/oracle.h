/*
 * oracle.h -- CPU restatement of MPLA's solver hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be called by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * never by the product (mpla_amd/, libmpla, libmplak).  It restates the
 * reference's algorithms in plain C (OpenMP for the O(N^2) loops), following
 * the reference files cited per function (paths relative to the MPLA source
 * tree).  Vectors are plain host arrays; matrices are column-major.
 *
 * Pinning (what checks each function against something other than itself) is
 * listed in the header comment of oracle.c.
 */
#ifndef MPLA_ORACLE_H
#define MPLA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- inputs: SplitMix64 counter generator (same recipe as the product) --- */
uint64_t oracle_rng_key(uint64_t seed, uint64_t stream);
uint64_t oracle_rng_u64(uint64_t key, uint64_t i);
void oracle_fill_uniform(double* out, int64_t n, uint64_t seed, uint64_t stream,
                         uint64_t first_index, double lo, double hi);
/* Box-Muller normals from outputs 2i, 2i+1 of the (seed, stream) generator. */
void oracle_fill_normal(double* out, int64_t n, uint64_t seed, uint64_t stream,
                        uint64_t first_index);

/* ---- block distribution (src/mpla.cpp:157-209, 338-367, 390-439) -------- */
/* counts[p] = floor(M/P) + (p < M mod P), offsets[p] = prefix sum; returns -1
 * when floor(M/P) == 0 (the reference prints and exits). */
int oracle_block_counts(int64_t M, int P, int64_t* counts, int64_t* offsets);

/* ---- dense level 1/2 (src/mpla.cpp:716-815) ------------------------------ */
/* y = A x, A m x n column-major with leading dimension lda.
 * mode 0: plain double sum in column order (the CPU baseline);
 * mode 1: long double accumulation (the accuracy oracle). */
void oracle_dgemv(int64_t m, int64_t n, const double* A, int64_t lda, const double* x, double* y,
                  int mode);
double oracle_ddot(int64_t n, const double* x, const double* y, int mode);
void oracle_daxpy(int64_t n, double alpha, const double* x, double* y);

/* ---- kernel-matrix entries and products (src/kernel.cu:7-68) ------------- */
#define ORACLE_KERNEL_GAUSSIAN 0
#define ORACLE_KERNEL_MATERN32 1

typedef struct oracle_kernel {
  const double* points; /* SoA, points[d * n_points + i] */
  int64_t n_points;
  int dim;
  int kind;
  double length_scale;
  double nugget;           /* added iff global row == global column */
  double coincident_shift; /* added iff distance < 1e-14 */
} oracle_kernel;

double oracle_kernel_entry(const oracle_kernel* k, int64_t i, int64_t j);

/* y[i] = sum_j K(row_off+i, col_off+j) x[j]; mode as oracle_dgemv. */
void oracle_kernel_matvec(const oracle_kernel* k, int64_t row_off, int64_t rows, int64_t col_off,
                          int64_t cols, const double* x, double* y, int mode);

/* b_i = prod_d (sqrt(pi)/2)(erf(1 - p_d) - erf(-p_d)) (src/kernel.cu:52-68) */
void oracle_gaussian_rhs(const oracle_kernel* k, int64_t row_off, int64_t rows, double* b);

/* ---- operators for the solvers ------------------------------------------ */
typedef struct oracle_op {
  int type; /* 0: dense A (column-major, lda), 1: kernel matrix */
  int64_t n;
  const double* A;
  int64_t lda;
  oracle_kernel kernel;
  int mode; /* accumulation mode of the product */
} oracle_op;

void oracle_apply(const oracle_op* op, const double* x, double* y);

/* Solver outcome. */
typedef struct oracle_solve_info {
  int iterations;   /* the reference's loop counter k at exit */
  int converged;    /* 1 if the stopping test fired */
  double final_rel; /* the stopping quantity at exit */
  int matvecs;      /* operator applications */
} oracle_solve_info;

/* CG with the RMS rule sqrt(<r,r>/N) < eps (src/mpla.cpp:817-885). */
void oracle_cg_rms(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                   oracle_solve_info* info);

/* CG with the relative rule ||r||/||b|| < eps and checkpoint/restart files
 * x.dat r.dat d.dat z.dat in `dir` (src/mpla.cpp:887-1001).
 * ckpt_every <= 0 never checkpoints. */
void oracle_cg_relative(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                        int ckpt_every, int restart, const char* dir, oracle_solve_info* info);

/* BiCGSTAB with the true-residual rule (src/mpla.cpp:1003-1183); checkpoint
 * files x r r_hat z zz s p in `dir`. */
void oracle_bicgstab(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                     int ckpt_every, int restart, const char* dir, oracle_solve_info* info);

/* ---- vector file format (src/mpla.cpp:605-652, 686-713) ------------------ */
/* int32 N (native endian) followed by N float64. Returns 0 on success. */
int oracle_save_vector(const char* path, const double* x, int32_t n);
/* Returns 0 on success, -1 on open/read error, -2 on length mismatch. */
int oracle_load_vector(const char* path, double* x, int32_t n);

/* ---- Halton points (src/generate_quadrature_weights.cpp:107-117) -------- */
/* The reference draws its quadrature nodes from GSL's gsl_qrng_halton (GSL,
 * version unpinned by the reference; not vendored).  Restated from the
 * published definition that generator implements: point k (k = 1, 2, ...,
 * the first call returns k = 1) has coordinate d equal to the radical
 * inverse of k in the d-th prime base b, evaluated as
 *   r = 0, v = 1, w = 1/b;  while k > 0: v *= w, r += v * (k mod b), k /= b.
 * out[d * ld + i] is coordinate d of point first + i + 1.  dim <= 64.
 * Returns 0, or -1 for an unsupported dim. */
int oracle_halton(double* out, int dim, int64_t n, int64_t first, int64_t ld);

/* ---- H-matrix operator (src/hmglib_adapter.cu:6-132; hmglib restated) --- */
/* Cluster tree, block partition and ACA as described in oracle.c, for one
 * point set used as rows and columns (the kernel k's points); entries are
 * oracle_kernel_entry.  NULL for invalid arguments. */
typedef struct oracle_hmatrix oracle_hmatrix;
oracle_hmatrix* oracle_hm_create(const oracle_kernel* k, double eta, int c_leaf, int kmax, double eps);
void oracle_hm_free(oracle_hmatrix* h);
/* perm[k] = point at position k of the cluster-tree order. */
void oracle_hm_perm(const oracle_hmatrix* h, int64_t* perm);
/* counts: leaves, near blocks, far blocks, near entries, far entries,
 * stored terms (sum (m + n) rank), largest rank. */
void oracle_hm_counts(const oracle_hmatrix* h, int64_t* counts);
int oracle_hm_cluster(const oracle_hmatrix* h, int c, double* out);
int oracle_hm_block(const oracle_hmatrix* h, int far, int b, int* rc);
/* y = H x in the caller's order. */
void oracle_hm_matvec(const oracle_hmatrix* h, const double* x, double* y);
/* ACA of an explicit m x n matrix (column-major, lda): returns the rank;
 * U (m x kmax) and V (n x kmax) column-major receive the factors. */
int oracle_aca(const double* A, int64_t lda, int64_t m, int64_t n, int kmax, double eps, double* U, double* V);

/* Number of OpenMP threads the oracle's parallel loops use. */
int oracle_threads(void);

#ifdef __cplusplus
}
#endif
#endif
