/*
 * oracle.c -- CPU restatement of MPLA's solver hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library; the product never does.  Every function cites the
 * reference source it restates (paths relative to the MPLA source tree).
 *
 * How each part is pinned (checked against something other than itself,
 * see tests/test_oracle.py):
 *   rng            SplitMix64 published test vector (seed 1234567) and
 *                  distribution moments.
 *   block counts   hand-evaluated cases of src/mpla.cpp:157-209 and the
 *                  invariants sum = M, |count_p - count_q| <= 1, prefix offsets.
 *   dgemv / ddot   closed forms of src/test_library.cpp (A_ij = j*m+i,
 *                  x_j = j, y_j = j+1), exact in both modes; numpy matmul.
 *   kernel entry   closed forms: k(0) = 1 (+ nugget on i==j, + shift on
 *                  coincident points), Gaussian/Matern values at known r,
 *                  the reference's unscaled forms at l = 1/sqrt(2) and sqrt(3).
 *   kernel matvec  brute force with numpy/scipy on small N; symmetry
 *                  x^T K y = y^T K x; mode 0 vs mode 1 agreement.
 *   gaussian rhs   the 1-D integral identity checked by scipy quadrature.
 *   CG (both)      1-D Laplacian manufactured solution (src/test_cg.cpp:57-113),
 *                  alpha*I converges in one iteration, k distinct eigenvalues
 *                  converge in <= k iterations, scipy.sparse.linalg.cg.
 *   BiCGSTAB       manufactured solution on D + R (nonsymmetric), true
 *                  residual below eps at exit, numpy.linalg.solve.
 *   vector files   committed byte fixture and round trip; length mismatch.
 *   halton         exact rational radical inverses (Python fractions): base 2
 *                  bitwise, every base within 4 ulp; the van der Corput
 *                  digits of known indices; the first points (1/2, 1/3, 1/5).
 *                  GSL itself is absent, so agreement with gsl_qrng_halton's
 *                  last-bit rounding is "parity unpinned" beyond that.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea, Flood 2014).                                      */

static const uint64_t GOLDEN = 0x9E3779B97F4A7C15ull;

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t oracle_rng_key(uint64_t seed, uint64_t stream) { return mix64(seed * GOLDEN + stream); }

uint64_t oracle_rng_u64(uint64_t key, uint64_t i) { return mix64(key + (i + 1) * GOLDEN); }

void oracle_fill_uniform(double* out, int64_t n, uint64_t seed, uint64_t stream,
                         uint64_t first_index, double lo, double hi) {
  const uint64_t key = oracle_rng_key(seed, stream);
  const double span = hi - lo;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double u = (double)(oracle_rng_u64(key, first_index + (uint64_t)i) >> 11) * 0x1.0p-53;
    out[i] = fma(span, u, lo);
  }
}

void oracle_fill_normal(double* out, int64_t n, uint64_t seed, uint64_t stream,
                        uint64_t first_index) {
  const uint64_t key = oracle_rng_key(seed, stream);
  const double two_pi = 6.283185307179586476925;
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t k = 2 * (first_index + (uint64_t)i);
    const double u1 = (double)((oracle_rng_u64(key, k) >> 11) + 1) * 0x1.0p-53;
    const double u2 = (double)(oracle_rng_u64(key, k + 1) >> 11) * 0x1.0p-53;
    out[i] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
}

/* ------------------------------------------------------------------------ */
/* Balanced block distribution: src/mpla.cpp:157-209 (matrix rows/columns),
 * :338-367 (vectors over process rows), :390-439 (vectors over process
 * columns) all use count_p = floor(M/P) + (p < M mod P), offsets by prefix
 * sums, and exit when floor(M/P) == 0. */
int oracle_block_counts(int64_t M, int P, int64_t* counts, int64_t* offsets) {
  if (P <= 0) return -1;
  const int64_t base = M / P;
  const int64_t rem = M % P;
  if (base == 0) return -1;
  int64_t off = 0;
  for (int p = 0; p < P; ++p) {
    counts[p] = base + (p < rem ? 1 : 0);
    offsets[p] = off;
    off += counts[p];
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Dense level 1/2: cublasDgemv / cublasDdot / cublasDaxpy of
 * src/mpla.cpp:729, :757, :809. */

void oracle_dgemv(int64_t m, int64_t n, const double* A, int64_t lda, const double* x, double* y,
                  int mode) {
  if (mode == 0) {
    /* row blocks per thread, columns streamed in order */
#pragma omp parallel
    {
#ifdef _OPENMP
      const int nt = omp_get_num_threads(), t = omp_get_thread_num();
#else
      const int nt = 1, t = 0;
#endif
      const int64_t chunk = (m + nt - 1) / nt;
      const int64_t i0 = t * chunk;
      const int64_t i1 = (i0 + chunk < m) ? i0 + chunk : m;
      for (int64_t i = i0; i < i1; ++i) y[i] = 0.0;
      for (int64_t j = 0; j < n; ++j) {
        const double xj = x[j];
        const double* a = A + j * lda;
        for (int64_t i = i0; i < i1; ++i) y[i] += a[i] * xj;
      }
    }
  } else {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; ++i) {
      long double s = 0.0L;
      for (int64_t j = 0; j < n; ++j) s += (long double)A[i + j * lda] * (long double)x[j];
      y[i] = (double)s;
    }
  }
}

double oracle_ddot(int64_t n, const double* x, const double* y, int mode) {
  if (mode == 0) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += x[i] * y[i];
    return s;
  }
  long double s = 0.0L;
  for (int64_t i = 0; i < n; ++i) s += (long double)x[i] * (long double)y[i];
  return (double)s;
}

void oracle_daxpy(int64_t n, double alpha, const double* x, double* y) {
  for (int64_t i = 0; i < n; ++i) y[i] = y[i] + alpha * x[i];
}

/* ------------------------------------------------------------------------ */
/* Kernel entries.  src/kernel.cu:7-15 defines (1+r)exp(-r) and exp(-r*r) of
 * the point distance r; :39-50 adds 1e-8 when r < 1e-14.  The length scale
 * generalises them: Gaussian exp(-r^2/(2 l^2)) (l = 1/sqrt(2) gives exp(-r^2)),
 * Matern-3/2 (1+s)exp(-s) with s = sqrt(3) r / l (l = sqrt(3) gives the
 * reference's (1+r)exp(-r)).  The nugget on i == j follows the assembler's
 * regularization rule (src/kernel_system_assembler.h:33-34). */

double oracle_kernel_entry(const oracle_kernel* k, int64_t i, int64_t j) {
  double r2 = 0.0;
  for (int d = 0; d < k->dim; ++d) {
    const double df = k->points[(int64_t)d * k->n_points + i] - k->points[(int64_t)d * k->n_points + j];
    r2 += df * df;
  }
  const double r = sqrt(r2);
  const double l = k->length_scale;
  double v;
  if (k->kind == ORACLE_KERNEL_GAUSSIAN) {
    v = exp(-r2 / (2.0 * l * l));
  } else {
    const double s = sqrt(3.0) * r / l;
    v = (1.0 + s) * exp(-s);
  }
  if (r < 1.0e-14) v += k->coincident_shift;
  if (i == j) v += k->nugget;
  return v;
}

void oracle_kernel_matvec(const oracle_kernel* k, int64_t row_off, int64_t rows, int64_t col_off,
                          int64_t cols, const double* x, double* y, int mode) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < rows; ++i) {
    if (mode == 0) {
      double s = 0.0;
      for (int64_t j = 0; j < cols; ++j) s += oracle_kernel_entry(k, row_off + i, col_off + j) * x[j];
      y[i] = s;
    } else {
      long double s = 0.0L;
      for (int64_t j = 0; j < cols; ++j)
        s += (long double)oracle_kernel_entry(k, row_off + i, col_off + j) * (long double)x[j];
      y[i] = (double)s;
    }
  }
}

void oracle_gaussian_rhs(const oracle_kernel* k, int64_t row_off, int64_t rows, double* b) {
  const double half_sqrt_pi = 0.88622692545275801365;
  for (int64_t i = 0; i < rows; ++i) {
    double v = 1.0;
    for (int d = 0; d < k->dim; ++d) {
      const double p = k->points[(int64_t)d * k->n_points + row_off + i];
      v = v * half_sqrt_pi * (erf(1.0 - p) - erf(0.0 - p));
    }
    b[i] = v;
  }
}

/* ------------------------------------------------------------------------ */

void oracle_apply(const oracle_op* op, const double* x, double* y) {
  if (op->type == 0)
    oracle_dgemv(op->n, op->n, op->A, op->lda, x, y, op->mode);
  else
    oracle_kernel_matvec(&op->kernel, 0, op->n, 0, op->n, x, y, op->mode);
}

static double dot(int64_t n, const double* a, const double* b) { return oracle_ddot(n, a, b, 1); }

static void set_zero(int64_t n, double* v) { memset(v, 0, sizeof(double) * (size_t)n); }

static void ckpt_path(char* buf, size_t len, const char* dir, const char* name) {
  snprintf(buf, len, "%s/%s", (dir && dir[0]) ? dir : ".", name);
}

static void save_named(const char* dir, const char* name, const double* v, int64_t n) {
  char p[4096];
  ckpt_path(p, sizeof p, dir, name);
  oracle_save_vector(p, v, (int32_t)n);
}

static int load_named(const char* dir, const char* name, double* v, int64_t n) {
  char p[4096];
  ckpt_path(p, sizeof p, dir, name);
  return oracle_load_vector(p, v, (int32_t)n);
}

/* src/mpla.cpp:817-885 */
void oracle_cg_rms(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                   oracle_solve_info* info) {
  const int64_t n = op->n;
  double* r = calloc((size_t)n, sizeof(double));
  double* d = calloc((size_t)n, sizeof(double));
  double* z = calloc((size_t)n, sizeof(double));
  int matvecs = 0;
  oracle_apply(op, x, z);
  ++matvecs;
  set_zero(n, r);
  oracle_daxpy(n, 1.0, b, r);
  oracle_daxpy(n, -1.0, z, r);
  set_zero(n, d);
  oracle_daxpy(n, 1.0, r, d);
  double res = dot(n, r, r);
  int k = 0, converged = 0;
  double rel = sqrt(res / (double)n);
  if (rel >= eps) {
    for (k = 1; k < iter_max; ++k) {
      oracle_apply(op, d, z);
      ++matvecs;
      const double t1 = dot(n, r, r);
      double t2 = dot(n, d, z);
      const double alpha = t1 / t2;
      oracle_daxpy(n, alpha, d, x);
      oracle_daxpy(n, -alpha, z, r);
      t2 = dot(n, r, r);
      const double beta = t2 / t1;
      rel = sqrt(t2 / (double)n);
      if (rel < eps) {
        converged = 1;
        break;
      }
      set_zero(n, z);
      oracle_daxpy(n, beta, d, z);
      oracle_daxpy(n, 1.0, r, z);
      set_zero(n, d);
      oracle_daxpy(n, 1.0, z, d);
    }
  } else {
    converged = 1;
  }
  if (info) {
    info->iterations = k;
    info->converged = converged;
    info->final_rel = rel;
    info->matvecs = matvecs;
  }
  free(r);
  free(d);
  free(z);
}

/* src/mpla.cpp:887-1001 */
void oracle_cg_relative(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                        int ckpt_every, int restart, const char* dir, oracle_solve_info* info) {
  const int64_t n = op->n;
  double* r = calloc((size_t)n, sizeof(double));
  double* d = calloc((size_t)n, sizeof(double));
  double* z = calloc((size_t)n, sizeof(double));
  int matvecs = 0;
  const double b_norm = sqrt(dot(n, b, b));
  oracle_apply(op, x, z);
  ++matvecs;
  set_zero(n, r);
  oracle_daxpy(n, 1.0, b, r);
  oracle_daxpy(n, -1.0, z, r);
  set_zero(n, d);
  oracle_daxpy(n, 1.0, r, d);
  const double res = dot(n, r, r);
  double rel = sqrt(res) / b_norm;
  if (restart) {
    /* loaded after the initial residual is formed (src/mpla.cpp:920-933) */
    load_named(dir, "x.dat", x, n);
    load_named(dir, "r.dat", r, n);
    load_named(dir, "d.dat", d, n);
    load_named(dir, "z.dat", z, n);
  }
  int k = 0, converged = 0;
  if (rel >= eps) {
    for (k = 1; k < iter_max; ++k) {
      oracle_apply(op, d, z);
      ++matvecs;
      const double t1 = dot(n, r, r);
      double t2 = dot(n, d, z);
      const double alpha = t1 / t2;
      oracle_daxpy(n, alpha, d, x);
      oracle_daxpy(n, -alpha, z, r);
      t2 = dot(n, r, r);
      const double beta = t2 / t1;
      rel = sqrt(t2) / b_norm;
      if (rel < eps) {
        converged = 1;
        break;
      }
      set_zero(n, z);
      oracle_daxpy(n, beta, d, z);
      oracle_daxpy(n, 1.0, r, z);
      set_zero(n, d);
      oracle_daxpy(n, 1.0, z, d);
      if (ckpt_every > 0 && k % ckpt_every == 0) {
        save_named(dir, "x.dat", x, n);
        save_named(dir, "r.dat", r, n);
        save_named(dir, "d.dat", d, n);
        save_named(dir, "z.dat", z, n);
      }
    }
  } else {
    converged = 1;
  }
  if (info) {
    info->iterations = k;
    info->converged = converged;
    info->final_rel = rel;
    info->matvecs = matvecs;
  }
  free(r);
  free(d);
  free(z);
}

/* src/mpla.cpp:1003-1183 */
void oracle_bicgstab(const oracle_op* op, const double* b, double* x, int iter_max, double eps,
                     int ckpt_every, int restart, const char* dir, oracle_solve_info* info) {
  const int64_t n = op->n;
  double* r = calloc((size_t)n, sizeof(double));
  double* r_hat = calloc((size_t)n, sizeof(double));
  double* p = calloc((size_t)n, sizeof(double));
  double* z = calloc((size_t)n, sizeof(double));
  double* zz = calloc((size_t)n, sizeof(double));
  double* s = calloc((size_t)n, sizeof(double));
  int matvecs = 0;
  const double b_norm = sqrt(dot(n, b, b));
  oracle_apply(op, x, z);
  ++matvecs;
  set_zero(n, r);
  oracle_daxpy(n, 1.0, b, r);
  oracle_daxpy(n, -1.0, z, r);
  set_zero(n, r_hat);
  oracle_daxpy(n, 1.0, r, r_hat);
  set_zero(n, p);
  oracle_daxpy(n, 1.0, r, p);
  const double res = dot(n, r, r);
  double rel = sqrt(res) / b_norm;
  if (restart) {
    load_named(dir, "x.dat", x, n);
    load_named(dir, "r.dat", r, n);
    load_named(dir, "r_hat.dat", r_hat, n);
    load_named(dir, "z.dat", z, n);
    load_named(dir, "zz.dat", zz, n);
    load_named(dir, "s.dat", s, n);
    load_named(dir, "p.dat", p, n);
  }
  int k = 0, converged = 0;
  if (rel >= eps) {
    for (k = 1; k < iter_max; ++k) {
      double tmp1 = dot(n, r_hat, r);
      oracle_apply(op, p, z);
      ++matvecs;
      double tmp2 = dot(n, z, r_hat);
      const double alpha = tmp1 / tmp2;
      const double r_i_r_hat = tmp1;
      set_zero(n, s);
      oracle_daxpy(n, 1.0, r, s);
      oracle_daxpy(n, -alpha, z, s);
      oracle_apply(op, s, zz);
      ++matvecs;
      tmp1 = dot(n, zz, s);
      tmp2 = dot(n, zz, zz);
      const double omega = tmp1 / tmp2;
      oracle_daxpy(n, alpha, p, x);
      oracle_daxpy(n, omega, s, x);
      set_zero(n, r);
      oracle_daxpy(n, 1.0, s, r);
      oracle_daxpy(n, -omega, zz, r);
      tmp1 = dot(n, r, r_hat);
      const double beta = (alpha / omega) * (tmp1 / r_i_r_hat);
      set_zero(n, zz);
      oracle_daxpy(n, 1.0, p, zz);
      oracle_daxpy(n, -omega, z, zz);
      set_zero(n, p);
      oracle_daxpy(n, 1.0, r, p);
      oracle_daxpy(n, beta, zz, p);
      oracle_apply(op, x, z);
      ++matvecs;
      set_zero(n, zz);
      oracle_daxpy(n, 1.0, b, zz);
      oracle_daxpy(n, -1.0, z, zz);
      rel = sqrt(dot(n, zz, zz)) / b_norm;
      if (rel < eps) {
        converged = 1;
        break;
      }
      if (ckpt_every > 0 && k % ckpt_every == 0) {
        save_named(dir, "x.dat", x, n);
        save_named(dir, "r.dat", r, n);
        save_named(dir, "r_hat.dat", r_hat, n);
        save_named(dir, "z.dat", z, n);
        save_named(dir, "zz.dat", zz, n);
        save_named(dir, "s.dat", s, n);
        save_named(dir, "p.dat", p, n);
      }
    }
  } else {
    converged = 1;
  }
  if (info) {
    info->iterations = k;
    info->converged = converged;
    info->final_rel = rel;
    info->matvecs = matvecs;
  }
  free(r);
  free(r_hat);
  free(p);
  free(z);
  free(zz);
  free(s);
}

/* ------------------------------------------------------------------------ */
/* src/mpla.cpp:605-652 (save: int32 count then raw doubles, native endian)
 * and :686-713 (load: exit(1) on a length mismatch). */

int oracle_save_vector(const char* path, const double* x, int32_t n) {
  FILE* f = fopen(path, "wb");
  if (!f) return -1;
  int ok = fwrite(&n, sizeof(int32_t), 1, f) == 1 &&
           fwrite(x, sizeof(double), (size_t)n, f) == (size_t)n;
  fclose(f);
  return ok ? 0 : -1;
}

int oracle_load_vector(const char* path, double* x, int32_t n) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  int32_t m = 0;
  if (fread(&m, sizeof(int32_t), 1, f) != 1) {
    fclose(f);
    return -1;
  }
  if (m != n) {
    fclose(f);
    return -2;
  }
  size_t got = fread(x, sizeof(double), (size_t)n, f);
  fclose(f);
  return got == (size_t)n ? 0 : -1;
}

int oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---- Halton points ------------------------------------------------------ */

static const int oracle_primes[64] = {
    2,   3,   5,   7,   11,  13,  17,  19,  23,  29,  31,  37,  41,  43,  47,  53,
    59,  61,  67,  71,  73,  79,  83,  89,  97,  101, 103, 107, 109, 113, 127, 131,
    137, 139, 149, 151, 157, 163, 167, 173, 179, 181, 191, 193, 197, 199, 211, 223,
    227, 229, 233, 239, 241, 251, 257, 263, 269, 271, 277, 281, 283, 293, 307, 311};

int oracle_halton(double* out, int dim, int64_t n, int64_t first, int64_t ld) {
  if (dim < 1 || dim > 64) return -1;
  for (int d = 0; d < dim; ++d) {
    const int64_t b = oracle_primes[d];
    const double w = 1.0 / (double)b;
    for (int64_t i = 0; i < n; ++i) {
      int64_t k = first + i + 1;
      double r = 0.0, v = 1.0;
      while (k > 0) {
        v *= w;
        r += v * (double)(k % b);
        k /= b;
      }
      out[(int64_t)d * ld + i] = r;
    }
  }
  return 0;
}

/* ---- H-matrix operator (the H-matrix plugin, src/hmglib_adapter.cu:6-132) --
 * The reference delegates setup and product to hmglib (P. Zaspel's GPU
 * H-matrix library; version unpinned, linked by src/Makefile:63,78-81, not
 * vendored).  Restated from the published construction of H-matrices
 * (Hackbusch, "Hierarchical Matrices", Springer 2015, ch. 5-6; Bebendorf,
 * Numer. Math. 86, 2000 for ACA), with the choices the build states:
 *   cluster tree   a cluster of more than c_leaf points is split along the
 *                  longest edge of its bounding box (lowest axis on ties)
 *                  into the floor(n/2) points of smallest (coordinate, index)
 *                  and the rest; a leaf lists its points by increasing index.
 *   admissibility  min(diam B_s, diam B_t) <= eta dist(B_s, B_t), dist > 0.
 *   block tree     admissible -> low rank; both leaves -> dense; otherwise the
 *                  non-leaf side(s) are refined.
 *   ACA            partial pivoting from the first row of the block: pivot
 *                  column = arg max |row residual| (smallest index on ties),
 *                  next pivot row = arg max |column residual| over unused rows;
 *                  stops at kmax terms, a zero pivot, or (eps > 0) once
 *                  |u_l||v_l| <= eps |S_l|_F.
 * The product uses the oracle's kernel entries (libm exp) for the dense
 * blocks and the ACA factors; sums in long double. */

typedef struct {
  int64_t lo, hi;
  int child[2];
  double bmin[8], bmax[8];
} hm_cluster;

struct oracle_hmatrix {
  oracle_kernel k;
  int64_t n;
  int dim, kmax;
  int64_t* perm;
  hm_cluster* nodes;
  int nnodes, nleaves;
  int (*near)[2];
  int nnear, cap_near;
  int (*far)[2];
  int nfar, cap_far;
  double** U; /* per far block, m x rank column-major */
  double** V; /* per far block, n x rank */
  int* rank;
};

static const double* hm_key_coords;
static int hm_key_cmp(const void* pa, const void* pb) {
  const int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
  const double ca = hm_key_coords[a], cb = hm_key_coords[b];
  if (ca < cb) return -1;
  if (ca > cb) return 1;
  return a < b ? -1 : (a > b ? 1 : 0);
}
static int hm_index_cmp(const void* pa, const void* pb) {
  const int64_t a = *(const int64_t*)pa, b = *(const int64_t*)pb;
  return a < b ? -1 : (a > b ? 1 : 0);
}

static int hm_split(struct oracle_hmatrix* h, int id, int c_leaf) {
  hm_cluster* c = &h->nodes[id];
  const int64_t lo = c->lo, hi = c->hi, n = h->n;
  for (int d = 0; d < h->dim; ++d) {
    double mn = INFINITY, mx = -INFINITY;
    for (int64_t q = lo; q < hi; ++q) {
      const double v = h->k.points[(int64_t)d * n + h->perm[q]];
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    c->bmin[d] = mn;
    c->bmax[d] = mx;
  }
  c->child[0] = c->child[1] = -1;
  if (hi - lo <= c_leaf) {
    qsort(h->perm + lo, (size_t)(hi - lo), sizeof(int64_t), hm_index_cmp);
    return 0;
  }
  int axis = 0;
  for (int d = 1; d < h->dim; ++d)
    if (c->bmax[d] - c->bmin[d] > c->bmax[axis] - c->bmin[axis]) axis = d;
  hm_key_coords = h->k.points + (int64_t)axis * n;
  /* a full sort selects the same two sets as any selection of the median */
  qsort(h->perm + lo, (size_t)(hi - lo), sizeof(int64_t), hm_key_cmp);
  const int64_t mid = lo + (hi - lo) / 2;
  const int c0 = h->nnodes;
  h->nodes[c0].lo = lo;
  h->nodes[c0].hi = mid;
  h->nodes[c0 + 1].lo = mid;
  h->nodes[c0 + 1].hi = hi;
  h->nnodes += 2;
  h->nodes[id].child[0] = c0;
  h->nodes[id].child[1] = c0 + 1;
  hm_split(h, c0, c_leaf);
  hm_split(h, c0 + 1, c_leaf);
  return 0;
}

static double hm_diam(const hm_cluster* c, int dim) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) s += (c->bmax[d] - c->bmin[d]) * (c->bmax[d] - c->bmin[d]);
  return sqrt(s);
}

static double hm_dist(const hm_cluster* a, const hm_cluster* b, int dim) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    double g = a->bmin[d] - b->bmax[d];
    if (b->bmin[d] - a->bmax[d] > g) g = b->bmin[d] - a->bmax[d];
    if (g > 0.0) s += g * g;
  }
  return sqrt(s);
}

static void hm_push(int (**list)[2], int* cnt, int* cap, int a, int b) {
  if (*cnt == *cap) {
    *cap = *cap ? 2 * *cap : 256;
    *list = realloc(*list, sizeof(int[2]) * (size_t)*cap);
  }
  (*list)[*cnt][0] = a;
  (*list)[*cnt][1] = b;
  ++*cnt;
}

static void hm_blocks(struct oracle_hmatrix* h, int a, int b, double eta) {
  const hm_cluster* A = &h->nodes[a];
  const hm_cluster* B = &h->nodes[b];
  const double ds = hm_dist(A, B, h->dim);
  const double da = hm_diam(A, h->dim), db = hm_diam(B, h->dim);
  if (ds > 0.0 && (da < db ? da : db) <= eta * ds) {
    hm_push(&h->far, &h->nfar, &h->cap_far, a, b);
    return;
  }
  const int la = A->child[0] < 0, lb = B->child[0] < 0;
  if (la && lb) {
    hm_push(&h->near, &h->nnear, &h->cap_near, a, b);
  } else if (la) {
    hm_blocks(h, a, B->child[0], eta);
    hm_blocks(h, a, B->child[1], eta);
  } else if (lb) {
    hm_blocks(h, A->child[0], b, eta);
    hm_blocks(h, A->child[1], b, eta);
  } else {
    const int a0 = A->child[0], a1 = A->child[1], b0 = B->child[0], b1 = B->child[1];
    hm_blocks(h, a0, b0, eta);
    hm_blocks(h, a0, b1, eta);
    hm_blocks(h, a1, b0, eta);
    hm_blocks(h, a1, b1, eta);
  }
}

/* entry (p, q) of the block with sorted rows r0 + p and columns c0 + q */
typedef struct {
  const struct oracle_hmatrix* h;
  const double* dense; /* explicit matrix (oracle_aca) instead of the kernel */
  int64_t ld, r0, c0;
} hm_src;

static double hm_entry(const hm_src* s, int64_t p, int64_t q) {
  if (s->dense) return s->dense[p + q * s->ld];
  return oracle_kernel_entry(&s->h->k, s->h->perm[s->r0 + p], s->h->perm[s->c0 + q]);
}

static int hm_aca(const hm_src* src, int64_t m, int64_t n, int kmax, double eps, double* U, double* V) {
  int64_t* used = malloc(sizeof(int64_t) * (size_t)(kmax > 0 ? kmax : 1));
  int64_t piv = 0;
  int r = 0;
  double s2 = 0.0;
  for (int l = 0; l < kmax; ++l) {
    int64_t js = -1;
    double best = -1.0;
    for (int64_t j = 0; j < n; ++j) {
      double v = hm_entry(src, piv, j);
      for (int q = 0; q < l; ++q) v -= U[piv + q * m] * V[j + q * n];
      V[j + l * n] = v;
      if (fabs(v) > best) {
        best = fabs(v);
        js = j;
      }
    }
    const double delta = V[js + l * n];
    if (!(fabs(delta) > 0.0) || !isfinite(delta)) break;
    double nv2 = 0.0, nu2 = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      V[j + l * n] /= delta;
      nv2 += V[j + l * n] * V[j + l * n];
    }
    int64_t is = -1;
    best = -1.0;
    for (int64_t i = 0; i < m; ++i) {
      double u = hm_entry(src, i, js);
      for (int q = 0; q < l; ++q) u -= V[js + q * n] * U[i + q * m];
      U[i + l * m] = u;
      nu2 += u * u;
      int fresh = i != piv;
      for (int q = 0; q < l; ++q) fresh = fresh && used[q] != i;
      if (fresh && fabs(u) > best) {
        best = fabs(u);
        is = i;
      }
    }
    r = l + 1;
    if (eps > 0.0) {
      double cross = 0.0;
      for (int q = 0; q < l; ++q) {
        double du = 0.0, dv = 0.0;
        for (int64_t i = 0; i < m; ++i) du += U[i + q * m] * U[i + l * m];
        for (int64_t j = 0; j < n; ++j) dv += V[j + q * n] * V[j + l * n];
        cross += du * dv;
      }
      s2 += 2.0 * cross + nu2 * nv2;
      if (sqrt(nu2 * nv2) <= eps * sqrt(fabs(s2))) break;
    }
    used[l] = piv;
    if (is < 0) break;
    piv = is;
  }
  free(used);
  return r;
}

int oracle_aca(const double* A, int64_t lda, int64_t m, int64_t n, int kmax, double eps, double* U, double* V) {
  hm_src s = {NULL, A, lda, 0, 0};
  return hm_aca(&s, m, n, kmax, eps, U, V);
}

struct oracle_hmatrix* oracle_hm_create(const oracle_kernel* k, double eta, int c_leaf, int kmax, double eps) {
  if (k->dim < 1 || k->dim > 8 || c_leaf < 1 || kmax < 1 || !(eta > 0.0)) return NULL;
  struct oracle_hmatrix* h = calloc(1, sizeof(*h));
  h->k = *k;
  h->n = k->n_points;
  h->dim = k->dim;
  h->kmax = kmax;
  h->perm = malloc(sizeof(int64_t) * (size_t)(h->n > 0 ? h->n : 1));
  for (int64_t i = 0; i < h->n; ++i) h->perm[i] = i;
  h->nodes = calloc((size_t)(2 * (h->n / (c_leaf > 1 ? c_leaf / 2 : 1) + 2) + 4), sizeof(hm_cluster));
  if (h->n > 0) {
    h->nodes[0].lo = 0;
    h->nodes[0].hi = h->n;
    h->nnodes = 1;
    hm_split(h, 0, c_leaf);
    for (int i = 0; i < h->nnodes; ++i) h->nleaves += h->nodes[i].child[0] < 0;
    hm_blocks(h, 0, 0, eta);
  }
  h->U = calloc((size_t)(h->nfar > 0 ? h->nfar : 1), sizeof(double*));
  h->V = calloc((size_t)(h->nfar > 0 ? h->nfar : 1), sizeof(double*));
  h->rank = calloc((size_t)(h->nfar > 0 ? h->nfar : 1), sizeof(int));
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < h->nfar; ++b) {
    const hm_cluster* A = &h->nodes[h->far[b][0]];
    const hm_cluster* B = &h->nodes[h->far[b][1]];
    const int64_t m = A->hi - A->lo, n = B->hi - B->lo;
    h->U[b] = malloc(sizeof(double) * (size_t)(m * kmax));
    h->V[b] = malloc(sizeof(double) * (size_t)(n * kmax));
    hm_src s = {h, NULL, 0, A->lo, B->lo};
    h->rank[b] = hm_aca(&s, m, n, kmax, eps, h->U[b], h->V[b]);
  }
  return h;
}

void oracle_hm_free(struct oracle_hmatrix* h) {
  if (!h) return;
  for (int b = 0; b < h->nfar; ++b) {
    free(h->U[b]);
    free(h->V[b]);
  }
  free(h->U);
  free(h->V);
  free(h->rank);
  free(h->near);
  free(h->far);
  free(h->nodes);
  free(h->perm);
  free(h);
}

void oracle_hm_perm(const struct oracle_hmatrix* h, int64_t* perm) {
  memcpy(perm, h->perm, sizeof(int64_t) * (size_t)h->n);
}

void oracle_hm_counts(const struct oracle_hmatrix* h, int64_t* counts) {
  int64_t near_entries = 0, far_entries = 0, stored = 0;
  int max_rank = 0;
  for (int b = 0; b < h->nnear; ++b) {
    const hm_cluster* A = &h->nodes[h->near[b][0]];
    const hm_cluster* B = &h->nodes[h->near[b][1]];
    near_entries += (A->hi - A->lo) * (B->hi - B->lo);
  }
  for (int b = 0; b < h->nfar; ++b) {
    const hm_cluster* A = &h->nodes[h->far[b][0]];
    const hm_cluster* B = &h->nodes[h->far[b][1]];
    far_entries += (A->hi - A->lo) * (B->hi - B->lo);
    stored += (A->hi - A->lo + B->hi - B->lo) * h->rank[b];
    if (h->rank[b] > max_rank) max_rank = h->rank[b];
  }
  counts[0] = h->nleaves;
  counts[1] = h->nnear;
  counts[2] = h->nfar;
  counts[3] = near_entries;
  counts[4] = far_entries;
  counts[5] = stored;
  counts[6] = max_rank;
}

/* Cluster c of the tree: lo, hi, child0, child1, bmin[0..dim), bmax[0..dim)
 * into out (4 + 2 dim doubles); returns the node count. */
int oracle_hm_cluster(const struct oracle_hmatrix* h, int c, double* out) {
  if (c >= 0 && c < h->nnodes) {
    const hm_cluster* x = &h->nodes[c];
    out[0] = (double)x->lo;
    out[1] = (double)x->hi;
    out[2] = x->child[0];
    out[3] = x->child[1];
    for (int d = 0; d < h->dim; ++d) {
      out[4 + d] = x->bmin[d];
      out[4 + h->dim + d] = x->bmax[d];
    }
  }
  return h->nnodes;
}

/* Block b of the near (far = 0) or far (far = 1) list: its row and column
 * clusters into rc[0..1]; returns the list length. */
int oracle_hm_block(const struct oracle_hmatrix* h, int far, int b, int* rc) {
  const int cnt = far ? h->nfar : h->nnear;
  if (b >= 0 && b < cnt) {
    rc[0] = far ? h->far[b][0] : h->near[b][0];
    rc[1] = far ? h->far[b][1] : h->near[b][1];
  }
  return cnt;
}

/* y = H x in the caller's order (x, y indexed by point). */
void oracle_hm_matvec(const struct oracle_hmatrix* h, const double* x, double* y) {
  const int64_t n = h->n;
  long double* acc = calloc((size_t)(n > 0 ? n : 1), sizeof(long double));
  for (int b = 0; b < h->nnear; ++b) {
    const hm_cluster* A = &h->nodes[h->near[b][0]];
    const hm_cluster* B = &h->nodes[h->near[b][1]];
#pragma omp parallel for schedule(static)
    for (int64_t p = A->lo; p < A->hi; ++p) {
      long double s = 0.0L;
      for (int64_t q = B->lo; q < B->hi; ++q)
        s += (long double)oracle_kernel_entry(&h->k, h->perm[p], h->perm[q]) * x[h->perm[q]];
      acc[p] += s;
    }
  }
  for (int b = 0; b < h->nfar; ++b) {
    const hm_cluster* A = &h->nodes[h->far[b][0]];
    const hm_cluster* B = &h->nodes[h->far[b][1]];
    const int64_t m = A->hi - A->lo, nn = B->hi - B->lo;
    for (int l = 0; l < h->rank[b]; ++l) {
      long double t = 0.0L;
      for (int64_t q = 0; q < nn; ++q) t += (long double)h->V[b][q + l * nn] * x[h->perm[B->lo + q]];
      for (int64_t p = 0; p < m; ++p) acc[A->lo + p] += (long double)h->U[b][p + l * m] * t;
    }
  }
  for (int64_t p = 0; p < n; ++p) y[h->perm[p]] = (double)acc[p];
  free(acc);
}
