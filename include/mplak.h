/*
 * mplak.h -- C ABI of the gfx950 kernel library (libmplak.so).
 *
 * This is the layer under the drop-in `mpla.h` API: plain pointers, 64-bit
 * sizes and a HIP stream handle, no C++ or torch types.  Every entry point
 *   - is stream-ordered (enqueues work on `stream`, never synchronises),
 *   - allocates nothing (workspace comes from the caller, sized by the
 *     matching *_workspace_bytes query),
 *   - returns an MPLAK_* status code instead of exiting.
 * All array arguments are device pointers unless stated otherwise.
 *
 * Each function names the reference operation it replaces (paths relative to
 * the MPLA source tree).
 */
#ifndef MPLAK_H
#define MPLAK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* mplak_stream_t; /* a hipStream_t; NULL is the default stream */

/* Status codes */
#define MPLAK_OK 0
#define MPLAK_ERR_ARG 1         /* invalid size, pointer or enum value */
#define MPLAK_ERR_HIP 2         /* a HIP runtime call or kernel launch failed */
#define MPLAK_ERR_WORKSPACE 3   /* caller workspace smaller than required */
#define MPLAK_ERR_UNSUPPORTED 4 /* valid request this build does not implement */
#define MPLAK_ERR_TOO_LARGE 5   /* beyond the fixed scratch of mplak_kernel_matvec_noalloc */

/* Kernel functions k(r) of the matrix-free operator, r = |p_i - p_j| */
#define MPLAK_KERNEL_GAUSSIAN 0 /* exp(-r^2 / (2 l^2)); l = 1/sqrt(2) is exp(-r^2) of src/kernel.cu:12-15 */
#define MPLAK_KERNEL_MATERN32 1 /* (1 + s) exp(-s), s = sqrt(3) r / l; l = sqrt(3) is src/kernel.cu:7-10 */

/* Arithmetic used for kernel-entry evaluation and the product */
#define MPLAK_PREC_F64 0    /* entries and accumulation in FP64 */
#define MPLAK_PREC_F32 1    /* entries in FP32, FP32 partial sums folded into FP64 */
#define MPLAK_PREC_BF16 2   /* multi-RHS only: BF16 operands on MFMA, FP32 accumulation */
#define MPLAK_PREC_BF16X3 3 /* multi-RHS only: split hi+lo BF16 operands, 3 MFMA products */

/* Stopping rules of the fused CG iteration */
#define MPLAK_STOP_RMS 0      /* sqrt(<r,r>/N) < eps      (src/mpla.cpp:843,868) */
#define MPLAK_STOP_RELATIVE 1 /* sqrt(<r,r>)/||b|| < eps  (src/mpla.cpp:920,963-965) */

/* Solver state flags (mplak_cg_state.done) */
#define MPLAK_RUNNING 0
#define MPLAK_CONVERGED 1
#define MPLAK_ITER_LIMIT 2
#define MPLAK_BREAKDOWN 3 /* mplak_cg1_iterate: step length not finite and positive; no update made */

const char* mplak_status_string(int status);

/* Library version as 10000*major + 100*minor + patch. */
int mplak_version(void);

/* ---------------------------------------------------------------------------
 * Synthetic inputs and plain vector fills
 * ------------------------------------------------------------------------ */

/* out[i] = lo + (hi - lo) * U(seed, stream, first_index + i), i in [0, n).
 * U is the SplitMix64 output number (first_index + i) of a generator keyed by
 * (seed, stream), mapped to [0,1) with 53 bits; the same integer recipe runs
 * on the host, so GPU and CPU inputs are bitwise identical.
 * Replaces curandGenerateUniformDouble (src/benchmark_dgemv.cpp:66-73,
 * src/benchmark_kernel_cg.cpp:75-103). */
int mplak_fill_uniform_f64(double* out, int64_t n, uint64_t seed, uint64_t stream,
                           uint64_t first_index, double lo, double hi, mplak_stream_t stream_);

/* out[i] = value.  Replaces cudaMemset in mpla_vector_set_zero
 * (src/mpla.cpp:812-815); a fill kernel is also safe inside HIP graphs. */
int mplak_fill_f64(double* out, int64_t n, double value, mplak_stream_t stream);

/* dst[i] = src[i] (device to device, kernel copy). */
int mplak_copy_f64(double* dst, const double* src, int64_t n, mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Dense FP64 level 1/2
 * ------------------------------------------------------------------------ */

/* Workspace for mplak_dgemv_f64 on an m x n block. */
size_t mplak_dgemv_workspace_bytes(int64_t m, int64_t n);

/* y = A x with A column-major (element (i,j) at A[i + j*lda]), beta = 0.
 * Replaces cublasDgemv(OP_N, m, n, 1, A, lda, x, 1, 0, y, 1) of
 * mpla_dgemv (src/mpla.cpp:729) and of the stored-dense plugin core
 * (src/test_cg.cpp:38-43, src/test_library.cpp:37-42).
 * The column sum is split over a fixed number of slabs that depends only on
 * (m, n); partial sums are combined in slab order, so results are
 * deterministic.  If `skip_if` is non-NULL and *skip_if != 0 when the kernels
 * run, y is left untouched (used by solvers that have already converged). */
int mplak_dgemv_f64(int64_t m, int64_t n, const double* A, int64_t lda, const double* x,
                    double* y, void* work, size_t work_bytes, const int32_t* skip_if,
                    mplak_stream_t stream);

size_t mplak_dot_workspace_bytes(int64_t n);

/* *result = sum_i x[i] y[i] (result is a device pointer), deterministic order.
 * Replaces cublasDdot of mpla_ddot (src/mpla.cpp:757). */
int mplak_ddot_f64(int64_t n, const double* x, const double* y, double* result, void* work,
                   size_t work_bytes, mplak_stream_t stream);

/* y += alpha x.  Replaces cublasDaxpy of mpla_daxpy (src/mpla.cpp:809). */
int mplak_daxpy_f64(int64_t n, double alpha, const double* x, double* y, mplak_stream_t stream);

/* w = a x + b y (w may alias x or y). */
int mplak_dwaxpby_f64(int64_t n, double a, const double* x, double b, const double* y, double* w,
                      mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Matrix-free kernel-matrix operator
 * ------------------------------------------------------------------------ */

typedef struct mplak_kernel_desc {
  const double* const* points; /* device array of `dim` device pointers: points[d][i] is
                                  coordinate d of global point i (the layout of
                                  kernel_matrix_data.points, src/kernel.h:3-8) */
  int32_t dim;                 /* 1 .. 16 */
  int32_t kind;                /* MPLAK_KERNEL_* */
  double length_scale;         /* > 0 */
  double nugget;               /* added to entries whose global row == global column */
  double coincident_shift;     /* added to entries whose point distance is < 1e-14
                                  (src/kernel.cu:45-48 uses 1e-8; 0 disables the test) */
  int64_t row_offset, row_count; /* global row range of the local block */
  int64_t col_offset, col_count; /* global column range of the local block */
} mplak_kernel_desc;

size_t mplak_kernel_matvec_workspace_bytes(const mplak_kernel_desc* desc);

/* y[i] = sum_j K(row_offset+i, col_offset+j) x[j] with entries computed on the
 * fly from the points and never stored.  `precision` is MPLAK_PREC_F64 or
 * MPLAK_PREC_F32.  Replaces the distance-matrix / apply_kernel / cublasDgemv
 * sequence of mpla_dgemv_core_kernel_cublas and
 * mpla_dgemv_core_kernel_streamed_cublas (src/kernel.cu:17-50, 87-171).
 * Deterministic; `skip_if` as for mplak_dgemv_f64. */
int mplak_kernel_matvec(const mplak_kernel_desc* desc, int precision, const double* x, double* y,
                        void* work, size_t work_bytes, const int32_t* skip_if,
                        mplak_stream_t stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  mplak_kernel_matvec with its workspace taken from a fixed
 * device array of mplak_kernel_matvec_noalloc_max_bytes() bytes (80 MiB)
 * that the library holds on each device it runs on: the same kernels, so y
 * is bitwise that of mplak_kernel_matvec.  Calls no hipMalloc or other
 * allocating routine.  Returns MPLAK_ERR_TOO_LARGE when the product's
 * workspace (mplak_kernel_matvec_workspace_bytes) exceeds the array: square
 * blocks up to n = 65536 fit.  All calls on one device share the array, so
 * calls that overlap on one device (two streams, two threads) race. */
int64_t mplak_kernel_matvec_noalloc_max_bytes(void);
int mplak_kernel_matvec_noalloc(const mplak_kernel_desc* desc, int precision, const double* x, double* y,
                                const int32_t* skip_if, mplak_stream_t stream);

/* No reference counterpart (an approximation the reference does not make):
 * the product of mplak_kernel_matvec with negligible entries dropped, for
 * inner solves whose result is corrected in FP64 (iterative refinement).
 * For a square block with row range == column range and dim <= 3, entries
 * between 128-point target blocks and 64-point source blocks whose bounding
 * boxes are so far apart that every entry is below 2^-drop_log2 (the kernel's
 * value at distance 0 is 1) are treated as zero; everything else is computed
 * by the symmetric product mplak_kernel_matvec uses for large square blocks
 * (at every size here).  Other blocks get mplak_kernel_matvec's product.
 * Pays off when consecutive points are spatially close (see
 * mplak_spatial_order).  0 < drop_log2 <= 2000.  Deterministic. */
size_t mplak_kernel_matvec_screened_workspace_bytes(const mplak_kernel_desc* desc);
int mplak_kernel_matvec_screened(const mplak_kernel_desc* desc, int precision, double drop_log2,
                                 const double* x, double* y, void* work, size_t work_bytes,
                                 const int32_t* skip_if, mplak_stream_t stream);

/* No reference counterpart.  A screening plan for repeated screened products
 * with the same points, precision and drop level (an inner solve): built once,
 * it lists the 512 x 512 tile pairs that hold kept entries, heaviest first,
 * and owns the product's partial sums (so one planned product at a time per
 * plan).  With a coincidence shift: mplak_kernel_matvec_planned gives
 * bitwise the result of mplak_kernel_matvec_screened with the same
 * arguments, without launching work for tile pairs whose entries are all
 * dropped.  FP64 with an FP64 plan and no coincidence shift: the screen is
 * applied per 16 x 16 block of points, with the entries from FP64
 * coordinate differences (Gaussian kernel on compact target tiles: FP64
 * squared distances from the matrix cores in tile-centred form), so it
 * drops only entries below 2^-drop_log2 and agrees with the screened
 * product to FP64 accuracy plus those entries (deterministic).  FP32 with an FP32 plan and no coincidence shift: on target tiles
 * whose bounding box is compact (half-diagonal <= 7 in the scaled
 * coordinates, the usual case after mplak_spatial_order) the squared
 * distances come from the matrix cores in tile-centred form and the screen
 * is applied per 16 x 16 block of points, so fewer entries are evaluated and
 * the result agrees with the screened product to FP32 accuracy (still
 * deterministic); other tiles are computed as by the screened product.  The
 * desc, precision and drop_log2 passed to the product must be the ones the
 * plan was built with.  Square blocks with row range == column range and
 * dim <= 3 only (mplak_screen_plan_bytes returns 0 otherwise). */
size_t mplak_screen_plan_bytes(const mplak_kernel_desc* desc);
int mplak_screen_plan_build(const mplak_kernel_desc* desc, int precision, double drop_log2, void* plan,
                            size_t plan_bytes, mplak_stream_t stream);
int mplak_kernel_matvec_planned(const mplak_kernel_desc* desc, int precision, double drop_log2, void* plan,
                                size_t plan_bytes, const double* x, double* y, const int32_t* skip_if,
                                mplak_stream_t stream);
/* No reference counterpart.  mplak_kernel_matvec_planned that also leaves
 * the partial sums of <x,x> and <x,y> for mplak_cg1_iterate: dots holds
 * 2 * mplak_planned_dots_count(n) doubles (all <x,x> partials, then all
 * <x,y> partials), each summed in a fixed order.  y is bitwise that of
 * mplak_kernel_matvec_planned. */
int64_t mplak_planned_dots_count(int64_t n);
int mplak_kernel_matvec_planned_dots(const mplak_kernel_desc* desc, int precision, double drop_log2, void* plan,
                                     size_t plan_bytes, const double* x, double* y, double* dots,
                                     const int32_t* skip_if, mplak_stream_t stream);
/* No reference counterpart.  Relaxed planned FP32 product for inexact inner
 * solves: nlev (1..3) FP32 plans of the same desc (no coincidence shift), at
 * drop levels drop_log2[0] > drop_log2[1] > ..., each of plan_bytes.  Level
 * l > 0 is used while *sel <= sel_below[l-1] (sel_below decreasing; sel is
 * device memory, read by the kernels, e.g. the inner solve's relative
 * residual in its mplak_cg_state), level 0 otherwise; nlev = 1 is
 * mplak_kernel_matvec_planned_dots on plans[0].  All levels use the partial
 * sums of plans[0].  dots as for mplak_kernel_matvec_planned_dots (may be
 * NULL). */
int mplak_kernel_matvec_planned_relaxed(const mplak_kernel_desc* desc, int nlev, const double* drop_log2,
                                        void* const* plans, size_t plan_bytes, const double* sel,
                                        const double* sel_below, const double* x, double* y, double* dots,
                                        const int32_t* skip_if, mplak_stream_t stream);
/* Writes counts[0] = tile pairs the plan lists, counts[1] = 128 x 64 blocks
 * they hold that pass the screen, counts[2] = those of counts[1] on diagonal
 * tile pairs, counts[3] and counts[4] = kernel entries the planned FP32
 * product (no coincidence shift) evaluates on off-diagonal and on diagonal
 * tile pairs (counts: 5 int64 in device memory). */
int mplak_screen_plan_counts(const mplak_kernel_desc* desc, const void* plan, size_t plan_bytes, int64_t* counts,
                             mplak_stream_t stream);

/* No reference counterpart.  Spatial (Hilbert-curve) permutation of the
 * points p[d][offset + i], i < n: perm[k] is the index i of the k-th point
 * when the points' bounding box is cut into 2^(30/dim) cells per axis and the
 * cells are visited along the Hilbert curve (Skilling's transpose form, axis
 * 0 the most significant within each level); points in one cell keep their
 * index order.  Deterministic.  1 <= dim <= 16, n < 2^31. */
size_t mplak_spatial_order_workspace_bytes(int64_t n, int32_t dim);
int mplak_spatial_order(const double* const* points, int32_t dim, int64_t offset, int64_t n, int64_t* perm,
                        void* work, size_t work_bytes, mplak_stream_t stream);

/* out[d * ld + k] = points[d][offset + perm[k]] (d < dim, k < n). */
int mplak_gather_points(const double* const* points, int32_t dim, int64_t offset, const int64_t* perm, int64_t n,
                        double* out, int64_t ld, mplak_stream_t stream);

/* dst[k] = src[perm[k]] (gather) and dst[perm[k]] = src[k] (scatter), k < n. */
int mplak_gather_f64(double* dst, const double* src, const int64_t* perm, int64_t n, mplak_stream_t stream);
int mplak_scatter_f64(double* dst, const double* src, const int64_t* perm, int64_t n, mplak_stream_t stream);

/* b[i] = prod_d (sqrt(pi)/2) (erf(1 - p_d) - erf(-p_d)) for global point
 * row_offset + i: the integral of exp(-|p - y|^2) over y in [0,1]^dim.
 * Replaces gen_gaussian_kernel_rhs (src/kernel.cu:52-68). */
int mplak_gaussian_kernel_rhs(const double* const* points, int32_t dim, int64_t row_offset,
                              int64_t row_count, double* b, mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Hierarchical-matrix (H-matrix) operator: the kernels behind the H-matrix
 * plugin cores of hmglib_adapter.h, which replace hmglib's setup and product
 * (src/hmglib_adapter.cu:6-132).  The host builds the cluster trees and the
 * block partition; these kernels compress the admissible blocks and apply
 * the operator.  Rows and columns are in cluster-tree order ("sorted").
 * ------------------------------------------------------------------------ */

#define MPLAK_HM_MAX_RANK 32

typedef struct mplak_hm_block {
  int64_t r0, m;    /* sorted rows [r0, r0 + m) */
  int64_t c0, n;    /* sorted columns [c0, c0 + n) */
  int64_t uoff;     /* U factor (m x kmax, column-major) at U + uoff */
  int64_t voff;     /* V factor (n x kmax, column-major) at V + voff */
} mplak_hm_block;

typedef struct mplak_hm_chunk {
  int32_t leaf;     /* row leaf whose near/far lists apply */
  int32_t count;    /* rows, 1 .. 128 */
  int64_t r0;       /* first sorted row */
} mplak_hm_chunk;

typedef struct mplak_hm_item {
  int32_t block;    /* admissible block */
  int32_t pad;
  int64_t j0, count; /* its local columns [j0, j0 + count) */
} mplak_hm_item;

/* A pair of near chunks A <= B (sorted rows [a0, a0 + am) and [b0, b0 + bm),
 * at most 512 each) of a symmetric operator's inadmissible blocks. */
typedef struct mplak_hm_pair {
  int64_t a0, am;
  int64_t b0, bm;
} mplak_hm_pair;

/* Closed-form entries: k(|p_i - p_j|) of MPLAK_KERNEL_* with the points'
 * coordinates already multiplied by the kernel's coordinate scale
 * (mplak_hm_gather_scaled).  The nugget is not part of these entries. */
typedef struct mplak_hm_points {
  const double* rows;  /* sorted row points: rows[d * row_ld + i] */
  const double* cols;  /* sorted column points */
  int64_t row_ld, col_ld;
  int32_t dim;         /* 1 .. 4 */
  int32_t kind;        /* MPLAK_KERNEL_* */
} mplak_hm_points;

/* out[d * ld + k] = scale(kind, length_scale) * points[d][offset + perm[k]]:
 * the sorted, scaled coordinates mplak_hm_points expects. */
int mplak_hm_gather_scaled(const double* const* points, int32_t dim, int64_t offset, const int64_t* perm, int64_t n,
                           int32_t kind, double length_scale, double* out, int64_t ld, mplak_stream_t stream);

/* Adaptive cross approximation with partial pivoting of every block
 * (one workgroup each): rank[b] <= kmax <= MPLAK_HM_MAX_RANK terms, fewer
 * when a pivot is exactly zero or, with eps > 0, once the newest term is
 * below eps times the Frobenius norm of the approximation.  Deterministic. */
int mplak_hm_aca(const mplak_hm_points* pts, const mplak_hm_block* blocks, int64_t nblocks, int32_t kmax,
                 double eps, double* U, double* V, int32_t* rank, mplak_stream_t stream);

/* part[i * kmax + l] = sum over the item's columns j of V_b(j, l) xs[c0_b + j]
 * (l < rank[b]), then t[b * kmax + l] = sum of the partials of block b's
 * items in item order (items of one block are consecutive; first_item has
 * nblocks + 1 entries). */
int mplak_hm_far_partials(const mplak_hm_block* blocks, const mplak_hm_item* items, int64_t nitems,
                          const int32_t* rank, int32_t kmax, const double* V, const double* xs, double* part,
                          mplak_stream_t stream);
int mplak_hm_far_sum(const int64_t* first_item, int64_t nblocks, const int32_t* rank, int32_t kmax,
                     const double* part, double* t, mplak_stream_t stream);

/* ys[row] = sum over the row leaf's inadmissible column ranges (near_cols
 * holds (c0, count) pairs, near_ptr the CSR offsets per leaf) of the
 * entries times xs, plus sum over its admissible blocks (far_blk, CSR
 * far_ptr) of U_b(row, :) t_b, for the rows of every chunk. */
int mplak_hm_apply(const mplak_hm_points* pts, const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* near_ptr,
                   const int64_t* near_cols, const int64_t* far_ptr, const int32_t* far_blk,
                   const mplak_hm_block* blocks, const int32_t* rank, int32_t kmax, const double* U, const double* t,
                   const double* xs, double* ys, mplak_stream_t stream);

/* Symmetric near field (rows and columns the same sorted point set, so
 * pts->rows and pts->cols hold the same coordinates): for pair k,
 *   part[(2 k) 512 + i]     = sum_{j < bm} K(a0 + i, b0 + j) xs[b0 + j]  (i < am),
 *   part[(2 k + 1) 512 + j] = sum_{i < am} K(a0 + i, b0 + j) xs[a0 + i]  (j < bm, a0 != b0),
 * one workgroup per pair; each entry of an off-diagonal pair is evaluated
 * once for both.  No reference counterpart (hmglib evaluates every
 * inadmissible block per row). */
int mplak_hm_near_pairs(const mplak_hm_points* pts, const mplak_hm_pair* pairs, int64_t npairs, const double* xs,
                        double* part, mplak_stream_t stream);

/* The apply of mplak_hm_apply with the near field taken from those
 * partials: chunk c lies in near chunk chunk_near[c], whose first row is
 * near_start[chunk_near[c]]; its near terms are the 512-row partial blocks
 * ent[k] (2 pair + side), k in [ent_ptr[nc], ent_ptr[nc + 1]), added in a
 * fixed order, then the far field as in mplak_hm_apply. */
int mplak_hm_apply_pairs(const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* chunk_near,
                         const int64_t* near_start, const int64_t* ent_ptr, const int64_t* ent, const double* part,
                         const int64_t* far_ptr, const int32_t* far_blk, const mplak_hm_block* blocks,
                         const int32_t* rank, int32_t kmax, const double* U, const double* t, double* ys,
                         mplak_stream_t stream);

/* The apply of mplak_hm_apply with the inadmissible blocks' entries read
 * from `stored` instead of evaluated: near range k (CSR order) of leaf l
 * holds entry (i, j) at stored[near_off[k] + (i - lo_l) + j m_l], with
 * leaf_range[2 l] = lo_l and leaf_range[2 l + 1] = m_l (the leaf's sorted
 * rows [lo_l, lo_l + m_l)).  For entries that are expensive to evaluate
 * (a virtual assembler): the near field is then HBM-bound. */
int mplak_hm_apply_stored(const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* leaf_range,
                          const int64_t* near_ptr, const int64_t* near_cols, const int64_t* near_off,
                          const double* stored, const int64_t* far_ptr, const int32_t* far_blk,
                          const mplak_hm_block* blocks, const int32_t* rank, int32_t kmax, const double* U,
                          const double* t, const double* xs, double* ys, mplak_stream_t stream);

/* y[perm[k]] = ys[k] + nugget * x[perm[k] + shift] (the second term only
 * where 0 <= perm[k] + shift < x_count): the sorted product back in the
 * caller's row order with the nugget on the global diagonal. */
int mplak_hm_scatter_nugget(const double* ys, const int64_t* perm, int64_t n, double nugget, const double* x,
                            int64_t shift, int64_t x_count, double* y, mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Multi-right-hand-side product (no reference counterpart)
 * ------------------------------------------------------------------------ */

size_t mplak_kernel_mrhs_workspace_bytes(const mplak_kernel_desc* desc, int precision, int64_t nrhs);

/* Y[:, r] = K V[:, r] for r < nrhs, V and Y column-major (leading dimensions
 * ldv >= col_count, ldy >= row_count).  precision:
 *   MPLAK_PREC_F64    FP64 entries and accumulation (VALU)
 *   MPLAK_PREC_F32    FP32 entries; MFMA v_mfma_f32_32x32x2_f32 when nrhs >= 16
 *                     and dim <= 4, VALU otherwise
 *   MPLAK_PREC_BF16   FP32 entries rounded to BF16, V in BF16, MFMA with FP32
 *                     accumulation (dim <= 4)
 *   MPLAK_PREC_BF16X3 hi+lo BF16 split of entries and V, three MFMA products
 * Each slab of sources accumulates in FP32 (FP64 for MPLAK_PREC_F64); slab
 * partials are added in FP64 in a fixed order; the nugget term is exact
 * FP64.  Deterministic; skip_if as for mplak_dgemv_f64. */
int mplak_kernel_mrhs(const mplak_kernel_desc* desc, int precision, int64_t nrhs, const double* V, int64_t ldv,
                      double* Y, int64_t ldy, void* work, size_t work_bytes, const int32_t* skip_if,
                      mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Fused conjugate-gradient steps with device-resident scalars
 *
 * One CG iteration of src/mpla.cpp:844-879 (and :937-975) is
 *     z = A d                         (caller's operator, skip_if = &in->done)
 *     mplak_cg_iterate(...)           (3 kernels: <d,z>; x,r update + <r,r>;
 *                                      stopping test + d update)
 * The state is double-buffered: iteration j reads state[j%2] and writes
 * state[(j+1)%2], so no launch both reads and writes a field.  Once `done`
 * is set, later iterations only carry the state forward, which makes a
 * captured chunk of iterations safe to replay past convergence.
 * ------------------------------------------------------------------------ */

typedef struct mplak_cg_state {
  double rr;       /* <r,r> of the current residual */
  double bnorm2;   /* <b,b> */
  double rel;      /* value of the stopping quantity after the last update */
  double eps;      /* tolerance */
  double n_global; /* global vector length (RMS rule) */
  double alpha;    /* last step length (diagnostic) */
  double beta;     /* last direction update coefficient (diagnostic) */
  int32_t k;       /* completed iterations */
  int32_t done;    /* MPLAK_RUNNING / MPLAK_CONVERGED / MPLAK_ITER_LIMIT */
  int32_t iter_max;
  int32_t rule;    /* MPLAK_STOP_* */
} mplak_cg_state;

size_t mplak_cg_workspace_bytes(int64_t n);

/* Start: r = b - z (z = A x0 computed by the caller), d = r, and state with
 * rr = <r,r>, bnorm2 = <b,b>, k = 0, done set when the initial residual
 * already meets the rule (src/mpla.cpp:829-843, 899-936).  `state` is a
 * device pointer to one mplak_cg_state. */
int mplak_cg_start(int64_t n, const double* b, const double* z, double* r, double* d,
                   mplak_cg_state* state, int rule, double eps, int32_t iter_max,
                   double n_global, void* work, size_t work_bytes, mplak_stream_t stream);

/* One iteration after z = A d: alpha = rr/<d,z>, x += alpha d, r -= alpha z,
 * rr' = <r,r>, stopping test, d = r + (rr'/rr) d.  Reads *in, writes *out. */
int mplak_cg_iterate(int64_t n, double* x, double* r, double* d, const double* z,
                     const mplak_cg_state* in, mplak_cg_state* out, void* work,
                     size_t work_bytes, mplak_stream_t stream);

/* No reference counterpart (used by the mixed-precision inner solves).  One
 * iteration of single-reduction CG (Chronopoulos & Gear 1989) after
 * w = A r with dots = the ndots-pair partials of <r,r> and <r,w> (as left by
 * mplak_kernel_matvec_planned_dots): beta = gamma/gamma_old, alpha =
 * gamma/(delta - beta gamma/alpha_old), p = r + beta p, s = w + beta s,
 * x += alpha p, r -= alpha s, in one kernel.  The stopping test is on
 * gamma = <r,r> of the residual entering the iteration (state.rr), so a
 * converged solve has applied the operator once more than it updated; k
 * counts updates, and k == 0 starts the recurrence (p and s need no initial
 * values).  A step length alpha that is not finite and positive (the
 * operator is not positive definite on the Krylov space, or a partial sum is
 * not finite) stops the solve with done = MPLAK_BREAKDOWN before any vector
 * is updated.  Same state double-buffering as mplak_cg_iterate. */
int mplak_cg1_iterate(int64_t n, double* x, double* r, double* p, double* s, const double* w, const double* dots,
                      int64_t ndots, const mplak_cg_state* in, mplak_cg_state* out, mplak_stream_t stream);

/* Recompute rr = <r,r> into *state (used after loading a checkpointed r). */
int mplak_cg_refresh_rr(int64_t n, const double* r, mplak_cg_state* state, void* work,
                        size_t work_bytes, mplak_stream_t stream);


/* Copy one state record (device to device); used to return a replayed chunk
 * with an odd number of iterations to buffer 0. */
int mplak_cg_copy_state(const mplak_cg_state* src, mplak_cg_state* dst, mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Batched CG over nrhs right-hand sides (no reference counterpart)
 *
 * Column c of the N x nrhs column-major blocks X, R, D, Z (leading dimension
 * n) runs the CG recurrence of src/mpla.cpp:937-975 with its own scalars in
 * states[c]; all columns share one multi-RHS product Z = A D per iteration
 * (skip_if = &group->all_done).  States are double-buffered per iteration as
 * for mplak_cg_iterate (in = states + p*nrhs, out = states + (1-p)*nrhs).
 * ------------------------------------------------------------------------ */

typedef struct mplak_bcg_group {
  int32_t all_done; /* 1 once no column is running */
  int32_t active;   /* columns still running */
  int32_t pad[2];
} mplak_bcg_group;

size_t mplak_bcg_workspace_bytes(int64_t n, int64_t nrhs);

/* R = B - Z, D = R, per-column states (rule, eps, iter_max), group flags. */
int mplak_bcg_start(int64_t n, int64_t nrhs, const double* B, const double* Z, double* R, double* D,
                    mplak_cg_state* states, mplak_bcg_group* group, int rule, double eps, int32_t iter_max,
                    double n_global, void* work, size_t work_bytes, mplak_stream_t stream);

/* One batched iteration after Z = A D; updates `group` from `out`. */
int mplak_bcg_iterate(int64_t n, int64_t nrhs, double* X, double* R, double* D, const double* Z,
                      const mplak_cg_state* in, mplak_cg_state* out, mplak_bcg_group* group, void* work,
                      size_t work_bytes, mplak_stream_t stream);

/* Copy nrhs state records and recompute the group flags from dst (group may be NULL). */
int mplak_bcg_copy_states(const mplak_cg_state* src, mplak_cg_state* dst, int64_t nrhs, mplak_bcg_group* group,
                          mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Ragged batches: many independent kernel systems in one launch (serving;
 * no reference counterpart)
 *
 * A batch is a device array of system records plus a device array of work
 * items; workgroup b of a launch processes items[b].  Each system's work split
 * (tiles, slabs, vector blocks) depends only on its own n, nrhs and
 * precision, so its arithmetic is identical alone and in any batch.
 * ------------------------------------------------------------------------ */

typedef struct mplak_ragged_system {
  const double* points;  /* SoA: coordinate d of point i at points[d * n + i] (dim <= 4) */
  const double* V;       /* product input, n x nrhs column-major (leading dimension n) */
  double* Y;             /* product output, n x nrhs column-major */
  double* part;          /* slab partials, mplak_ragged_plan() bytes */
  const int32_t* skip;   /* the system's product is skipped when non-NULL and *skip != 0 */
  int64_t n;             /* points (the system is n x n) */
  int64_t slab_cols;     /* planned */
  double length_scale;
  double nugget;         /* added on the diagonal */
  double scale;          /* planned: coordinate scale of the entry formula */
  int32_t dim;
  int32_t kind;          /* MPLAK_KERNEL_* */
  int32_t nrhs;
  int32_t precision;     /* MPLAK_PREC_F64 or MPLAK_PREC_F32 */
  int32_t rt, ctiles, tiles, slabs; /* planned */
  double drop_log2;      /* > 0: drop every block of entries that are all below 2^-drop_log2 (64 targets
                            x 64 sources in the slab scheme, 128 x 64 in the pair scheme), as
                            mplak_kernel_matvec_screened (useful on spatially ordered points; no
                            reference counterpart); 0: exact product */
  double drop_s;         /* planned */
  int32_t scheme;        /* planned: MPLAK_RAGGED_SLABS or MPLAK_RAGGED_PAIRS */
  int32_t reserved;
} mplak_ragged_system;

/* Work split of a ragged product.
 * SLABS: item = (target tile of 256 rows, source slab, column tile); every
 *   entry is evaluated once per column tile.
 * PAIRS: the points in tiles of 512 (`tiles` of them), item = (tile pair
 *   I <= J as pair index J (J + 1) / 2 + I, column tile); an off-diagonal
 *   block is evaluated once and used for the rows of I (sums over J) and the
 *   rows of J (sums over I).  Planned for n >= 8192; MPLAK_RAGGED_PAIRS=0 / 1
 *   selects SLABS / PAIRS for every n. */
#define MPLAK_RAGGED_SLABS 0
#define MPLAK_RAGGED_PAIRS 1

typedef struct mplak_ragged_item {
  int32_t sys;   /* index into the system array */
  int32_t tile;  /* product: target tile of 256 rows (SLABS) or tile pair (PAIRS); finish: block of 256
                    entries; CG: vector block */
  int32_t slab;  /* product: source slab (SLABS; 0 for PAIRS); CG: column */
  int32_t ctile; /* product: column tile of rt right-hand sides */
} mplak_ragged_item;

/* Host only: fills the planned fields of *s from (n, nrhs, precision, kind,
 * length_scale, drop_log2) and returns the bytes the system's `part` buffer
 * needs. */
size_t mplak_ragged_plan(mplak_ragged_system* s);

/* Host only: number of product / finish work items of a planned system; when
 * `out` is non-NULL also writes them (sys field = sys_index). */
int64_t mplak_ragged_product_items(const mplak_ragged_system* s, mplak_ragged_item* out, int32_t sys_index);
int64_t mplak_ragged_finish_items(const mplak_ragged_system* s, mplak_ragged_item* out, int32_t sys_index);

/* Product partials for `count` items whose systems share rt, precision and
 * scheme (systems and items are device arrays); then mplak_ragged_finish adds
 * the partials in a fixed order and the nugget term: Y = (K + nugget I) V per
 * system.  dim: 1..4 when every system of the launch has that dimension (a
 * kernel specialised for it; PAIRS items of a system of another dimension
 * are left undone), 0 for any mix of dimensions (SLABS only). */
int mplak_ragged_product(const mplak_ragged_system* systems, const mplak_ragged_item* items, int64_t count, int rt,
                         int dim, int precision, int scheme, mplak_stream_t stream);
int mplak_ragged_finish(const mplak_ragged_system* systems, const mplak_ragged_item* items, int64_t count,
                        mplak_stream_t stream);

typedef struct mplak_ragged_cg {
  double* X;             /* the iterate CG updates (the correction in refinement) */
  double* R;             /* residual, n x nrhs */
  double* D;             /* direction, n x nrhs */
  const double* Z;       /* operator applied to D (to the solution for the start step) */
  const double* B;       /* right-hand sides for the start step */
  double* Xout;          /* MPLAK_RAGGED_AXPY adds X into Xout */
  mplak_cg_state* states; /* 2 x nrhs records, double-buffered by iteration parity */
  mplak_bcg_group* group; /* all_done / active of the system */
  double* part;          /* mplak_ragged_cg_part_bytes(n, nrhs) */
  int64_t n;
  int32_t nrhs;
  int32_t G;             /* vector blocks per column: mplak_ragged_cg_blocks(n) */
  double eps;            /* tolerance written by MPLAK_RAGGED_START_FINAL */
} mplak_ragged_cg;

#define MPLAK_RAGGED_START 0       /* R = B - Z, D = R, partial <r,r>, <b,b>; items (sys, block, column) */
#define MPLAK_RAGGED_START_FINAL 1 /* states[column] of buffer 0 with rel = ||r||/||b||; items (sys, -, column) */
#define MPLAK_RAGGED_GROUP 2       /* group flags from buffer `parity`; items (sys, -, -) */
#define MPLAK_RAGGED_DZ 3          /* the three kernels of one CG iteration reading buffer `parity` */
#define MPLAK_RAGGED_UPDATE 4
#define MPLAK_RAGGED_DIRECT 5
#define MPLAK_RAGGED_AXPY 6        /* Xout += X */
#define MPLAK_RAGGED_FILL 7        /* X = value */

int32_t mplak_ragged_cg_blocks(int64_t n);
size_t mplak_ragged_cg_part_bytes(int64_t n, int32_t nrhs);
/* Host only: the (sys, vector block, column) items of one system. */
int64_t mplak_ragged_cg_items(const mplak_ragged_cg* s, mplak_ragged_item* out, int32_t sys_index);
int mplak_ragged_cg_op(int op, const mplak_ragged_cg* systems, const mplak_ragged_item* items, int64_t count,
                       int parity, double value, mplak_stream_t stream);

/* ---------------------------------------------------------------------------
 * Fused BiCGSTAB steps (src/mpla.cpp:1003-1183), device-resident scalars,
 * double-buffered state as for CG.  One iteration:
 *     z  = A p ;  mplak_bicgstab_step_s     alpha = rho/<z,r_hat>, s = r - alpha z
 *     zz = A s ;  mplak_bicgstab_step_x     omega = <zz,s>/<zz,zz>,
 *                                           x += alpha p + omega s, r = s - omega zz,
 *                                           rho' = <r,r_hat>, beta, p = r + beta (p - omega z)
 *     z  = A x ;  mplak_bicgstab_step_true  zz = b - z, stopping test on ||zz||/||b||
 * With true_residual = 0 the third product is skipped and step_x tests the
 * recursive residual ||r||/||b|| and writes the state itself.
 * ------------------------------------------------------------------------ */

typedef struct mplak_bicgstab_state {
  double rho;    /* <r, r_hat> of the current residual */
  double bnorm2; /* <b,b> */
  double rel;    /* stopping quantity after the last iteration */
  double eps;
  double alpha, omega, beta; /* last coefficients (diagnostic) */
  int32_t k;
  int32_t done;
  int32_t iter_max;
  int32_t true_residual; /* 1: stop on ||b - A x||/||b|| (reference), 0: on ||r||/||b|| */
} mplak_bicgstab_state;

size_t mplak_bicgstab_workspace_bytes(int64_t n);

/* r = b - z (z = A x0), r_hat = r, p = r, rho = <r,r>, bnorm2, stopping test. */
int mplak_bicgstab_start(int64_t n, const double* b, const double* z, double* r, double* r_hat,
                         double* p, mplak_bicgstab_state* state, double eps, int32_t iter_max,
                         int32_t true_residual, void* work, size_t work_bytes,
                         mplak_stream_t stream);

int mplak_bicgstab_step_s(int64_t n, const double* r, const double* z, double* s,
                          const double* r_hat, const mplak_bicgstab_state* in, void* work,
                          size_t work_bytes, mplak_stream_t stream);

int mplak_bicgstab_step_x(int64_t n, double* x, double* r, const double* r_hat, double* p,
                          const double* s, const double* z, const double* zz,
                          const mplak_bicgstab_state* in, mplak_bicgstab_state* out, void* work,
                          size_t work_bytes, mplak_stream_t stream);

int mplak_bicgstab_step_true(int64_t n, const double* b, const double* z, double* zz,
                             const mplak_bicgstab_state* in, mplak_bicgstab_state* out,
                             void* work, size_t work_bytes, mplak_stream_t stream);

/* rho = <r, r_hat> into *state (after loading checkpointed vectors). */
int mplak_bicgstab_refresh_rho(int64_t n, const double* r, const double* r_hat,
                               mplak_bicgstab_state* state, void* work, size_t work_bytes,
                               mplak_stream_t stream);

int mplak_bicgstab_copy_state(const mplak_bicgstab_state* src, mplak_bicgstab_state* dst,
                              mplak_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* MPLAK_H */
