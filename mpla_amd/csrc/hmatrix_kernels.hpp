// Device kernels of the hierarchical-matrix (H-matrix) operator, templated on
// the source of the matrix entries: closed-form kernel entries from sorted,
// pre-scaled point coordinates (libmplak), or a device-side virtual
// system_assembler (libmpla, relocatable device code).
//
// The operator replaces hmglib's H-matrix product behind
// mpla_dgemv_core_hmglib_h_matrix (src/hmglib_adapter.cu:117-124): rows and
// columns are permuted into cluster-tree order, admissible blocks are stored
// as U V^T from adaptive cross approximation with partial pivoting
// (M. Bebendorf, "Approximation of boundary element matrices", Numer. Math.
// 86, 2000), inadmissible leaf blocks are evaluated on the fly.  Every sum
// below has a fixed order that depends only on the block structure.
#pragma once

#include "common.hpp"
#include "entry.hpp"

namespace mplak {
namespace hm {

constexpr int kHmThreads = 256;                 // 4 waves
constexpr int kHmApplyThreads = 512;             // 8 waves per apply work item
constexpr int kHmLaneRows = 2;                   // rows per lane
constexpr int kHmRows = kHmLaneRows * kWave;     // rows of one apply work item
constexpr int kHmParts = kHmApplyThreads / kWave;  // waves splitting one row's sources
constexpr int kHmStage = 512;                   // sources staged in LDS at a time
constexpr int kHmMaxDim = 4;

// Signed value of largest magnitude and its index; ties go to the smaller
// index, so the choice does not depend on the reduction order.
struct ArgMax {
  double v;
  int64_t i;
};

__device__ __forceinline__ ArgMax better(ArgMax a, ArgMax b) {
  const double fa = fabs(a.v), fb = fabs(b.v);
  if (b.i < 0) return a;
  if (a.i < 0) return b;
  if (fb > fa || (fb == fa && b.i < a.i)) return b;
  return a;
}

// Block-wide arg-max of |v| (every thread gets the result).  `sv`/`si` hold
// kHmThreads / kWave entries.
__device__ __forceinline__ ArgMax block_argmax(ArgMax a, double* sv, int64_t* si) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    ArgMax o;
    o.v = __shfl_xor(a.v, off, kWave);
    o.i = __shfl_xor(a.i, off, kWave);
    a = better(a, o);
  }
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if ((threadIdx.x & (kWave - 1)) == 0) {
    sv[w] = a.v;
    si[w] = a.i;
  }
  __syncthreads();
  ArgMax r{sv[0], si[0]};
#pragma unroll
  for (int k = 1; k < kHmThreads / kWave; ++k) r = better(r, ArgMax{sv[k], si[k]});
  return r;
}

// Closed-form kernel entries on sorted points whose coordinates are already
// multiplied by coord_scale(kind, length scale): entry = k(s) with s the
// squared distance.  The nugget is not part of these entries (it is added on
// the caller's diagonal after the product).
template <int KIND>
struct KernelEntries {
  const double* rows;  // rows[d * row_ld + i]
  const double* cols;  // cols[d * col_ld + j]
  int64_t row_ld, col_ld;
  int dim;
  static constexpr bool kTable = true;
  __device__ __forceinline__ double operator()(int64_t i, int64_t j, const double* tab) const {
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
      const double df = rows[d * row_ld + i] - cols[d * col_ld + j];
      s = fma(df, df, s);
    }
    return kernel_entry<double, KIND>(s, tab);
  }
};

// ---------------------------------------------------------------------------
// Adaptive cross approximation, one workgroup per admissible block.
//
// Block b covers sorted rows [r0, r0 + m) and columns [c0, c0 + n); its
// factors are U (m x kmax) and V (n x kmax), column-major.  Step l:
//   row residual    R_j = A(i, j) - sum_{q<l} U(i,q) V(j,q)   (pivot row i)
//   column pivot    j* = arg max |R_j|, delta = R_j*; stop when delta = 0
//   V(:, l) = R / delta
//   column residual U(i', l) = A(i', j*) - sum_{q<l} V(j*,q) U(i',q)
//   next pivot row  arg max |U(i', l)| over rows not yet used.
// With eps > 0 the block also stops once |u_l| |v_l| <= eps |S_l|_F, with
// |S_l|_F^2 updated by the cross terms 2 sum_q (u_q.u_l)(v_q.v_l) +
// |u_l|^2 |v_l|^2 (Bebendorf's stopping rule); eps = 0 runs to kmax.
// Each thread owns rows ii = tid (mod 256) of U and columns j = tid (mod
// 256) of V, so it reads back only what it wrote; the few values other
// threads need (the pivot's row of U and column of V) go through LDS.
template <class E>
__device__ void aca_block(const E& e, const mplak_hm_block& B, int kmax, double eps, double* __restrict__ U,
                          double* __restrict__ V, int32_t* __restrict__ rank_out, const double* tab) {
  __shared__ double urow[MPLAK_HM_MAX_RANK];
  __shared__ double vcol[MPLAK_HM_MAX_RANK];
  __shared__ int64_t used[MPLAK_HM_MAX_RANK];
  __shared__ double sv[kHmThreads / kWave];
  __shared__ int64_t si[kHmThreads / kWave];
  __shared__ double red[kHmThreads / kWave];
  const int tid = threadIdx.x;
  const int64_t m = B.m, n = B.n;
  double* Ub = U + B.uoff;
  double* Vb = V + B.voff;
  int64_t piv = 0;
  int r = 0;
  double s2 = 0.0;
  for (int l = 0; l < kmax; ++l) {
    // U(piv, q) for q < l, from the thread that owns row piv
    if (tid == (int)(piv % kHmThreads))
      for (int q = 0; q < l; ++q) urow[q] = Ub[piv + q * m];
    __syncthreads();
    ArgMax best{0.0, -1};
    for (int64_t j = tid; j < n; j += kHmThreads) {
      double v = e(B.r0 + piv, B.c0 + j, tab);
      for (int q = 0; q < l; ++q) v = fma(-urow[q], Vb[j + q * n], v);
      Vb[j + l * n] = v;
      best = better(best, ArgMax{v, j});
    }
    const ArgMax jb = block_argmax(best, sv, si);
    const double delta = jb.v;
    if (jb.i < 0 || !(fabs(delta) > 0.0) || !isfinite(delta)) break;
    const double inv = 1.0 / delta;
    double nv2 = 0.0;
    for (int64_t j = tid; j < n; j += kHmThreads) {
      const double v = Vb[j + l * n] * inv;
      Vb[j + l * n] = v;
      nv2 = fma(v, v, nv2);
    }
    // V(j*, q) for q <= l, from the thread that owns column j*
    const int64_t js = jb.i;
    if (tid == (int)(js % kHmThreads))
      for (int q = 0; q <= l; ++q) vcol[q] = Vb[js + q * n];
    __syncthreads();
    ArgMax ib{0.0, -1};
    double nu2 = 0.0;
    for (int64_t ii = tid; ii < m; ii += kHmThreads) {
      double u = e(B.r0 + ii, B.c0 + js, tab);
      for (int q = 0; q < l; ++q) u = fma(-vcol[q], Ub[ii + q * m], u);
      Ub[ii + l * m] = u;
      nu2 = fma(u, u, nu2);
      bool fresh = ii != piv;
      for (int q = 0; q < l; ++q) fresh = fresh && used[q] != ii;
      if (fresh) ib = better(ib, ArgMax{u, ii});
    }
    r = l + 1;
    if (eps > 0.0) {
      nu2 = block_sum<kHmThreads>(nu2, red);
      nv2 = block_sum<kHmThreads>(nv2, red);
      double cross = 0.0;
      for (int q = 0; q < l; ++q) {
        double du = 0.0, dv = 0.0;
        for (int64_t ii = tid; ii < m; ii += kHmThreads) du = fma(Ub[ii + q * m], Ub[ii + l * m], du);
        for (int64_t j = tid; j < n; j += kHmThreads) dv = fma(Vb[j + q * n], Vb[j + l * n], dv);
        du = block_sum<kHmThreads>(du, red);
        dv = block_sum<kHmThreads>(dv, red);
        cross = fma(du, dv, cross);
      }
      s2 += 2.0 * cross + nu2 * nv2;
      if (sqrt(nu2 * nv2) <= eps * sqrt(fabs(s2))) break;
    }
    const ArgMax nb = block_argmax(ib, sv, si);
    if (tid == 0) used[l] = piv;
    if (nb.i < 0) break;  // every row is a pivot row: the block is exact
    piv = nb.i;
  }
  if (tid == 0) rank_out[0] = r;
}

template <class E>
__global__ __launch_bounds__(kHmThreads) void aca_kernel(E e, const mplak_hm_block* __restrict__ blocks, int kmax,
                                                         double eps, double* __restrict__ U, double* __restrict__ V,
                                                         int32_t* __restrict__ rank) {
  __shared__ double tab[kExp2Table];
  if (E::kTable) {
    fill_exp2_table(tab);
    __syncthreads();
  }
  const mplak_hm_block B = blocks[blockIdx.x];
  aca_block(e, B, kmax, eps, U, V, rank + blockIdx.x, tab);
}

// Far field for the rows of an apply item: sum over the leaf's admissible
// blocks b (those of part `part` of kHmParts) of U_b(row - r0_b, :) t_b.
__device__ __forceinline__ void far_rows(const int64_t (&row)[kHmLaneRows], const bool (&valid)[kHmLaneRows],
                                         const int64_t* __restrict__ far_ptr, const int32_t* __restrict__ far_blk,
                                         int32_t leaf, const mplak_hm_block* __restrict__ blocks,
                                         const int32_t* __restrict__ rank, int kmax, const double* __restrict__ U,
                                         const double* __restrict__ t, int part, double (&acc)[kHmLaneRows]) {
  const int64_t f0 = far_ptr[leaf], f1 = far_ptr[leaf + 1];
  for (int64_t f = f0 + part; f < f1; f += kHmParts) {
    const int32_t b = far_blk[f];
    const mplak_hm_block B = blocks[b];
    const int rb = rank[b];
    const double* tb = t + (int64_t)b * kmax;
#pragma unroll
    for (int h = 0; h < kHmLaneRows; ++h) {
      if (!valid[h]) continue;
      const double* u = U + B.uoff + (row[h] - B.r0);
      double a = acc[h];
      for (int q = 0; q < rb; ++q) a = fma(u[q * B.m], tb[q], a);
      acc[h] = a;
    }
  }
}

// Adds the kHmParts partial sums of each row in part order and stores them.
__device__ __forceinline__ void finish_rows(const double (&acc)[kHmLaneRows], const int64_t (&row)[kHmLaneRows],
                                            const bool (&valid)[kHmLaneRows], double* __restrict__ ys,
                                            double (*red)[kHmRows]) {
  const int lane = threadIdx.x % kWave, part = threadIdx.x / kWave;
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h) red[part][h * kWave + lane] = acc[h];
  __syncthreads();
  if (part == 0) {
#pragma unroll
    for (int h = 0; h < kHmLaneRows; ++h) {
      if (!valid[h]) continue;
      double v = red[0][h * kWave + lane];
#pragma unroll
      for (int p = 1; p < kHmParts; ++p) v += red[p][h * kWave + lane];
      ys[row[h]] = v;
    }
  }
}

// Row h of the lane within the item: c.r0 + 64 h + lane.
__device__ __forceinline__ void item_rows(const mplak_hm_chunk& c, int64_t (&row)[kHmLaneRows],
                                          bool (&valid)[kHmLaneRows]) {
  const int lane = threadIdx.x % kWave;
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h) {
    valid[h] = h * kWave + lane < c.count;
    row[h] = c.r0 + (valid[h] ? h * kWave + lane : 0);
  }
}

// Near field plus far field of one apply item (up to 128 rows of one leaf,
// two per lane), closed-form entries: the leaf's inadmissible column ranges
// are staged in LDS 512 sources at a time; wave p of the workgroup's 8 takes
// sources p, p + 8, ... of each stage, and every staged source serves both of
// the lane's rows (half the LDS reads of one row per lane, at the same 8
// waves per SIMD).  The eight partial sums of a row are added in wave order.
template <int KIND, int DIM>
__global__ __launch_bounds__(kHmApplyThreads) void apply_kernel(
    KernelEntries<KIND> e, const mplak_hm_chunk* __restrict__ chunks, const int64_t* __restrict__ near_ptr,
    const int64_t* __restrict__ near_cols, const int64_t* __restrict__ far_ptr, const int32_t* __restrict__ far_blk,
    const mplak_hm_block* __restrict__ blocks, const int32_t* __restrict__ rank, int kmax,
    const double* __restrict__ U, const double* __restrict__ t, const double* __restrict__ xs,
    double* __restrict__ ys) {
  constexpr int D = DIM > 0 ? DIM : kHmMaxDim;
  __shared__ double tab[kExp2Table];
  __shared__ double src[D + 1][kHmStage];
  __shared__ double red[kHmParts][kHmRows];
  const int dim = DIM > 0 ? DIM : e.dim;
  fill_exp2_table(tab);
  const mplak_hm_chunk c = chunks[blockIdx.x];
  const int part = threadIdx.x / kWave;
  int64_t row[kHmLaneRows];
  bool valid[kHmLaneRows];
  item_rows(c, row, valid);
  double p[kHmLaneRows][D];
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h)
#pragma unroll
    for (int d = 0; d < D; ++d) p[h][d] = d < dim ? e.rows[d * e.row_ld + row[h]] : 0.0;
  double acc[kHmLaneRows];
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h) acc[h] = 0.0;
  const int64_t n0 = near_ptr[c.leaf], n1 = near_ptr[c.leaf + 1];
  for (int64_t k = n0; k < n1; ++k) {
    const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
    for (int64_t base = 0; base < cn; base += kHmStage) {
      const int cnt = (int)(cn - base < kHmStage ? cn - base : kHmStage);
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        const int64_t j = c0 + base + threadIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (d < dim) src[d][threadIdx.x] = e.cols[d * e.col_ld + j];
        src[D][threadIdx.x] = xs[j];
      }
      __syncthreads();
      for (int s = part; s < cnt; s += kHmParts) {
        double q[D];
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = d < dim ? src[d][s] : 0.0;
        const double xv = src[D][s];
#pragma unroll
        for (int h = 0; h < kHmLaneRows; ++h) {
          double r2 = 0.0;
#pragma unroll
          for (int d = 0; d < D; ++d)
            if (d < dim) {
              const double df = p[h][d] - q[d];
              r2 = fma(df, df, r2);
            }
          acc[h] = fma(kernel_entry<double, KIND>(r2, tab), xv, acc[h]);
        }
      }
    }
  }
  far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part, acc);
  finish_rows(acc, row, valid, ys, red);
}

// The same with entries from an entry object evaluated per (row, column)
// (e.g. a virtual system assembler): no staging.
template <class E>
__global__ __launch_bounds__(kHmApplyThreads) void apply_generic_kernel(
    E e, const mplak_hm_chunk* __restrict__ chunks, const int64_t* __restrict__ near_ptr,
    const int64_t* __restrict__ near_cols, const int64_t* __restrict__ far_ptr, const int32_t* __restrict__ far_blk,
    const mplak_hm_block* __restrict__ blocks, const int32_t* __restrict__ rank, int kmax,
    const double* __restrict__ U, const double* __restrict__ t, const double* __restrict__ xs,
    double* __restrict__ ys) {
  __shared__ double tab[kExp2Table];
  __shared__ double red[kHmParts][kHmRows];
  if (E::kTable) {
    fill_exp2_table(tab);
    __syncthreads();
  }
  const mplak_hm_chunk c = chunks[blockIdx.x];
  const int part = threadIdx.x / kWave;
  int64_t row[kHmLaneRows];
  bool valid[kHmLaneRows];
  item_rows(c, row, valid);
  double acc[kHmLaneRows];
  const int64_t n0 = near_ptr[c.leaf], n1 = near_ptr[c.leaf + 1];
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h) {
    acc[h] = 0.0;
    if (!valid[h]) continue;
    for (int64_t k = n0; k < n1; ++k) {
      const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
      for (int64_t s = part; s < cn; s += kHmParts) acc[h] = fma(e(row[h], c0 + s, tab), xs[c0 + s], acc[h]);
    }
  }
  far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part, acc);
  finish_rows(acc, row, valid, ys, red);
}

// Stored near field: entry (i, j) of near range k of leaf l (rows
// [lo_l, hi_l), columns [c0, c0 + cn)) at st[off_k + (i - lo_l) + j (hi_l -
// lo_l)].  One workgroup per near range fills it from an entry object.
template <class E>
__global__ __launch_bounds__(kHmThreads) void near_fill_kernel(E e, const int64_t* __restrict__ range_leaf_lo,
                                                               const int64_t* __restrict__ near_cols,
                                                               const int64_t* __restrict__ near_off,
                                                               double* __restrict__ st) {
  __shared__ double tab[kExp2Table];
  if (E::kTable) {
    fill_exp2_table(tab);
    __syncthreads();
  }
  const int64_t k = blockIdx.x;
  const int64_t lo = range_leaf_lo[2 * k], m = range_leaf_lo[2 * k + 1];
  const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
  double* out = st + near_off[k];
  for (int64_t idx = threadIdx.x; idx < m * cn; idx += kHmThreads) {
    const int64_t i = idx % m, j = idx / m;
    out[idx] = e(lo + i, c0 + j, tab);
  }
}

// Near field from the stored entries plus the far field (lanes are rows, so
// every load of the stored block is coalesced; wave p takes every eighth
// column).
__global__ __launch_bounds__(kHmApplyThreads) void apply_stored_kernel(
    const mplak_hm_chunk* __restrict__ chunks, const int64_t* __restrict__ leaf_range,
    const int64_t* __restrict__ near_ptr, const int64_t* __restrict__ near_cols,
    const int64_t* __restrict__ near_off, const double* __restrict__ st, const int64_t* __restrict__ far_ptr,
    const int32_t* __restrict__ far_blk, const mplak_hm_block* __restrict__ blocks, const int32_t* __restrict__ rank,
    int kmax, const double* __restrict__ U, const double* __restrict__ t, const double* __restrict__ xs,
    double* __restrict__ ys) {
  __shared__ double red[kHmParts][kHmRows];
  const mplak_hm_chunk c = chunks[blockIdx.x];
  const int part = threadIdx.x / kWave;
  int64_t row[kHmLaneRows];
  bool valid[kHmLaneRows];
  item_rows(c, row, valid);
  const int64_t lo = leaf_range[2 * c.leaf], m = leaf_range[2 * c.leaf + 1];
  double acc[kHmLaneRows];
#pragma unroll
  for (int h = 0; h < kHmLaneRows; ++h) acc[h] = 0.0;
  const int64_t n0 = near_ptr[c.leaf], n1 = near_ptr[c.leaf + 1];
  for (int64_t k = n0; k < n1; ++k) {
    const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
    const double* a = st + near_off[k];
    for (int64_t s = part; s < cn; s += kHmParts) {
      const double xv = xs[c0 + s];
#pragma unroll
      for (int h = 0; h < kHmLaneRows; ++h)
        if (valid[h]) acc[h] = fma(a[(row[h] - lo) + s * m], xv, acc[h]);
    }
  }
  far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part, acc);
  finish_rows(acc, row, valid, ys, red);
}

}  // namespace hm
}  // namespace mplak
