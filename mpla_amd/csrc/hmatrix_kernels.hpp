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
constexpr int kHmRows = 64;                     // rows of one apply work item (one per lane)
constexpr int kHmParts = kHmThreads / kHmRows;  // waves splitting one row's sources
constexpr int kHmStage = 256;                   // sources staged in LDS at a time
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

// Far field for one row of an apply item: sum over the leaf's admissible
// blocks b (those of part `part` of kHmParts) of U_b(row - r0_b, :) t_b.
__device__ __forceinline__ double far_rows(int64_t row, bool valid, const int64_t* __restrict__ far_ptr,
                                           const int32_t* __restrict__ far_blk, int32_t leaf,
                                           const mplak_hm_block* __restrict__ blocks,
                                           const int32_t* __restrict__ rank, int kmax, const double* __restrict__ U,
                                           const double* __restrict__ t, int part) {
  double acc = 0.0;
  const int64_t f0 = far_ptr[leaf], f1 = far_ptr[leaf + 1];
  for (int64_t f = f0 + part; f < f1; f += kHmParts) {
    const int32_t b = far_blk[f];
    const mplak_hm_block B = blocks[b];
    const int rb = rank[b];
    const double* tb = t + (int64_t)b * kmax;
    if (valid) {
      const double* u = U + B.uoff + (row - B.r0);
      for (int q = 0; q < rb; ++q) acc = fma(u[q * B.m], tb[q], acc);
    }
  }
  return acc;
}

// Adds the kHmParts partial sums of each row in part order and stores them.
__device__ __forceinline__ void finish_rows(double acc, int64_t row, bool valid, double* __restrict__ ys,
                                            double (*red)[kHmRows]) {
  const int lane = threadIdx.x % kHmRows, part = threadIdx.x / kHmRows;
  red[part][lane] = acc;
  __syncthreads();
  if (part == 0 && valid) {
    double s = red[0][lane];
#pragma unroll
    for (int p = 1; p < kHmParts; ++p) s += red[p][lane];
    ys[row] = s;
  }
}

// Near field plus far field of one apply item (up to 64 rows of one leaf),
// closed-form entries: the leaf's inadmissible column ranges are staged in
// LDS 256 sources at a time; wave p of the workgroup takes sources p, p+4,
// ... of each stage for its 64 rows.
template <int KIND, int DIM>
__global__ __launch_bounds__(kHmThreads) void apply_kernel(
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
  const int lane = threadIdx.x % kHmRows, part = threadIdx.x / kHmRows;
  const int64_t row = c.r0 + lane;
  const bool valid = lane < c.count;
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = (d < dim && valid) ? e.rows[d * e.row_ld + row] : 0.0;
  double acc = 0.0;
  const int64_t n0 = near_ptr[c.leaf], n1 = near_ptr[c.leaf + 1];
  for (int64_t k = n0; k < n1; ++k) {
    const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
    for (int64_t base = 0; base < cn; base += kHmStage) {
      const int cnt = (int)(cn - base < kHmStage ? cn - base : kHmStage);
      __syncthreads();
      if (threadIdx.x < cnt) {
        const int64_t j = c0 + base + threadIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (d < dim) src[d][threadIdx.x] = e.cols[d * e.col_ld + j];
        src[D][threadIdx.x] = xs[j];
      }
      __syncthreads();
      for (int s = part; s < cnt; s += kHmParts) {
        double q = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d)
          if (d < dim) {
            const double df = p[d] - src[d][s];
            q = fma(df, df, q);
          }
        acc = fma(kernel_entry<double, KIND>(q, tab), src[D][s], acc);
      }
    }
  }
  acc += far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part);
  finish_rows(acc, row, valid, ys, red);
}

// The same with entries from an entry object evaluated per (row, column)
// (e.g. a virtual system assembler): no staging.
template <class E>
__global__ __launch_bounds__(kHmThreads) void apply_generic_kernel(
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
  const int lane = threadIdx.x % kHmRows, part = threadIdx.x / kHmRows;
  const int64_t row = c.r0 + lane;
  const bool valid = lane < c.count;
  double acc = 0.0;
  if (valid) {
    const int64_t n0 = near_ptr[c.leaf], n1 = near_ptr[c.leaf + 1];
    for (int64_t k = n0; k < n1; ++k) {
      const int64_t c0 = near_cols[2 * k], cn = near_cols[2 * k + 1];
      for (int64_t s = part; s < cn; s += kHmParts) acc = fma(e(row, c0 + s, tab), xs[c0 + s], acc);
    }
  }
  acc += far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part);
  finish_rows(acc, row, valid, ys, red);
}

}  // namespace hm
}  // namespace mplak
