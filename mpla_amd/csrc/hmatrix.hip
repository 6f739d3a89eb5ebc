// H-matrix operator kernels with closed-form kernel entries (C ABI of
// include/mplak.h): ACA compression of the admissible blocks, the far-field
// coefficient products and the near + far apply.  The block partition comes
// from the host (libmpla's H-matrix plugin, which replaces hmglib behind
// src/hmglib_adapter.cu:6-132).

#include "common.hpp"
#include "entry.hpp"
#include "hmatrix_kernels.hpp"

namespace mplak {
namespace {

using hm::kHmThreads;

__global__ __launch_bounds__(kBlock) void gather_scaled_kernel(const double* const* __restrict__ points, int dim,
                                                               int64_t offset, const int64_t* __restrict__ perm,
                                                               int64_t n, double scale, double* __restrict__ out,
                                                               int64_t ld) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const int64_t i = offset + perm[k];
    for (int d = 0; d < dim; ++d) out[d * ld + k] = scale * points[d][i];
  }
}

// One workgroup per item: the item's columns against the block's rank terms.
// Wave w takes the terms q = w, w + 4, ...: its lanes walk the item's columns
// (coalesced loads of V's column q and of x), then a butterfly sums the
// lanes.  No barriers; every sum has a fixed order.
__global__ __launch_bounds__(kHmThreads) void far_partial_kernel(const mplak_hm_block* __restrict__ blocks,
                                                                 const mplak_hm_item* __restrict__ items,
                                                                 const int32_t* __restrict__ rank, int kmax,
                                                                 const double* __restrict__ V,
                                                                 const double* __restrict__ xs,
                                                                 double* __restrict__ part) {
  const mplak_hm_item it = items[blockIdx.x];
  const mplak_hm_block B = blocks[it.block];
  const int rb = rank[it.block];
  const double* Vb = V + B.voff;
  const double* x = xs + B.c0;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  double* out = part + (int64_t)blockIdx.x * kmax;
  for (int q = w; q < rb; q += kHmThreads / kWave) {
    const double* v = Vb + (int64_t)q * B.n;
    double s = 0.0;
    for (int64_t j = it.j0 + lane; j < it.j0 + it.count; j += kWave) s = fma(v[j], x[j], s);
    s = wave_sum(s);
    if (lane == 0) out[q] = s;
  }
}

// One thread per (block, term): the block's item partials in item order.
__global__ __launch_bounds__(kBlock) void far_sum_kernel(const int64_t* __restrict__ first_item, int64_t nblocks,
                                                         const int32_t* __restrict__ rank, int kmax,
                                                         const double* __restrict__ part, double* __restrict__ t) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nblocks * kmax) return;
  const int64_t b = g / kmax;
  const int q = (int)(g % kmax);
  double s = 0.0;
  if (q < rank[b])
    for (int64_t i = first_item[b]; i < first_item[b + 1]; ++i) s += part[i * kmax + q];
  t[g] = s;
}

__global__ __launch_bounds__(kBlock) void scatter_nugget_kernel(const double* __restrict__ ys,
                                                                const int64_t* __restrict__ perm, int64_t n,
                                                                double nugget, const double* __restrict__ x,
                                                                int64_t shift, int64_t x_count,
                                                                double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
    const int64_t i = perm[k];
    const int64_t j = i + shift;
    double v = ys[k];
    if (nugget != 0.0 && j >= 0 && j < x_count) v = fma(nugget, x[j], v);
    y[i] = v;
  }
}

// ---- symmetric near field --------------------------------------------------
//
// One workgroup per pair (A, B) of near chunks (at most 512 sorted points
// each).  Wave w holds rows 128 w + 64 h + lane of A (h < 2) and walks the
// eight 64-point groups of B staged in LDS.  Off the diagonal, lane l takes
// source (l + s) mod 64 at step s, so one evaluation of each entry feeds the
// row sums of its two A rows and a column sum for its B row that rotates one
// lane per step and ends in the lane of its source; the four waves' column
// sums are added in wave order.  The diagonal pair (A = B) evaluates its
// whole block for the row sums.
constexpr int kPairRows = 2 * kBlock;
constexpr int kPairGroups = kPairRows / kWave;

template <int KIND, int DIM>
__global__ __launch_bounds__(kBlock) void near_pair_kernel(const double* __restrict__ pts, int64_t ld,
                                                           const mplak_hm_pair* __restrict__ pairs,
                                                           const double* __restrict__ xs, double* __restrict__ part) {
  __shared__ double4 src[kPairRows];  // coordinates (up to 3) and the source value in w
  __shared__ double crd4[DIM > 3 ? kPairRows : 1];
  __shared__ double red[kBlock / kWave][kWave];
  __shared__ double tab[kExp2Table];
  fill_exp2_table(tab);  // first read after the staging barrier
  const mplak_hm_pair P = pairs[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  // every global load of the staging is issued before the first use
  double pi[2][DIM], xi[2], sc[2][DIM], sv[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t k = w * 2 * kWave + h * kWave + lane;
    const int64_t i = P.a0 + (k < P.am ? k : P.am - 1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) pi[h][d] = pts[d * ld + i];
    xi[h] = k < P.am ? xs[i] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t k = u * kBlock + tid;
    const int64_t j = P.b0 + (k < P.bm ? k : P.bm - 1);
#pragma unroll
    for (int d = 0; d < DIM; ++d) sc[u][d] = pts[d * ld + j];
    sv[u] = k < P.bm ? xs[j] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int k = u * kBlock + tid;
    double4 q;
    q.x = sc[u][0];
    q.y = DIM > 1 ? sc[u][DIM > 1 ? 1 : 0] : 0.0;
    q.z = DIM > 2 ? sc[u][DIM > 2 ? 2 : 0] : 0.0;
    q.w = sv[u];
    src[k] = q;
    if (DIM > 3) crd4[DIM > 3 ? k : 0] = sc[u][DIM - 1];
  }
  __syncthreads();
  auto entries = [&](int jj, double (&e)[2]) {
    const double4 q = src[jj];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double dx = pi[h][0] - q.x;
      double s = dx * dx;
      if constexpr (DIM > 1) {
        const double dy = pi[h][1] - q.y;
        s = fma(dy, dy, s);
      }
      if constexpr (DIM > 2) {
        const double dz = pi[h][2] - q.z;
        s = fma(dz, dz, s);
      }
      if constexpr (DIM > 3) {
        const double dw = pi[h][DIM - 1] - crd4[DIM > 3 ? jj : 0];
        s = fma(dw, dw, s);
      }
      e[h] = kernel_entry<double, KIND>(s, tab);
    }
  };
  double acc[2] = {0.0, 0.0};
  double* rows_a = part + (int64_t)blockIdx.x * 2 * kPairRows;
  // rows and sources past am / bm are padding: a wave without rows and a
  // group without sources skip their evaluations (not their barriers)
  const bool wave_rows = w * 2 * kWave < P.am;
  if (P.a0 == P.b0) {
    const int cnt = wave_rows ? (int)P.bm : 0;
#pragma unroll 1
    for (int jj = 0; jj < cnt; ++jj) {
      double e[2];
      entries(jj, e);
      const double v = src[jj].w;
      acc[0] = fma(e[0], v, acc[0]);
      acc[1] = fma(e[1], v, acc[1]);
    }
  } else {
    double* rows_b = rows_a + kPairRows;
#pragma unroll 1
    for (int g = 0; g < kPairGroups; ++g) {
      double t = 0.0;
      const int steps = wave_rows && g * kWave < P.bm ? kWave : 0;
#pragma unroll 4
      for (int s = 0; s < steps; ++s) {
        const int jj = g * kWave + ((lane + s) & (kWave - 1));
        double e[2];
        entries(jj, e);
        const double v = src[jj].w;
        acc[0] = fma(e[0], v, acc[0]);
        acc[1] = fma(e[1], v, acc[1]);
        t = fma(e[1], xi[1], fma(e[0], xi[0], t));
        t = wave_rotate(t);
      }
      red[w][lane] = t;
      __syncthreads();
      if (tid < kWave) {
        double sum = red[0][tid];
#pragma unroll
        for (int ww = 1; ww < kBlock / kWave; ++ww) sum += red[ww][tid];
        rows_b[g * kWave + tid] = sum;
      }
      __syncthreads();
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) rows_a[w * 2 * kWave + h * kWave + lane] = acc[h];
}

// Far field plus the near-pair partials of one apply chunk: wave p of the
// workgroup's 8 adds the chunk's partial blocks p, p + 8, ... and its share
// of the admissible blocks; the eight sums of a row are added in wave order.
__global__ __launch_bounds__(hm::kHmApplyThreads) void apply_pairs_kernel(
    const mplak_hm_chunk* __restrict__ chunks, const int64_t* __restrict__ chunk_near,
    const int64_t* __restrict__ near_start, const int64_t* __restrict__ ent_ptr, const int64_t* __restrict__ ent,
    const double* __restrict__ part, const int64_t* __restrict__ far_ptr, const int32_t* __restrict__ far_blk,
    const mplak_hm_block* __restrict__ blocks, const int32_t* __restrict__ rank, int kmax,
    const double* __restrict__ U, const double* __restrict__ t, double* __restrict__ ys) {
  __shared__ double red[hm::kHmParts][hm::kHmRows];
  const mplak_hm_chunk c = chunks[blockIdx.x];
  const int part_w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  int64_t row[hm::kHmLaneRows];
  bool valid[hm::kHmLaneRows];
  hm::item_rows(c, row, valid);
  const int64_t nc = chunk_near[blockIdx.x];
  const int64_t off = c.r0 - near_start[nc];
  double acc[hm::kHmLaneRows];
#pragma unroll
  for (int h = 0; h < hm::kHmLaneRows; ++h) acc[h] = 0.0;
  for (int64_t k = ent_ptr[nc] + part_w; k < ent_ptr[nc + 1]; k += hm::kHmParts) {
    const double* blk = part + ent[k] * kPairRows + off;
#pragma unroll
    for (int h = 0; h < hm::kHmLaneRows; ++h)
      if (valid[h]) acc[h] += blk[h * kWave + lane];
  }
  hm::far_rows(row, valid, far_ptr, far_blk, c.leaf, blocks, rank, kmax, U, t, part_w, acc);
  hm::finish_rows(acc, row, valid, ys, red);
}

template <int KIND, int DIM>
void launch_near_pairs(const mplak_hm_points* p, const mplak_hm_pair* pairs, int64_t npairs, const double* xs,
                       double* part, hipStream_t s) {
  hipLaunchKernelGGL((near_pair_kernel<KIND, DIM>), dim3((unsigned)npairs), dim3(kBlock), 0, s, p->rows, p->row_ld,
                     pairs, xs, part);
}

template <int KIND>
void near_pairs_kind(const mplak_hm_points* p, const mplak_hm_pair* pairs, int64_t npairs, const double* xs,
                     double* part, hipStream_t s) {
  switch (p->dim) {
    case 1: launch_near_pairs<KIND, 1>(p, pairs, npairs, xs, part, s); break;
    case 2: launch_near_pairs<KIND, 2>(p, pairs, npairs, xs, part, s); break;
    case 3: launch_near_pairs<KIND, 3>(p, pairs, npairs, xs, part, s); break;
    default: launch_near_pairs<KIND, 4>(p, pairs, npairs, xs, part, s); break;
  }
}

template <int KIND>
int launch_aca(const mplak_hm_points* p, const mplak_hm_block* blocks, int64_t nblocks, int kmax, double eps,
               double* U, double* V, int32_t* rank, hipStream_t s) {
  hm::KernelEntries<KIND> e{p->rows, p->cols, p->row_ld, p->col_ld, p->dim};
  hipLaunchKernelGGL(hm::aca_kernel<hm::KernelEntries<KIND>>, dim3((unsigned)nblocks), dim3(kHmThreads), 0, s, e,
                     blocks, kmax, eps, U, V, rank);
  return launch_status();
}

template <int KIND, int DIM>
void launch_apply(const hm::KernelEntries<KIND>& e, int64_t nchunks, hipStream_t s, const mplak_hm_chunk* chunks,
                  const int64_t* near_ptr, const int64_t* near_cols, const int64_t* far_ptr, const int32_t* far_blk,
                  const mplak_hm_block* blocks, const int32_t* rank, int kmax, const double* U, const double* t,
                  const double* xs, double* ys) {
  hipLaunchKernelGGL((hm::apply_kernel<KIND, DIM>), dim3((unsigned)nchunks), dim3(hm::kHmApplyThreads), 0, s, e, chunks,
                     near_ptr, near_cols, far_ptr, far_blk, blocks, rank, kmax, U, t, xs, ys);
}

template <int KIND>
int apply_kind(const mplak_hm_points* p, const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* near_ptr,
               const int64_t* near_cols, const int64_t* far_ptr, const int32_t* far_blk,
               const mplak_hm_block* blocks, const int32_t* rank, int kmax, const double* U, const double* t,
               const double* xs, double* ys, hipStream_t s) {
  hm::KernelEntries<KIND> e{p->rows, p->cols, p->row_ld, p->col_ld, p->dim};
  switch (p->dim) {
    case 1: launch_apply<KIND, 1>(e, nchunks, s, chunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank, kmax,
                                  U, t, xs, ys); break;
    case 2: launch_apply<KIND, 2>(e, nchunks, s, chunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank, kmax,
                                  U, t, xs, ys); break;
    case 3: launch_apply<KIND, 3>(e, nchunks, s, chunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank, kmax,
                                  U, t, xs, ys); break;
    default: launch_apply<KIND, 0>(e, nchunks, s, chunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank, kmax,
                                   U, t, xs, ys); break;
  }
  return launch_status();
}

int check_points(const mplak_hm_points* p) {
  if (p == nullptr || p->rows == nullptr || p->cols == nullptr || p->dim < 1 || p->dim > hm::kHmMaxDim ||
      p->row_ld < 0 || p->col_ld < 0)
    return MPLAK_ERR_ARG;
  if (p->kind != MPLAK_KERNEL_GAUSSIAN && p->kind != MPLAK_KERNEL_MATERN32) return MPLAK_ERR_ARG;
  return MPLAK_OK;
}

int64_t grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  return g < 1 ? 1 : (g > 4096 ? 4096 : g);
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

int mplak_hm_gather_scaled(const double* const* points, int32_t dim, int64_t offset, const int64_t* perm, int64_t n,
                           int32_t kind, double length_scale, double* out, int64_t ld, mplak_stream_t stream) {
  if (points == nullptr || perm == nullptr || out == nullptr || dim < 1 || n < 0 || offset < 0 || ld < n ||
      !(length_scale > 0.0) || (kind != MPLAK_KERNEL_GAUSSIAN && kind != MPLAK_KERNEL_MATERN32))
    return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  hipLaunchKernelGGL(gather_scaled_kernel, dim3((unsigned)grid_for(n)), dim3(kBlock), 0, as_stream(stream), points,
                     dim, offset, perm, n, coord_scale(kind, length_scale), out, ld);
  return launch_status();
}

int mplak_hm_aca(const mplak_hm_points* pts, const mplak_hm_block* blocks, int64_t nblocks, int32_t kmax, double eps,
                 double* U, double* V, int32_t* rank, mplak_stream_t stream) {
  int st = check_points(pts);
  if (st != MPLAK_OK) return st;
  if (nblocks < 0 || kmax < 1 || kmax > MPLAK_HM_MAX_RANK || !(eps >= 0.0) || nblocks > 0x7fffffff) return MPLAK_ERR_ARG;
  if (nblocks == 0) return MPLAK_OK;
  if (blocks == nullptr || U == nullptr || V == nullptr || rank == nullptr) return MPLAK_ERR_ARG;
  hipStream_t s = as_stream(stream);
  if (pts->kind == MPLAK_KERNEL_GAUSSIAN)
    return launch_aca<MPLAK_KERNEL_GAUSSIAN>(pts, blocks, nblocks, kmax, eps, U, V, rank, s);
  return launch_aca<MPLAK_KERNEL_MATERN32>(pts, blocks, nblocks, kmax, eps, U, V, rank, s);
}

int mplak_hm_far_partials(const mplak_hm_block* blocks, const mplak_hm_item* items, int64_t nitems,
                          const int32_t* rank, int32_t kmax, const double* V, const double* xs, double* part,
                          mplak_stream_t stream) {
  if (nitems < 0 || nitems > 0x7fffffff || kmax < 1 || kmax > MPLAK_HM_MAX_RANK) return MPLAK_ERR_ARG;
  if (nitems == 0) return MPLAK_OK;
  if (blocks == nullptr || items == nullptr || rank == nullptr || V == nullptr || xs == nullptr || part == nullptr)
    return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(far_partial_kernel, dim3((unsigned)nitems), dim3(kHmThreads), 0, as_stream(stream), blocks,
                     items, rank, kmax, V, xs, part);
  return launch_status();
}

int mplak_hm_far_sum(const int64_t* first_item, int64_t nblocks, const int32_t* rank, int32_t kmax,
                     const double* part, double* t, mplak_stream_t stream) {
  if (nblocks < 0 || kmax < 1 || kmax > MPLAK_HM_MAX_RANK) return MPLAK_ERR_ARG;
  if (nblocks == 0) return MPLAK_OK;
  if (first_item == nullptr || rank == nullptr || part == nullptr || t == nullptr) return MPLAK_ERR_ARG;
  const int64_t total = nblocks * kmax;
  const int64_t g = (total + kBlock - 1) / kBlock;
  if (g > 0x7fffffff) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(far_sum_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(stream), first_item, nblocks, rank,
                     kmax, part, t);
  return launch_status();
}

int mplak_hm_apply(const mplak_hm_points* pts, const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* near_ptr,
                   const int64_t* near_cols, const int64_t* far_ptr, const int32_t* far_blk,
                   const mplak_hm_block* blocks, const int32_t* rank, int32_t kmax, const double* U, const double* t,
                   const double* xs, double* ys, mplak_stream_t stream) {
  int st = check_points(pts);
  if (st != MPLAK_OK) return st;
  if (nchunks < 0 || nchunks > 0x7fffffff || kmax < 1 || kmax > MPLAK_HM_MAX_RANK) return MPLAK_ERR_ARG;
  if (nchunks == 0) return MPLAK_OK;
  if (chunks == nullptr || near_ptr == nullptr || far_ptr == nullptr || xs == nullptr || ys == nullptr)
    return MPLAK_ERR_ARG;
  hipStream_t s = as_stream(stream);
  if (pts->kind == MPLAK_KERNEL_GAUSSIAN)
    return apply_kind<MPLAK_KERNEL_GAUSSIAN>(pts, chunks, nchunks, near_ptr, near_cols, far_ptr, far_blk, blocks,
                                             rank, kmax, U, t, xs, ys, s);
  return apply_kind<MPLAK_KERNEL_MATERN32>(pts, chunks, nchunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank,
                                           kmax, U, t, xs, ys, s);
}

int mplak_hm_near_pairs(const mplak_hm_points* pts, const mplak_hm_pair* pairs, int64_t npairs, const double* xs,
                        double* part, mplak_stream_t stream) {
  int st = check_points(pts);
  if (st != MPLAK_OK) return st;
  if (npairs < 0 || npairs > 0x7fffffff) return MPLAK_ERR_ARG;
  if (npairs == 0) return MPLAK_OK;
  if (pairs == nullptr || xs == nullptr || part == nullptr) return MPLAK_ERR_ARG;
  hipStream_t s = as_stream(stream);
  if (pts->kind == MPLAK_KERNEL_GAUSSIAN)
    near_pairs_kind<MPLAK_KERNEL_GAUSSIAN>(pts, pairs, npairs, xs, part, s);
  else
    near_pairs_kind<MPLAK_KERNEL_MATERN32>(pts, pairs, npairs, xs, part, s);
  return launch_status();
}

int mplak_hm_apply_pairs(const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* chunk_near,
                         const int64_t* near_start, const int64_t* ent_ptr, const int64_t* ent, const double* part,
                         const int64_t* far_ptr, const int32_t* far_blk, const mplak_hm_block* blocks,
                         const int32_t* rank, int32_t kmax, const double* U, const double* t, double* ys,
                         mplak_stream_t stream) {
  if (nchunks < 0 || nchunks > 0x7fffffff || kmax < 1 || kmax > MPLAK_HM_MAX_RANK) return MPLAK_ERR_ARG;
  if (nchunks == 0) return MPLAK_OK;
  if (chunks == nullptr || chunk_near == nullptr || near_start == nullptr || ent_ptr == nullptr || far_ptr == nullptr ||
      ys == nullptr)
    return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(apply_pairs_kernel, dim3((unsigned)nchunks), dim3(hm::kHmApplyThreads), 0, as_stream(stream),
                     chunks, chunk_near, near_start, ent_ptr, ent, part, far_ptr, far_blk, blocks, rank, kmax, U, t,
                     ys);
  return launch_status();
}

int mplak_hm_apply_stored(const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* leaf_range,
                          const int64_t* near_ptr, const int64_t* near_cols, const int64_t* near_off,
                          const double* stored, const int64_t* far_ptr, const int32_t* far_blk,
                          const mplak_hm_block* blocks, const int32_t* rank, int32_t kmax, const double* U,
                          const double* t, const double* xs, double* ys, mplak_stream_t stream) {
  if (nchunks < 0 || nchunks > 0x7fffffff || kmax < 1 || kmax > MPLAK_HM_MAX_RANK) return MPLAK_ERR_ARG;
  if (nchunks == 0) return MPLAK_OK;
  if (chunks == nullptr || leaf_range == nullptr || near_ptr == nullptr || far_ptr == nullptr || xs == nullptr ||
      ys == nullptr)
    return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(hm::apply_stored_kernel, dim3((unsigned)nchunks), dim3(hm::kHmApplyThreads), 0, as_stream(stream), chunks,
                     leaf_range, near_ptr, near_cols, near_off, stored, far_ptr, far_blk, blocks, rank, kmax, U, t, xs,
                     ys);
  return launch_status();
}

int mplak_hm_scatter_nugget(const double* ys, const int64_t* perm, int64_t n, double nugget, const double* x,
                            int64_t shift, int64_t x_count, double* y, mplak_stream_t stream) {
  if (n < 0 || x_count < 0) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  if (ys == nullptr || perm == nullptr || y == nullptr || (nugget != 0.0 && x == nullptr)) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(scatter_nugget_kernel, dim3((unsigned)grid_for(n)), dim3(kBlock), 0, as_stream(stream), ys, perm,
                     n, nugget, x, shift, x_count, y);
  return launch_status();
}

}  // extern "C"
