// Spatial ordering of points (no counterpart in the reference).
//
// The matrix-free products touch every (target, source) pair; when the points
// are stored in spatial order, consecutive blocks of points are compact
// clusters, and a screened product can skip whole blocks whose entries are
// negligible.  The order is the Hilbert curve over a uniform grid on the
// points' bounding box, with 30 / dim bits per axis; points in one grid cell
// keep their index order.  Unlike the Z-order (Morton) curve, the Hilbert
// curve never jumps: consecutive cells are neighbours, so every run of
// consecutive points is one connected, compact cluster (a Z-order run that
// crosses a quadrant or octant boundary spans two distant regions, and its
// bounding box covers both).  The cell index along the curve is Skilling's
// transpose form (J. Skilling, "Programming the Hilbert curve", AIP Conf.
// Proc. 707, 381, 2004), bit-interleaved.  Sorting uses rocPRIM's device
// radix sort on 64-bit keys (cell index << 32 | point index) over the
// cell-index bits only; the sort is stable, so the permutation is fully
// determined by the points.

#include <algorithm>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"

namespace mplak {
namespace {

constexpr int kBoxBlocks = 256;
constexpr int kMaxDim = 16;

// Per-block min/max of every coordinate: part[blk][d] = min, part[blk][dim + d] = max.
__global__ __launch_bounds__(kBlock) void box_partial_kernel(const double* const* __restrict__ points, int dim,
                                                             int64_t offset, int64_t n, double* __restrict__ part) {
  __shared__ double red[2][kBlock / kWave];
  for (int d = 0; d < dim; ++d) {
    const double* p = points[d] + offset;
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
      lo = fmin(lo, p[i]);
      hi = fmax(hi, p[i]);
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
      lo = fmin(lo, __shfl_xor(lo, m));
      hi = fmax(hi, __shfl_xor(hi, m));
    }
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
      red[0][w] = lo;
      red[1][w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < kBlock / kWave; ++k) {
        lo = fmin(lo, red[0][k]);
        hi = fmax(hi, red[1][k]);
      }
      part[(size_t)blockIdx.x * 2 * dim + d] = lo;
      part[(size_t)blockIdx.x * 2 * dim + dim + d] = hi;
    }
  }
}

// Spreads the low `bits` bits of v so that bit b lands at position b * dim + slot.
__device__ __forceinline__ uint64_t spread_bits(uint64_t v, int bits, int dim, int slot) {
  uint64_t out = 0;
  for (int b = 0; b < bits; ++b) out |= ((v >> b) & 1ull) << (b * dim + slot);
  return out;
}

// Hilbert index of the cell with coordinates X[0..D) (b bits each):
// Skilling's axes-to-transpose map (undo the excess work of the recursive
// rotations and reflections, then Gray-encode), after which interleaving the
// bits of X[0], X[1], ... from the top gives the index along the curve.
template <int D>
__device__ __forceinline__ uint64_t hilbert_index(uint32_t (&X)[D], int dim, int b) {
  const uint32_t M = 1u << (b - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    const uint32_t P = Q - 1;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      if (i >= dim) break;
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
#pragma unroll
  for (int i = 1; i < D; ++i)
    if (i < dim) X[i] ^= X[i - 1];
  uint32_t t = 0, last = 0;
#pragma unroll
  for (int i = 0; i < D; ++i)
    if (i == dim - 1) last = X[i];
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (last & Q) t ^= Q - 1;
#pragma unroll
  for (int i = 0; i < D; ++i)
    if (i < dim) X[i] ^= t;
  uint64_t code = 0;
#pragma unroll
  for (int i = 0; i < D; ++i)
    if (i < dim) code |= spread_bits(X[i], b, dim, dim - 1 - i);
  return code;
}

template <int D>
__global__ __launch_bounds__(kBlock) void hilbert_key_kernel(const double* const* __restrict__ points, int dim,
                                                             int64_t offset, int64_t n,
                                                             const double* __restrict__ part, int nparts,
                                                             uint64_t* __restrict__ keys) {
  // the bounding box from the per-block partial boxes, reduced by the whole
  // block (minima and maxima are exact, so any order gives the same box)
  __shared__ double box[2 * kMaxDim];
  __shared__ double red[2][kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  for (int d = 0; d < dim; ++d) {
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < nparts; b += kBlock) {
      lo = fmin(lo, part[(size_t)b * 2 * dim + d]);
      hi = fmax(hi, part[(size_t)b * 2 * dim + dim + d]);
    }
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
      lo = fmin(lo, __shfl_xor(lo, m));
      hi = fmax(hi, __shfl_xor(hi, m));
    }
    if (lane == 0) {
      red[0][w] = lo;
      red[1][w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < kBlock / kWave; ++k) {
        lo = fmin(lo, red[0][k]);
        hi = fmax(hi, red[1][k]);
      }
      box[d] = lo;
      box[kMaxDim + d] = hi;
    }
    __syncthreads();
  }
  const int bits = 30 / dim;
  const double cells = (double)(1u << bits);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    uint32_t X[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      X[d] = 0;
      if (d >= dim) continue;
      const double lo = box[d], ext = box[kMaxDim + d] - lo;
      double c = ext > 0.0 ? floor((points[d][offset + i] - lo) / ext * cells) : 0.0;
      c = fmin(fmax(c, 0.0), cells - 1.0);
      X[d] = (uint32_t)c;
    }
    const uint64_t code = hilbert_index<D>(X, dim, bits);
    keys[i] = (code << 32) | (uint64_t)i;
  }
}

__global__ __launch_bounds__(kBlock) void key_index_kernel(const uint64_t* __restrict__ keys, int64_t n,
                                                           int64_t* __restrict__ perm) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    perm[i] = (int64_t)(keys[i] & 0xffffffffull);
}

__global__ __launch_bounds__(kBlock) void gather_points_kernel(const double* const* __restrict__ points, int dim,
                                                               int64_t offset, const int64_t* __restrict__ perm,
                                                               int64_t n, double* __restrict__ out, int64_t ld) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t j = offset + perm[i];
    for (int d = 0; d < dim; ++d) out[d * ld + i] = points[d][j];
  }
}

__global__ __launch_bounds__(kBlock) void gather_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                        const int64_t* __restrict__ perm, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[perm[i]];
}

__global__ __launch_bounds__(kBlock) void scatter_kernel(double* __restrict__ dst, const double* __restrict__ src,
                                                         const int64_t* __restrict__ perm, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    dst[perm[i]] = src[i];
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  if (g < 1) g = 1;
  return (unsigned)g;
}

size_t sort_bytes(int64_t n) {
  size_t tb = 0;
  if (rocprim::radix_sort_keys(nullptr, tb, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                               (unsigned)n, 32, 62) != hipSuccess)
    return 0;
  return tb;
}

struct OrderLayout {
  size_t keys_in, keys_out, part, tmp, total, tmp_bytes;
};

OrderLayout order_layout(int64_t n, int dim) {
  OrderLayout L;
  L.keys_in = 0;
  L.keys_out = align_up(sizeof(uint64_t) * (size_t)n, 256);
  L.part = L.keys_out + align_up(sizeof(uint64_t) * (size_t)n, 256);
  L.tmp = L.part + align_up(sizeof(double) * 2 * (size_t)dim * kBoxBlocks, 256);
  L.tmp_bytes = sort_bytes(n);
  L.total = L.tmp_bytes == 0 ? 0 : L.tmp + align_up(L.tmp_bytes, 256);
  return L;
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

size_t mplak_spatial_order_workspace_bytes(int64_t n, int32_t dim) {
  if (n <= 0 || n >= (int64_t(1) << 31) || dim < 1 || dim > kMaxDim) return 0;
  return order_layout(n, dim).total;
}

int mplak_spatial_order(const double* const* points, int32_t dim, int64_t offset, int64_t n, int64_t* perm,
                        void* work, size_t work_bytes, mplak_stream_t s) {
  if (points == nullptr || dim < 1 || dim > kMaxDim || offset < 0 || n < 0 || n >= (int64_t(1) << 31))
    return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  if (perm == nullptr) return MPLAK_ERR_ARG;
  const OrderLayout L = order_layout(n, dim);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (work == nullptr || work_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  char* base = static_cast<char*>(work);
  uint64_t* kin = reinterpret_cast<uint64_t*>(base + L.keys_in);
  uint64_t* kout = reinterpret_cast<uint64_t*>(base + L.keys_out);
  double* part = reinterpret_cast<double*>(base + L.part);
  hipStream_t hs = as_stream(s);
  const int nb = (int)std::min<int64_t>(kBoxBlocks, (n + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(box_partial_kernel, dim3(nb), dim3(kBlock), 0, hs, points, dim, offset, n, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  // dims 1-3 with their own register arrays; larger dims share one of kMaxDim
  if (dim == 1)
    hipLaunchKernelGGL(hilbert_key_kernel<1>, dim3(grid_for(n)), dim3(kBlock), 0, hs, points, dim, offset, n, part,
                       nb, kin);
  else if (dim == 2)
    hipLaunchKernelGGL(hilbert_key_kernel<2>, dim3(grid_for(n)), dim3(kBlock), 0, hs, points, dim, offset, n, part,
                       nb, kin);
  else if (dim == 3)
    hipLaunchKernelGGL(hilbert_key_kernel<3>, dim3(grid_for(n)), dim3(kBlock), 0, hs, points, dim, offset, n, part,
                       nb, kin);
  else
    hipLaunchKernelGGL(hilbert_key_kernel<kMaxDim>, dim3(grid_for(n)), dim3(kBlock), 0, hs, points, dim, offset, n,
                       part, nb, kin);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  size_t tb = L.tmp_bytes;
  if (rocprim::radix_sort_keys(base + L.tmp, tb, kin, kout, (unsigned)n, 32, 62, hs) != hipSuccess)
    return MPLAK_ERR_HIP;
  hipLaunchKernelGGL(key_index_kernel, dim3(grid_for(n)), dim3(kBlock), 0, hs, kout, n, perm);
  return launch_status();
}

int mplak_gather_points(const double* const* points, int32_t dim, int64_t offset, const int64_t* perm, int64_t n,
                        double* out, int64_t ld, mplak_stream_t s) {
  if (points == nullptr || dim < 1 || offset < 0 || n < 0 || ld < n) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  if (perm == nullptr || out == nullptr) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(gather_points_kernel, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), points, dim, offset,
                     perm, n, out, ld);
  return launch_status();
}

int mplak_gather_f64(double* dst, const double* src, const int64_t* perm, int64_t n, mplak_stream_t s) {
  if (n < 0) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  if (dst == nullptr || src == nullptr || perm == nullptr || dst == src) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(gather_kernel, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), dst, src, perm, n);
  return launch_status();
}

int mplak_scatter_f64(double* dst, const double* src, const int64_t* perm, int64_t n, mplak_stream_t s) {
  if (n < 0) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  if (dst == nullptr || src == nullptr || perm == nullptr || dst == src) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(s), dst, src, perm, n);
  return launch_status();
}

}  // extern "C"
