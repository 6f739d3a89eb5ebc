// Dense FP64 level-1/2 kernels and synthetic-input fills for gfx950.
//
//   dgemv  : y = A x, A column-major, HBM-streaming row panels with split-K
//   ddot   : deterministic two-pass reduction
//   daxpy / waxpby / fill / copy : grid-stride elementwise kernels
//   fill_uniform : counter-based SplitMix64 inputs (bitwise equal to the host)

#include "common.hpp"
#include "rng.hpp"

namespace mplak {
namespace {

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------

inline int elementwise_grid(int64_t n) {
  int64_t blocks = (n + kBlock * 4 - 1) / (kBlock * 4);
  if (blocks < 1) blocks = 1;
  if (blocks > 8192) blocks = 8192;
  return static_cast<int>(blocks);
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double* __restrict__ out, int64_t n,
                                                      double v) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = v;
}

__global__ __launch_bounds__(kBlock) void copy_kernel(double* __restrict__ dst,
                                                      const double* __restrict__ src, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void axpy_kernel(int64_t n, double alpha,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = fma(alpha, x[i], y[i]);
}

// w may alias x or y, so no __restrict__ here.
__global__ __launch_bounds__(kBlock) void waxpby_kernel(int64_t n, double a, const double* x,
                                                        double b, const double* y, double* w) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    w[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(kBlock) void uniform_kernel(double* __restrict__ out, int64_t n,
                                                         uint64_t key, uint64_t first, double lo,
                                                         double span) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = rng_uniform_ab(key, first + (uint64_t)i, lo, span);
}

// ---------------------------------------------------------------------------
// Deterministic dot product
// ---------------------------------------------------------------------------

constexpr int kDotMaxBlocks = 1024;

inline int dot_blocks(int64_t n) {
  int64_t b = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (b < 1) b = 1;
  if (b > kDotMaxBlocks) b = kDotMaxBlocks;
  return static_cast<int>(b);
}

__global__ __launch_bounds__(kBlock) void dot_partial_kernel(int64_t n,
                                                             const double* __restrict__ x,
                                                             const double* __restrict__ y,
                                                             double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = fma(x[i], y[i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void dot_final_kernel(const double* __restrict__ part,
                                                           int count, double* __restrict__ out) {
  __shared__ double scratch[kBlock / kWave];
  double s = reduce_partials<kBlock>(part, count, scratch);
  if (threadIdx.x == 0) *out = s;
}

// ---------------------------------------------------------------------------
// dgemv: y = A x, column-major A.
//
// A block of 4 waves owns a 128-row panel; lane l of every wave handles rows
// 2l, 2l+1 of the panel with one 16-byte load per column, so each wave reads
// 1 KiB contiguous per column.  The block's column slab is split into four
// contiguous quarters, one per wave; x[j] is wave-uniform and comes through
// the scalar cache.  The four wave sums are combined in LDS in wave order and
// written as one partial per (slab, row); a second kernel adds the slabs in
// slab order.  The slab count depends only on (m, n): about 256 workgroups
// (one per CU) with at least 256 columns per wave measured fastest from
// N = 4096 (cache-resident) to N = 131072 (HBM).
//
// A matrix larger than the 256 MiB Infinity Cache is read with nontemporal
// loads (STREAM): every element is used once, and the hint keeps the stream
// from churning L2 and the Infinity Cache.  At N = 65536 that raises the read
// rate from 6.26 to about 6.85 TB/s; a smaller matrix that stays cached
// between calls reads faster with ordinary loads (N = 4096: 29 us against
// 32 us).
// ---------------------------------------------------------------------------

constexpr int kGvRows = 128;
constexpr int kGvTargetBlocks = 256;  // ~one 4-wave workgroup per CU (swept on MI355X)
constexpr int64_t kGvStreamBytes = int64_t(256) << 20;  // Infinity Cache capacity

typedef double gv_f64x2 __attribute__((ext_vector_type(2)));

template <bool STREAM>
__device__ inline gv_f64x2 gv_load2(const double* p) {
  const gv_f64x2* q = reinterpret_cast<const gv_f64x2*>(p);
  if constexpr (STREAM) return __builtin_nontemporal_load(q);
  else return *q;
}

template <bool STREAM>
__device__ inline double gv_load1(const double* p) {
  if constexpr (STREAM) return __builtin_nontemporal_load(p);
  else return *p;
}

struct GemvPlan {
  int64_t row_tiles;
  int64_t slabs;
  int64_t slab_cols;
};

inline GemvPlan gemv_plan(int64_t m, int64_t n) {
  GemvPlan p;
  p.row_tiles = (m + kGvRows - 1) / kGvRows;
  int64_t ks = (kGvTargetBlocks + p.row_tiles - 1) / p.row_tiles;
  int64_t max_ks = n / 1024;  // keep at least 256 columns per wave
  if (max_ks < 1) max_ks = 1;
  if (ks > max_ks) ks = max_ks;
  if (ks < 1) ks = 1;
  if (ks > 65535) ks = 65535;
  p.slab_cols = (n + ks - 1) / ks;
  p.slabs = (n + p.slab_cols - 1) / p.slab_cols;
  return p;
}

template <bool VEC2, bool STREAM>
__global__ __launch_bounds__(kBlock) void dgemv_partial_kernel(
    int64_t m, int64_t n, const double* __restrict__ A, int64_t lda,
    const double* __restrict__ x, int64_t slab_cols, double* __restrict__ out,
    const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  __shared__ double red[kBlock / kWave][kGvRows];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t panel = (int64_t)blockIdx.x * kGvRows;
  const int64_t row = panel + 2 * lane;
  const int64_t c0 = (int64_t)blockIdx.y * slab_cols;
  const int64_t c1 = (c0 + slab_cols < n) ? c0 + slab_cols : n;
  const int64_t qlen = (slab_cols + 3) / 4;
  const int64_t j0 = (c0 + w * qlen < c1) ? c0 + w * qlen : c1;
  const int64_t j1 = (j0 + qlen < c1) ? j0 + qlen : c1;

  double acc0 = 0.0, acc1 = 0.0;
  if (row + 1 < m) {
    const double* a = A + row + j0 * lda;
    int64_t j = j0;
    if (VEC2) {
      for (; j + 8 <= j1; j += 8) {
        gv_f64x2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = gv_load2<STREAM>(a + u * lda);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double xj = x[j + u];
          acc0 = fma(v[u].x, xj, acc0);
          acc1 = fma(v[u].y, xj, acc1);
        }
        a += 8 * lda;
      }
      for (; j < j1; ++j) {
        const gv_f64x2 v = gv_load2<STREAM>(a);
        const double xj = x[j];
        acc0 = fma(v.x, xj, acc0);
        acc1 = fma(v.y, xj, acc1);
        a += lda;
      }
    } else {
      for (; j + 4 <= j1; j += 4) {
        double v0[4], v1[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          v0[u] = gv_load1<STREAM>(a + u * lda);
          v1[u] = gv_load1<STREAM>(a + u * lda + 1);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const double xj = x[j + u];
          acc0 = fma(v0[u], xj, acc0);
          acc1 = fma(v1[u], xj, acc1);
        }
        a += 4 * lda;
      }
      for (; j < j1; ++j) {
        const double xj = x[j];
        acc0 = fma(gv_load1<STREAM>(a), xj, acc0);
        acc1 = fma(gv_load1<STREAM>(a + 1), xj, acc1);
        a += lda;
      }
    }
  } else if (row < m) {
    const double* a = A + row + j0 * lda;
    for (int64_t j = j0; j < j1; ++j) {
      acc0 = fma(gv_load1<STREAM>(a), x[j], acc0);
      a += lda;
    }
  }
  red[w][2 * lane] = acc0;
  red[w][2 * lane + 1] = acc1;
  __syncthreads();
  if (threadIdx.x < kGvRows) {
    const int r = threadIdx.x;
    const int64_t gr = panel + r;
    if (gr < m) {
      double s = red[0][r];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) s += red[q][r];
      out[(int64_t)blockIdx.y * m + gr] = s;
    }
  }
}

__global__ __launch_bounds__(kBlock) void slab_sum_kernel(int64_t m, int64_t slabs,
                                                          const double* __restrict__ part,
                                                          double* __restrict__ y,
                                                          const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
    double s = part[i];
    for (int64_t k = 1; k < slabs; ++k) s += part[k * m + i];
    y[i] = s;
  }
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

const char* mplak_status_string(int status) {
  switch (status) {
    case MPLAK_OK: return "ok";
    case MPLAK_ERR_ARG: return "invalid argument";
    case MPLAK_ERR_HIP: return "HIP runtime error";
    case MPLAK_ERR_WORKSPACE: return "workspace too small";
    case MPLAK_ERR_UNSUPPORTED: return "unsupported request";
    case MPLAK_ERR_TOO_LARGE: return "problem needs more scratch than the no-allocation entry point holds";
    default: return "unknown status";
  }
}

int mplak_version(void) { return 100; }

int mplak_fill_uniform_f64(double* out, int64_t n, uint64_t seed, uint64_t stream,
                           uint64_t first_index, double lo, double hi, mplak_stream_t s) {
  if (n < 0 || (n > 0 && out == nullptr)) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  const uint64_t key = rng_key(seed, stream);
  hipLaunchKernelGGL(uniform_kernel, dim3(elementwise_grid(n)), dim3(kBlock), 0, as_stream(s), out,
                     n, key, first_index, lo, hi - lo);
  return launch_status();
}

int mplak_fill_f64(double* out, int64_t n, double value, mplak_stream_t s) {
  if (n < 0 || (n > 0 && out == nullptr)) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  hipLaunchKernelGGL(fill_kernel, dim3(elementwise_grid(n)), dim3(kBlock), 0, as_stream(s), out, n,
                     value);
  return launch_status();
}

int mplak_copy_f64(double* dst, const double* src, int64_t n, mplak_stream_t s) {
  if (n < 0 || (n > 0 && (dst == nullptr || src == nullptr))) return MPLAK_ERR_ARG;
  if (n == 0 || dst == src) return MPLAK_OK;
  hipLaunchKernelGGL(copy_kernel, dim3(elementwise_grid(n)), dim3(kBlock), 0, as_stream(s), dst,
                     src, n);
  return launch_status();
}

size_t mplak_dgemv_workspace_bytes(int64_t m, int64_t n) {
  if (m <= 0 || n <= 0) return 0;
  const GemvPlan p = gemv_plan(m, n);
  return p.slabs > 1 ? (size_t)p.slabs * (size_t)m * sizeof(double) : 0;
}

int mplak_dgemv_f64(int64_t m, int64_t n, const double* A, int64_t lda, const double* x, double* y,
                    void* work, size_t work_bytes, const int32_t* skip_if, mplak_stream_t s) {
  if (m < 0 || n < 0 || lda < (m > 1 ? m : 1)) return MPLAK_ERR_ARG;
  if (m == 0) return MPLAK_OK;
  if (y == nullptr) return MPLAK_ERR_ARG;
  if (n == 0) return mplak_fill_f64(y, m, 0.0, s);
  if (A == nullptr || x == nullptr) return MPLAK_ERR_ARG;
  const GemvPlan p = gemv_plan(m, n);
  const size_t need = mplak_dgemv_workspace_bytes(m, n);
  if (work_bytes < need || (need > 0 && work == nullptr)) return MPLAK_ERR_WORKSPACE;
  double* out = p.slabs > 1 ? static_cast<double*>(work) : y;
  const bool vec2 = (lda % 2 == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0);
  const dim3 grid((unsigned)p.row_tiles, (unsigned)p.slabs);
  const bool stream = (double)m * (double)n * sizeof(double) > (double)kGvStreamBytes;
  auto kern = vec2 ? (stream ? dgemv_partial_kernel<true, true> : dgemv_partial_kernel<true, false>)
                   : (stream ? dgemv_partial_kernel<false, true> : dgemv_partial_kernel<false, false>);
  hipLaunchKernelGGL(kern, grid, dim3(kBlock), 0, as_stream(s), m, n, A, lda, x, p.slab_cols, out,
                     skip_if);
  int st = launch_status();
  if (st != MPLAK_OK || p.slabs == 1) return st;
  hipLaunchKernelGGL(slab_sum_kernel, dim3(elementwise_grid(m)), dim3(kBlock), 0, as_stream(s), m,
                     p.slabs, out, y, skip_if);
  return launch_status();
}

size_t mplak_dot_workspace_bytes(int64_t n) {
  return (size_t)dot_blocks(n > 0 ? n : 1) * sizeof(double);
}

int mplak_ddot_f64(int64_t n, const double* x, const double* y, double* result, void* work,
                   size_t work_bytes, mplak_stream_t s) {
  if (n < 0 || result == nullptr || (n > 0 && (x == nullptr || y == nullptr))) return MPLAK_ERR_ARG;
  if (n == 0) return mplak_fill_f64(result, 1, 0.0, s);
  const int g = dot_blocks(n);
  if (work == nullptr || work_bytes < (size_t)g * sizeof(double)) return MPLAK_ERR_WORKSPACE;
  double* part = static_cast<double*>(work);
  hipLaunchKernelGGL(dot_partial_kernel, dim3(g), dim3(kBlock), 0, as_stream(s), n, x, y, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(dot_final_kernel, dim3(1), dim3(kBlock), 0, as_stream(s), part, g, result);
  return launch_status();
}

int mplak_daxpy_f64(int64_t n, double alpha, const double* x, double* y, mplak_stream_t s) {
  if (n < 0 || (n > 0 && (x == nullptr || y == nullptr))) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  hipLaunchKernelGGL(axpy_kernel, dim3(elementwise_grid(n)), dim3(kBlock), 0, as_stream(s), n,
                     alpha, x, y);
  return launch_status();
}

int mplak_dwaxpby_f64(int64_t n, double a, const double* x, double b, const double* y, double* w,
                      mplak_stream_t s) {
  if (n < 0 || (n > 0 && (x == nullptr || y == nullptr || w == nullptr))) return MPLAK_ERR_ARG;
  if (n == 0) return MPLAK_OK;
  hipLaunchKernelGGL(waxpby_kernel, dim3(elementwise_grid(n)), dim3(kBlock), 0, as_stream(s), n, a,
                     x, b, y, w);
  return launch_status();
}

}  // extern "C"
