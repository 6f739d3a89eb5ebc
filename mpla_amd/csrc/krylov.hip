// Fused Krylov vector kernels with device-resident scalars.
//
// A CG iteration of the reference (one dgemv, three ddot, five daxpy and two
// memset, each dot a blocking device-to-host read: src/mpla.cpp:844-879) is
// done here as the caller's operator plus three kernels:
//   cg_dz      : partial <d,z>
//   cg_update  : alpha = rr/<d,z> (every block reduces the partials itself),
//                x += alpha d, r -= alpha z, partial <r,r>
//   cg_direct  : rr' = <r,r>, stopping test, d = r + (rr'/rr) d, state out
// All scalars stay on the device, so an iteration needs no host round trip and
// a chunk of iterations can be captured into a HIP graph.  The vector grid is
// a fixed function of n and partial sums are always combined in the same
// order, so every block recomputes identical alpha/beta values.

#include <math.h>

#include "common.hpp"

namespace mplak {
namespace {

constexpr int kVecMaxBlocks = 1024;

inline int vec_blocks(int64_t n) {
  int64_t b = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (b < 1) b = 1;
  if (b > kVecMaxBlocks) b = kVecMaxBlocks;
  return static_cast<int>(b);
}

__device__ __forceinline__ double stop_value(int rule, double rr, double bnorm2, double n_global) {
  return rule == MPLAK_STOP_RMS ? sqrt(rr / n_global) : sqrt(rr) / sqrt(bnorm2);
}

__global__ __launch_bounds__(kBlock) void cg_start_kernel(int64_t n, const double* __restrict__ b,
                                                          const double* __restrict__ z,
                                                          double* __restrict__ r,
                                                          double* __restrict__ d,
                                                          double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    const double ri = bi - z[i];
    r[i] = ri;
    d[i] = ri;
    rr = fma(ri, ri, rr);
    bb = fma(bi, bi, bb);
  }
  rr = block_sum<kBlock>(rr, scratch);
  bb = block_sum<kBlock>(bb, scratch);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rr;
    part[gridDim.x + blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(kBlock) void cg_start_final_kernel(const double* __restrict__ part,
                                                                int g, mplak_cg_state* st, int rule,
                                                                double eps, int iter_max,
                                                                double n_global) {
  __shared__ double scratch[kBlock / kWave];
  const double rr = reduce_partials<kBlock>(part, g, scratch);
  const double bb = reduce_partials<kBlock>(part + g, g, scratch);
  if (threadIdx.x == 0) {
    mplak_cg_state s;
    s.rr = rr;
    s.bnorm2 = bb;
    s.eps = eps;
    s.n_global = n_global;
    s.rel = stop_value(rule, rr, bb, n_global);
    s.alpha = 0.0;
    s.beta = 0.0;
    s.k = 0;
    s.iter_max = iter_max;
    s.rule = rule;
    // The loop runs only while rel >= eps (a NaN skips it, as the reference's
    // `if (rel >= eps) for (...)` does) and only for k = 1 .. iter_max-1.
    if (!(s.rel >= eps))
      s.done = MPLAK_CONVERGED;
    else if (iter_max <= 1)
      s.done = MPLAK_ITER_LIMIT;
    else
      s.done = MPLAK_RUNNING;
    *st = s;
  }
}

__global__ __launch_bounds__(kBlock) void cg_dz_kernel(int64_t n, const double* __restrict__ d,
                                                       const double* __restrict__ z,
                                                       const mplak_cg_state* __restrict__ in,
                                                       double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = fma(d[i], z[i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void cg_update_kernel(int64_t n, double* __restrict__ x,
                                                           double* __restrict__ r,
                                                           const double* __restrict__ d,
                                                           const double* __restrict__ z,
                                                           const mplak_cg_state* __restrict__ in,
                                                           double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int g = gridDim.x;
  const double dz = reduce_partials<kBlock>(part, g, scratch);
  const double alpha = in->rr / dz;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x[i] = fma(alpha, d[i], x[i]);
    const double ri = fma(-alpha, z[i], r[i]);
    r[i] = ri;
    rr = fma(ri, ri, rr);
  }
  rr = block_sum<kBlock>(rr, scratch);
  if (threadIdx.x == 0) part[g + blockIdx.x] = rr;
}

__global__ __launch_bounds__(kBlock) void cg_direct_kernel(int64_t n, const double* __restrict__ r,
                                                           double* __restrict__ d,
                                                           const mplak_cg_state* __restrict__ in,
                                                           mplak_cg_state* __restrict__ out,
                                                           const double* __restrict__ part) {
  const mplak_cg_state s_in = *in;
  if (s_in.done != MPLAK_RUNNING) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  const int g = gridDim.x;
  const double dz = reduce_partials<kBlock>(part, g, scratch);
  const double rr_new = reduce_partials<kBlock>(part + g, g, scratch);
  const double alpha = s_in.rr / dz;
  const double beta = rr_new / s_in.rr;
  const int k = s_in.k + 1;
  const double rel = stop_value(s_in.rule, rr_new, s_in.bnorm2, s_in.n_global);
  const bool conv = rel < s_in.eps;
  const bool limit = k >= s_in.iter_max - 1;
  if (!conv && !limit) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
      d[i] = fma(beta, d[i], r[i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    mplak_cg_state s = s_in;
    s.rr = rr_new;
    s.rel = rel;
    s.alpha = alpha;
    s.beta = beta;
    s.k = k;
    s.done = conv ? MPLAK_CONVERGED : (limit ? MPLAK_ITER_LIMIT : MPLAK_RUNNING);
    *out = s;
  }
}

__global__ __launch_bounds__(kBlock) void dot_into_kernel(int64_t n, const double* __restrict__ a,
                                                          double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = fma(a[i], a[i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void refresh_rr_final_kernel(const double* __restrict__ part,
                                                                  int g, mplak_cg_state* st) {
  __shared__ double scratch[kBlock / kWave];
  const double rr = reduce_partials<kBlock>(part, g, scratch);
  if (threadIdx.x == 0) st->rr = rr;
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

size_t mplak_cg_workspace_bytes(int64_t n) {
  return (size_t)2 * vec_blocks(n > 0 ? n : 1) * sizeof(double);
}

int mplak_cg_start(int64_t n, const double* b, const double* z, double* r, double* d,
                   mplak_cg_state* state, int rule, double eps, int32_t iter_max, double n_global,
                   void* work, size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !b || !z || !r || !d || !state) return MPLAK_ERR_ARG;
  if (rule != MPLAK_STOP_RMS && rule != MPLAK_STOP_RELATIVE) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_cg_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(cg_start_kernel, dim3(g), dim3(kBlock), 0, hs, n, b, z, r, d, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(cg_start_final_kernel, dim3(1), dim3(kBlock), 0, hs, part, g, state, rule, eps,
                     (int)iter_max, n_global);
  return launch_status();
}

int mplak_cg_iterate(int64_t n, double* x, double* r, double* d, const double* z,
                     const mplak_cg_state* in, mplak_cg_state* out, void* work, size_t work_bytes,
                     mplak_stream_t s) {
  if (n <= 0 || !x || !r || !d || !z || !in || !out || in == out) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_cg_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(cg_dz_kernel, dim3(g), dim3(kBlock), 0, hs, n, d, z, in, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(cg_update_kernel, dim3(g), dim3(kBlock), 0, hs, n, x, r, d, z, in, part);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(cg_direct_kernel, dim3(g), dim3(kBlock), 0, hs, n, r, d, in, out, part);
  return launch_status();
}

int mplak_cg_refresh_rr(int64_t n, const double* r, mplak_cg_state* state, void* work,
                        size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !r || !state) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_cg_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(dot_into_kernel, dim3(g), dim3(kBlock), 0, hs, n, r, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(refresh_rr_final_kernel, dim3(1), dim3(kBlock), 0, hs, part, g, state);
  return launch_status();
}

}  // extern "C"
