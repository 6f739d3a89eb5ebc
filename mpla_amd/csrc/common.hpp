// Shared device/host helpers for the gfx950 kernel library (libmplak).
//
// Everything here is stream-ordered and allocation-free: the C-ABI entry
// points receive every buffer they touch (inputs, outputs, workspace) from
// the caller and only enqueue kernels on the caller's stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/mplak.h"

namespace mplak {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kBlock = 256;        // default workgroup: 4 waves

// Status helpers -----------------------------------------------------------

inline int hip_status(hipError_t e) { return e == hipSuccess ? MPLAK_OK : MPLAK_ERR_HIP; }

inline int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? MPLAK_OK : MPLAK_ERR_HIP;
}

inline hipStream_t as_stream(mplak_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Deterministic reductions -------------------------------------------------
//
// Every reduction below combines values in an order that depends only on the
// launch shape, never on timing: wave reductions use a fixed xor-butterfly and
// block reductions sum the per-wave results in wave order.  Repeated calls on
// the same inputs therefore return bitwise-identical sums.

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Minimum / maximum over the wave (every lane gets the result).
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v = fmin(v, __shfl_xor(v, m));
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// Sum over a block of NT threads; the result is valid in every thread.
// `scratch` must hold NT / kWave elements of T and is reused after a barrier.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
  constexpr int NW = NT / kWave;
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) scratch[w] = v;
  __syncthreads();
  T s = scratch[0];
#pragma unroll
  for (int i = 1; i < NW; ++i) s += scratch[i];
  return s;
}

// This thread's share of a fixed-order sum of `count` partials: p[t],
// p[t + NT], p[t + 2 NT], ... added in that order from +0.0.  The loads go
// out kPartialBatch at a time (one round trip for up to kPartialBatch * NT
// partials instead of one per partial); a batch past the end adds +0.0,
// which leaves the sum as it is (a sum that starts at +0.0 is never -0.0 in
// round-to-nearest), so the bits are those of the one-by-one loop.
constexpr int kPartialBatch = 4;

template <int NT>
__device__ __forceinline__ double thread_partials(const double* __restrict__ p, int count) {
  double s = 0.0;
  for (int i0 = threadIdx.x; i0 < count; i0 += kPartialBatch * NT) {
    double v[kPartialBatch];
#pragma unroll
    for (int u = 0; u < kPartialBatch; ++u) {
      const int i = i0 + u * NT;
      v[u] = i < count ? p[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPartialBatch; ++u) s += v[u];
  }
  return s;
}

// The same for two partial arrays of equal length, their loads in flight
// together.
template <int NT>
__device__ __forceinline__ void thread_partials2(const double* __restrict__ p, const double* __restrict__ q,
                                                 int count, double& sp, double& sq) {
  sp = 0.0;
  sq = 0.0;
  for (int i0 = threadIdx.x; i0 < count; i0 += kPartialBatch * NT) {
    double vp[kPartialBatch], vq[kPartialBatch];
#pragma unroll
    for (int u = 0; u < kPartialBatch; ++u) {
      const int i = i0 + u * NT;
      vp[u] = i < count ? p[i] : 0.0;
      vq[u] = i < count ? q[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPartialBatch; ++u) {
      sp += vp[u];
      sq += vq[u];
    }
  }
}

// Sum of `count` partials laid out contiguously, evaluated by one block in a
// fixed order (strided per-thread sums, then block_sum).  Every block that
// calls this on the same data obtains the same bits.
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* __restrict__ p, int count,
                                                  double* scratch) {
  return block_sum<NT>(thread_partials<NT>(p, count), scratch);
}

// Both sums of two partial arrays of equal length (the bits of two
// reduce_partials calls).
template <int NT>
__device__ __forceinline__ void reduce_partials2(const double* __restrict__ p, const double* __restrict__ q,
                                                 int count, double* scratch, double& sp, double& sq) {
  double a, b;
  thread_partials2<NT>(p, q, count, a, b);
  sp = block_sum<NT>(a, scratch);
  sq = block_sum<NT>(b, scratch);
}

// Symmetric products -------------------------------------------------------

template <typename T>
struct Quad;
template <>
struct Quad<float> {
  typedef float4 type;
};
template <>
struct Quad<double> {
  typedef double4 type;
};

// Value of lane l + 1 (mod 64) into lane l (DPP wave_rol:1).  A column sum
// that a lane adds to at step s for source (l + s) mod 64 and then rotates
// ends, after 64 steps, in the lane of its source.
__device__ __forceinline__ float wave_rotate(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134 /* wave_rol:1 */, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_rotate(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Tile pair p = J (J + 1) / 2 + I, I <= J.
__device__ __forceinline__ int64_t sym_pair_index(int64_t I, int64_t J) { return J * (J + 1) / 2 + I; }

__device__ __forceinline__ void decode_pair(int64_t p, int64_t& I, int64_t& J) {
  J = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (J * (J + 1) / 2 > p) --J;
  while ((J + 1) * (J + 2) / 2 <= p) ++J;
  I = p - J * (J + 1) / 2;
}

}  // namespace mplak
