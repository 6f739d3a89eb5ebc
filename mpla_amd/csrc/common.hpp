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

// Sum of `count` partials laid out contiguously, evaluated by one block in a
// fixed order (strided per-thread sums, then block_sum).  Every block that
// calls this on the same data obtains the same bits.
template <int NT>
__device__ __forceinline__ double reduce_partials(const double* __restrict__ p, int count,
                                                  double* scratch) {
  double s = 0.0;
  for (int i = threadIdx.x; i < count; i += NT) s += p[i];
  return block_sum<NT>(s, scratch);
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
