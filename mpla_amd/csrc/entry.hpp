// Kernel-matrix entries shared by every matrix-free kernel (single-vector,
// multi-RHS, ragged).
//
// Coordinates are scaled once so that an entry is a function of a plain sum
// of squares s of scaled coordinate differences:
//   Gaussian  exp(-|d|^2 / (2 l^2)) = 2^-s,   s = |d|^2 log2(e) / (2 l^2)
//   Matern32  (1 + r) exp(-r),  r = sqrt(s),  s = 3 |d|^2 / l^2
// FP32 evaluates 2^-x with the hardware v_exp_f32.  FP64 uses exp2_neg below:
// a table-driven exponential (P. T. P. Tang, "Table-driven implementation of
// the exponential function in IEEE floating-point arithmetic", ACM TOMS 15,
// 1989) specialised to non-negative arguments, with a 256-entry table of
// 2^(j/256) that each workgroup builds in LDS and a degree-4 Taylor polynomial
// of 2^f on |f| <= 1/512 (truncation < 4e-17 relative; a few ulp overall).
#pragma once

#include "common.hpp"

namespace mplak {

constexpr double kLog2e = 1.4426950408889634074;
constexpr int kExp2Table = 256;

// Scale applied to every coordinate (both precisions).
inline double coord_scale(int kind, double l) {
  if (kind == MPLAK_KERNEL_GAUSSIAN) return sqrt(kLog2e) / (sqrt(2.0) * l);
  return sqrt(3.0) / l;
}

// Squared scaled distance (in coord_scale coordinates) beyond which every
// entry is below 2^-drop_log2; the entry at distance 0 is 1.
inline double drop_distance2(int kind, double drop_log2) {
  if (kind == MPLAK_KERNEL_GAUSSIAN) return drop_log2;  // entry 2^-s
  // (1 + r) exp(-r) is decreasing for r > 0: bisection for (1 + r) exp(-r) = 2^-drop_log2
  const double target = -drop_log2 * 0.69314718055994530942;
  double lo = 0.0, hi = 1.0;
  while (log1p(hi) - hi > target) hi *= 2.0;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (lo + hi);
    if (log1p(mid) - mid > target)
      lo = mid;
    else
      hi = mid;
  }
  return hi * hi;
}

// Fills tab[j] = 2^(j/256) (j < 256) with the threads of the calling block;
// the caller synchronises the block before the first use.
__device__ __forceinline__ void fill_exp2_table(double* tab) {
  for (int j = threadIdx.x; j < kExp2Table; j += blockDim.x) tab[j] = exp2((double)j / kExp2Table);
}

// 2^-s for s >= 0 (+inf and s > 1100 give 0).  -s = k/256 + f with
// k = round(-256 s) and |f| <= 1/512 (both exact), 2^(k/256) = 2^floor(k/256)
// * tab[k mod 256], 2^f = sum_{n<=4} (f ln 2)^n / n!.
__device__ __forceinline__ double exp2_neg(double s, const double* __restrict__ tab) {
  s = fmin(s, 1100.0);
  const double k = rint(s * -256.0);
  const double f = fma(k, -1.0 / 256.0, -s);
  const int ki = (int)k;
  // Taylor coefficients ln(2)^n / n!
  double p = fma(f, 0.0096181291076284772, 0.055504108664821580);
  p = fma(f, p, 0.24022650695910071);
  p = fma(f, p, 0.69314718055994531);
  p = fma(f, p, 1.0);
  return __builtin_amdgcn_ldexp(tab[ki & (kExp2Table - 1)] * p, ki >> 8);  // one v_ldexp_f64, any exponent
}

template <typename T, int KIND>
__device__ __forceinline__ T kernel_entry(T s, const double* tab);

template <>
__device__ __forceinline__ double kernel_entry<double, MPLAK_KERNEL_GAUSSIAN>(double s, const double* tab) {
  return exp2_neg(s, tab);
}
template <>
__device__ __forceinline__ double kernel_entry<double, MPLAK_KERNEL_MATERN32>(double s, const double* tab) {
  const double r = sqrt(s);
  const double e = exp2_neg(r * kLog2e, tab);
  return fma(r, e, e);
}
template <>
__device__ __forceinline__ float kernel_entry<float, MPLAK_KERNEL_GAUSSIAN>(float s, const double*) {
  return __builtin_amdgcn_exp2f(-s);
}
template <>
__device__ __forceinline__ float kernel_entry<float, MPLAK_KERNEL_MATERN32>(float s, const double*) {
  const float r = __builtin_amdgcn_sqrtf(s);
  const float e = __builtin_amdgcn_exp2f(-r * (float)kLog2e);
  return fmaf(r, e, e);
}

// Runtime kind (ragged kernels, one launch for many systems).
template <typename T>
__device__ __forceinline__ T kernel_entry_rt(int kind, T s, const double* tab) {
  return kind == MPLAK_KERNEL_GAUSSIAN ? kernel_entry<T, MPLAK_KERNEL_GAUSSIAN>(s, tab)
                                       : kernel_entry<T, MPLAK_KERNEL_MATERN32>(s, tab);
}

// LDS for the table where T needs it (one element otherwise).
template <typename T>
struct Exp2Lds {
  static constexpr int size = sizeof(T) == 8 ? kExp2Table : 1;
};

}  // namespace mplak
