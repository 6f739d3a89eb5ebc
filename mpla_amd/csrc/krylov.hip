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

#include <float.h>
#include <math.h>

#include "common.hpp"

namespace mplak {
namespace {

constexpr int kVecMaxBlocks = 1024;
// Elements per thread of the fused vector kernels: few, so that a vector of
// 64 Ki elements spreads over 128 workgroups (these kernels are latency-bound
// at solver sizes; the partial sums stay a few KB).
constexpr int kVecPerThread = 2;

inline int vec_blocks(int64_t n) {
  int64_t b = (n + kBlock * kVecPerThread - 1) / (kBlock * kVecPerThread);
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
  double rr, bb;
  reduce_partials2<kBlock>(part, part + g, g, scratch, rr, bb);
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
  double dz, rr_new;
  reduce_partials2<kBlock>(part, part + g, g, scratch, dz, rr_new);
  const double alpha = s_in.rr / dz;
  const double beta = rr_new / s_in.rr;
  const int k = s_in.k + 1;
  const double rel = stop_value(s_in.rule, rr_new, s_in.bnorm2, s_in.n_global);
  const bool conv = rel < s_in.eps;
  const bool limit = k >= s_in.iter_max - 1;
  // The reference breaks before the direction update only on convergence;
  // on its last iteration it still updates d (and may checkpoint it).
  if (!conv) {
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

// Single-reduction CG (Chronopoulos & Gear, J. Comput. Appl. Math. 25, 1989;
// the form of Ghysels & Vanroose, Parallel Comput. 40, 2014, Alg. 2).  The
// operator is applied to r, and the two inner products the iteration needs,
// gamma = <r,r> and delta = <r,Ar>, arrive as partial sums from the
// operator's own finish, so one kernel per iteration does the rest:
//   beta = gamma/gamma_old, alpha = gamma/(delta - beta gamma/alpha_old)
//   (first iteration: beta = 0, alpha = gamma/delta)
//   p = r + beta p, s = w + beta s, x += alpha p, r -= alpha s
// The stopping test is on gamma, the residual entering the iteration; on
// convergence nothing is updated.  State fields: rr = gamma, alpha and beta
// of the last update, k = updates done.
__global__ __launch_bounds__(kBlock) void cg1_update_kernel(int64_t n, double* __restrict__ x,
                                                            double* __restrict__ r, double* __restrict__ p,
                                                            double* __restrict__ s, const double* __restrict__ w,
                                                            const double* __restrict__ dots, int nd,
                                                            const mplak_cg_state* __restrict__ in,
                                                            mplak_cg_state* __restrict__ out) {
  const mplak_cg_state s_in = *in;
  if (s_in.done != MPLAK_RUNNING) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *out = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  // The vector loads of the usual launch shape (every element of this thread
  // in one pass) are issued before the partial sums are reduced, so the two
  // latency chains overlap; larger vectors finish in the strided loop.
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool first = s_in.k == 0;
  double rv[kVecPerThread], wv[kVecPerThread], pv[kVecPerThread], sv[kVecPerThread], xv[kVecPerThread];
#pragma unroll
  for (int u = 0; u < kVecPerThread; ++u) {
    const int64_t i = i0 + u * stride;
    const bool in_range = i < n;
    rv[u] = in_range ? r[i] : 0.0;
    wv[u] = in_range ? w[i] : 0.0;
    pv[u] = in_range && !first ? p[i] : 0.0;
    sv[u] = in_range && !first ? s[i] : 0.0;
    xv[u] = in_range ? x[i] : 0.0;
  }
  double gam, del;
  reduce_partials2<kBlock>(dots, dots + nd, nd, scratch, gam, del);
  const double rel = stop_value(s_in.rule, gam, s_in.bnorm2, s_in.n_global);
  const bool conv = rel < s_in.eps;
  const double beta = first ? 0.0 : gam / s_in.rr;
  const double alpha = first ? gam / del : gam / (del - beta * gam / s_in.alpha);
  // breakdown: <p, A p> = gamma/alpha is not positive (or not finite)
  const bool broke = !conv && !(alpha > 0.0 && alpha <= DBL_MAX);
  if (!conv && !broke) {
#pragma unroll
    for (int u = 0; u < kVecPerThread; ++u) {
      const int64_t i = i0 + u * stride;
      if (i >= n) break;
      const double pi = first ? rv[u] : fma(beta, pv[u], rv[u]);
      const double si = first ? wv[u] : fma(beta, sv[u], wv[u]);
      p[i] = pi;
      s[i] = si;
      x[i] = fma(alpha, pi, xv[u]);
      r[i] = fma(-alpha, si, rv[u]);
    }
    for (int64_t i = i0 + kVecPerThread * stride; i < n; i += stride) {
      const double ri = r[i], wi = w[i];
      const double pi = first ? ri : fma(beta, p[i], ri);
      const double si = first ? wi : fma(beta, s[i], wi);
      p[i] = pi;
      s[i] = si;
      x[i] = fma(alpha, pi, x[i]);
      r[i] = fma(-alpha, si, ri);
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    mplak_cg_state o = s_in;
    o.rr = gam;
    o.rel = rel;
    if (conv) {
      o.done = MPLAK_CONVERGED;
    } else if (broke) {
      o.done = MPLAK_BREAKDOWN;
    } else {
      o.alpha = alpha;
      o.beta = beta;
      o.k = s_in.k + 1;
      o.done = o.k >= s_in.iter_max - 1 ? MPLAK_ITER_LIMIT : MPLAK_RUNNING;
    }
    *out = o;
  }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (src/mpla.cpp:1003-1183).  Per iteration, with the caller's
// operator applications in between:
//   z  = A p ; bicg_alpha : alpha = rho/<z,r_hat>, s = r - alpha z
//   zz = A s ; bicg_omega : partials <zz,s>, <zz,zz>
//              bicg_update: omega, x += alpha p + omega s, r = s - omega zz,
//                           partial <r,r_hat> (and <r,r> for the cheap rule)
//              bicg_direct: beta, p = r + beta (p - omega z)
//                           [recursive rule: stopping test + state out]
//   z  = A x ; bicg_true  : zz = b - z, partial <zz,zz>
//              bicg_final : true relative residual, stopping test, state out
// Partial-sum slots: 0 <z,r_hat>, 1 <zz,s>, 2 <zz,zz>, 3 <r,r_hat>,
// 4 <r,r> or <b-Ax,b-Ax>; each slot holds one value per block.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void bicg_state_out(const mplak_bicgstab_state& s_in, double rho_new,
                                               double alpha, double omega, double beta, double rel,
                                               mplak_bicgstab_state* out) {
  mplak_bicgstab_state s = s_in;
  s.k = s_in.k + 1;
  s.rho = rho_new;
  s.alpha = alpha;
  s.omega = omega;
  s.beta = beta;
  s.rel = rel;
  const bool conv = rel < s_in.eps;
  const bool limit = s.k >= s_in.iter_max - 1;
  s.done = conv ? MPLAK_CONVERGED : (limit ? MPLAK_ITER_LIMIT : MPLAK_RUNNING);
  *out = s;
}

__global__ __launch_bounds__(kBlock) void bicg_start_kernel(int64_t n, const double* __restrict__ b,
                                                            const double* __restrict__ z,
                                                            double* __restrict__ r,
                                                            double* __restrict__ r_hat,
                                                            double* __restrict__ p,
                                                            double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double bi = b[i];
    const double ri = bi - z[i];
    r[i] = ri;
    r_hat[i] = ri;
    p[i] = ri;
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

__global__ __launch_bounds__(kBlock) void bicg_start_final_kernel(const double* __restrict__ part,
                                                                  int g, mplak_bicgstab_state* st,
                                                                  double eps, int iter_max,
                                                                  int true_residual) {
  __shared__ double scratch[kBlock / kWave];
  double rr, bb;
  reduce_partials2<kBlock>(part, part + g, g, scratch, rr, bb);
  if (threadIdx.x == 0) {
    mplak_bicgstab_state s;
    s.rho = rr;  // <r_hat, r> with r_hat = r
    s.bnorm2 = bb;
    s.rel = sqrt(rr) / sqrt(bb);
    s.eps = eps;
    s.alpha = s.omega = s.beta = 0.0;
    s.k = 0;
    s.iter_max = iter_max;
    s.true_residual = true_residual;
    if (!(s.rel >= eps))
      s.done = MPLAK_CONVERGED;
    else if (iter_max <= 1)
      s.done = MPLAK_ITER_LIMIT;
    else
      s.done = MPLAK_RUNNING;
    *st = s;
  }
}

template <int NDOT>
__device__ __forceinline__ void write_partials(double (&v)[NDOT], double* scratch, double* part,
                                               const int* slots) {
#pragma unroll
  for (int q = 0; q < NDOT; ++q) {
    const double s = block_sum<kBlock>(v[q], scratch);
    if (threadIdx.x == 0) part[slots[q] * gridDim.x + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void bicg_zr_kernel(int64_t n, const double* __restrict__ z,
                                                         const double* __restrict__ r_hat,
                                                         const mplak_bicgstab_state* __restrict__ in,
                                                         double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    v[0] = fma(z[i], r_hat[i], v[0]);
  const int slots[1] = {0};
  write_partials<1>(v, scratch, part, slots);
}

__global__ __launch_bounds__(kBlock) void bicg_alpha_kernel(int64_t n, const double* __restrict__ r,
                                                            const double* __restrict__ z,
                                                            double* __restrict__ s,
                                                            const mplak_bicgstab_state* __restrict__ in,
                                                            const double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const double alpha = in->rho / reduce_partials<kBlock>(part, gridDim.x, scratch);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s[i] = fma(-alpha, z[i], r[i]);
}

__global__ __launch_bounds__(kBlock) void bicg_omega_kernel(int64_t n, const double* __restrict__ zz,
                                                            const double* __restrict__ s,
                                                            const mplak_bicgstab_state* __restrict__ in,
                                                            double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double q = zz[i];
    v[0] = fma(q, s[i], v[0]);
    v[1] = fma(q, q, v[1]);
  }
  const int slots[2] = {1, 2};
  write_partials<2>(v, scratch, part, slots);
}

__global__ __launch_bounds__(kBlock) void bicg_update_kernel(
    int64_t n, double* __restrict__ x, double* __restrict__ r, const double* __restrict__ r_hat,
    const double* __restrict__ p, const double* __restrict__ s, const double* __restrict__ zz,
    const mplak_bicgstab_state* __restrict__ in, double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int g = gridDim.x;
  const double alpha = in->rho / reduce_partials<kBlock>(part, g, scratch);
  const double omega =
      reduce_partials<kBlock>(part + g, g, scratch) / reduce_partials<kBlock>(part + 2 * g, g, scratch);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double si = s[i];
    x[i] = fma(omega, si, fma(alpha, p[i], x[i]));
    const double ri = fma(-omega, zz[i], si);
    r[i] = ri;
    v[0] = fma(ri, r_hat[i], v[0]);
    v[1] = fma(ri, ri, v[1]);
  }
  const int slots[2] = {3, 4};
  write_partials<2>(v, scratch, part, slots);
}

__global__ __launch_bounds__(kBlock) void bicg_direct_kernel(
    int64_t n, const double* __restrict__ r, double* __restrict__ p, const double* __restrict__ z,
    const mplak_bicgstab_state* __restrict__ in, mplak_bicgstab_state* __restrict__ out,
    const double* __restrict__ part) {
  const mplak_bicgstab_state s_in = *in;
  if (s_in.done != MPLAK_RUNNING) {
    if (!s_in.true_residual && blockIdx.x == 0 && threadIdx.x == 0) *out = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  const int g = gridDim.x;
  const double alpha = s_in.rho / reduce_partials<kBlock>(part, g, scratch);
  const double omega =
      reduce_partials<kBlock>(part + g, g, scratch) / reduce_partials<kBlock>(part + 2 * g, g, scratch);
  const double rho_new = reduce_partials<kBlock>(part + 3 * g, g, scratch);
  const double beta = (alpha / omega) * (rho_new / s_in.rho);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = fma(beta, fma(-omega, z[i], p[i]), r[i]);
  if (!s_in.true_residual) {
    const double rr = reduce_partials<kBlock>(part + 4 * g, g, scratch);
    if (blockIdx.x == 0 && threadIdx.x == 0)
      bicg_state_out(s_in, rho_new, alpha, omega, beta, sqrt(rr) / sqrt(s_in.bnorm2), out);
  }
}

__global__ __launch_bounds__(kBlock) void bicg_true_kernel(int64_t n, const double* __restrict__ b,
                                                           const double* __restrict__ z,
                                                           double* __restrict__ zz,
                                                           const mplak_bicgstab_state* __restrict__ in,
                                                           double* __restrict__ part) {
  if (in->done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double q = b[i] - z[i];
    zz[i] = q;
    v[0] = fma(q, q, v[0]);
  }
  const int slots[1] = {4};
  write_partials<1>(v, scratch, part, slots);
}

__global__ __launch_bounds__(kBlock) void bicg_final_kernel(int g,
                                                            const mplak_bicgstab_state* __restrict__ in,
                                                            mplak_bicgstab_state* __restrict__ out,
                                                            const double* __restrict__ part) {
  const mplak_bicgstab_state s_in = *in;
  if (s_in.done != MPLAK_RUNNING) {
    if (threadIdx.x == 0) *out = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  const double alpha = s_in.rho / reduce_partials<kBlock>(part, g, scratch);
  const double omega =
      reduce_partials<kBlock>(part + g, g, scratch) / reduce_partials<kBlock>(part + 2 * g, g, scratch);
  const double rho_new = reduce_partials<kBlock>(part + 3 * g, g, scratch);
  const double beta = (alpha / omega) * (rho_new / s_in.rho);
  const double tt = reduce_partials<kBlock>(part + 4 * g, g, scratch);
  if (threadIdx.x == 0)
    bicg_state_out(s_in, rho_new, alpha, omega, beta, sqrt(tt) / sqrt(s_in.bnorm2), out);
}

__global__ __launch_bounds__(kBlock) void dot2_partial_kernel(int64_t n, const double* __restrict__ a,
                                                              const double* __restrict__ b,
                                                              double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = fma(a[i], b[i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void refresh_rho_final_kernel(const double* __restrict__ part,
                                                                   int g, mplak_bicgstab_state* st) {
  __shared__ double scratch[kBlock / kWave];
  const double v = reduce_partials<kBlock>(part, g, scratch);
  if (threadIdx.x == 0) st->rho = v;
}

__global__ void copy_state_kernel(const mplak_cg_state* __restrict__ src, mplak_cg_state* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}

__global__ void copy_bicg_state_kernel(const mplak_bicgstab_state* __restrict__ src,
                                       mplak_bicgstab_state* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = *src;
}


// ---------------------------------------------------------------------------
// Batched CG over nrhs columns (no reference counterpart): every column runs
// the recurrence of src/mpla.cpp:937-975 with its own scalars, all columns
// share one multi-RHS operator launch per iteration.  Grids are
// (vector blocks, columns); partial sums of column c live at part[c*G + g].
// A finished column is frozen (its kernels return at once); bcg_group sets
// all_done once every column has stopped, which lets the shared product skip.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void bcg_start_kernel(int64_t n, const double* __restrict__ B,
                                                           const double* __restrict__ Z, double* __restrict__ R,
                                                           double* __restrict__ D, double* __restrict__ part) {
  __shared__ double scratch[kBlock / kWave];
  const int64_t c = blockIdx.y;
  const int G = gridDim.x;
  const int64_t off = c * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double bi = B[off + i];
    const double ri = bi - Z[off + i];
    R[off + i] = ri;
    D[off + i] = ri;
    rr = fma(ri, ri, rr);
    bb = fma(bi, bi, bb);
  }
  rr = block_sum<kBlock>(rr, scratch);
  bb = block_sum<kBlock>(bb, scratch);
  if (threadIdx.x == 0) {
    part[(2 * c) * G + blockIdx.x] = rr;
    part[(2 * c + 1) * G + blockIdx.x] = bb;
  }
}

__global__ __launch_bounds__(kBlock) void bcg_start_final_kernel(const double* __restrict__ part, int G,
                                                                 mplak_cg_state* st, int rule, double eps,
                                                                 int iter_max, double n_global) {
  __shared__ double scratch[kBlock / kWave];
  const int c = blockIdx.x;
  const double rr = reduce_partials<kBlock>(part + (2 * c) * G, G, scratch);
  const double bb = reduce_partials<kBlock>(part + (2 * c + 1) * G, G, scratch);
  if (threadIdx.x == 0) {
    mplak_cg_state s;
    s.rr = rr;
    s.bnorm2 = bb;
    s.eps = eps;
    s.n_global = n_global;
    s.rel = stop_value(rule, rr, bb, n_global);
    s.alpha = s.beta = 0.0;
    s.k = 0;
    s.iter_max = iter_max;
    s.rule = rule;
    s.done = !(s.rel >= eps) ? MPLAK_CONVERGED : (iter_max <= 1 ? MPLAK_ITER_LIMIT : MPLAK_RUNNING);
    st[c] = s;
  }
}

__global__ __launch_bounds__(kBlock) void bcg_group_kernel(const mplak_cg_state* __restrict__ st, int nrhs,
                                                           mplak_bcg_group* __restrict__ group) {
  __shared__ int scratch[kBlock / kWave];
  int running = 0;
  for (int c = threadIdx.x; c < nrhs; c += kBlock) running += st[c].done == MPLAK_RUNNING ? 1 : 0;
  running = block_sum<kBlock>(running, scratch);
  if (threadIdx.x == 0) {
    group->active = running;
    group->all_done = running == 0 ? 1 : 0;
  }
}

__global__ __launch_bounds__(kBlock) void bcg_dz_kernel(int64_t n, const double* __restrict__ D,
                                                        const double* __restrict__ Z,
                                                        const mplak_cg_state* __restrict__ in,
                                                        double* __restrict__ part) {
  const int64_t c = blockIdx.y;
  if (in[c].done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t off = c * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s = fma(D[off + i], Z[off + i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) part[c * gridDim.x + blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void bcg_update_kernel(int64_t n, double* __restrict__ X,
                                                            double* __restrict__ R, const double* __restrict__ D,
                                                            const double* __restrict__ Z,
                                                            const mplak_cg_state* __restrict__ in, int64_t nrhs,
                                                            double* __restrict__ part) {
  const int64_t c = blockIdx.y;
  if (in[c].done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int G = gridDim.x;
  const double dz = reduce_partials<kBlock>(part + c * G, G, scratch);
  const double alpha = in[c].rr / dz;
  const int64_t off = c * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  double rr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    X[off + i] = fma(alpha, D[off + i], X[off + i]);
    const double ri = fma(-alpha, Z[off + i], R[off + i]);
    R[off + i] = ri;
    rr = fma(ri, ri, rr);
  }
  rr = block_sum<kBlock>(rr, scratch);
  if (threadIdx.x == 0) part[(nrhs + c) * G + blockIdx.x] = rr;
}

__global__ __launch_bounds__(kBlock) void bcg_direct_kernel(int64_t n, const double* __restrict__ R,
                                                            double* __restrict__ D,
                                                            const mplak_cg_state* __restrict__ in,
                                                            mplak_cg_state* __restrict__ out, int64_t nrhs,
                                                            const double* __restrict__ part) {
  const int64_t c = blockIdx.y;
  const mplak_cg_state s_in = in[c];
  if (s_in.done != MPLAK_RUNNING) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[c] = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  const int G = gridDim.x;
  const double dz = reduce_partials<kBlock>(part + c * G, G, scratch);
  const double rr_new = reduce_partials<kBlock>(part + (nrhs + c) * G, G, scratch);
  const double beta = rr_new / s_in.rr;
  const int k = s_in.k + 1;
  const double rel = stop_value(s_in.rule, rr_new, s_in.bnorm2, s_in.n_global);
  const bool conv = rel < s_in.eps;
  const bool limit = k >= s_in.iter_max - 1;
  if (!conv) {
    const int64_t off = c * n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
      D[off + i] = fma(beta, D[off + i], R[off + i]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    mplak_cg_state s = s_in;
    s.rr = rr_new;
    s.rel = rel;
    s.alpha = s_in.rr / dz;
    s.beta = beta;
    s.k = k;
    s.done = conv ? MPLAK_CONVERGED : (limit ? MPLAK_ITER_LIMIT : MPLAK_RUNNING);
    out[c] = s;
  }
}

__global__ void copy_states_kernel(const mplak_cg_state* __restrict__ src, mplak_cg_state* __restrict__ dst,
                                   int count) {
  for (int c = threadIdx.x; c < count; c += blockDim.x) dst[c] = src[c];
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

int mplak_cg1_iterate(int64_t n, double* x, double* r, double* p, double* sv, const double* w, const double* dots,
                      int64_t ndots, const mplak_cg_state* in, mplak_cg_state* out, mplak_stream_t s) {
  if (n <= 0 || !x || !r || !p || !sv || !w || !dots || !in || !out || in == out) return MPLAK_ERR_ARG;
  if (ndots <= 0 || ndots > (int64_t)1 << 20) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(cg1_update_kernel, dim3(vec_blocks(n)), dim3(kBlock), 0, as_stream(s), n, x, r, p, sv, w, dots,
                     (int)ndots, in, out);
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

extern "C" {

int mplak_cg_copy_state(const mplak_cg_state* src, mplak_cg_state* dst, mplak_stream_t s) {
  if (!src || !dst) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(copy_state_kernel, dim3(1), dim3(kWave), 0, as_stream(s), src, dst);
  return launch_status();
}

size_t mplak_bicgstab_workspace_bytes(int64_t n) {
  return (size_t)5 * vec_blocks(n > 0 ? n : 1) * sizeof(double);
}

int mplak_bicgstab_start(int64_t n, const double* b, const double* z, double* r, double* r_hat,
                         double* p, mplak_bicgstab_state* state, double eps, int32_t iter_max,
                         int32_t true_residual, void* work, size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !b || !z || !r || !r_hat || !p || !state) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bicgstab_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bicg_start_kernel, dim3(g), dim3(kBlock), 0, hs, n, b, z, r, r_hat, p, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bicg_start_final_kernel, dim3(1), dim3(kBlock), 0, hs, part, g, state, eps,
                     (int)iter_max, (int)(true_residual != 0));
  return launch_status();
}

int mplak_bicgstab_step_s(int64_t n, const double* r, const double* z, double* s_vec,
                          const double* r_hat, const mplak_bicgstab_state* in, void* work,
                          size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !r || !z || !s_vec || !r_hat || !in) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bicgstab_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bicg_zr_kernel, dim3(g), dim3(kBlock), 0, hs, n, z, r_hat, in, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bicg_alpha_kernel, dim3(g), dim3(kBlock), 0, hs, n, r, z, s_vec, in, part);
  return launch_status();
}

int mplak_bicgstab_step_x(int64_t n, double* x, double* r, const double* r_hat, double* p,
                          const double* s_vec, const double* z, const double* zz,
                          const mplak_bicgstab_state* in, mplak_bicgstab_state* out, void* work,
                          size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !x || !r || !r_hat || !p || !s_vec || !z || !zz || !in || !out || in == out)
    return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bicgstab_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bicg_omega_kernel, dim3(g), dim3(kBlock), 0, hs, n, zz, s_vec, in, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bicg_update_kernel, dim3(g), dim3(kBlock), 0, hs, n, x, r, r_hat, p, s_vec, zz, in,
                     part);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bicg_direct_kernel, dim3(g), dim3(kBlock), 0, hs, n, r, p, z, in, out, part);
  return launch_status();
}

int mplak_bicgstab_step_true(int64_t n, const double* b, const double* z, double* zz,
                             const mplak_bicgstab_state* in, mplak_bicgstab_state* out, void* work,
                             size_t work_bytes, mplak_stream_t s) {
  if (n <= 0 || !b || !z || !zz || !in || !out || in == out) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bicgstab_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bicg_true_kernel, dim3(g), dim3(kBlock), 0, hs, n, b, z, zz, in, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bicg_final_kernel, dim3(1), dim3(kBlock), 0, hs, g, in, out, part);
  return launch_status();
}

int mplak_bicgstab_refresh_rho(int64_t n, const double* r, const double* r_hat,
                               mplak_bicgstab_state* state, void* work, size_t work_bytes,
                               mplak_stream_t s) {
  if (n <= 0 || !r || !r_hat || !state) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bicgstab_workspace_bytes(n)) return MPLAK_ERR_WORKSPACE;
  const int g = vec_blocks(n);
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(dot2_partial_kernel, dim3(g), dim3(kBlock), 0, hs, n, r, r_hat, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(refresh_rho_final_kernel, dim3(1), dim3(kBlock), 0, hs, part, g, state);
  return launch_status();
}

int mplak_bicgstab_copy_state(const mplak_bicgstab_state* src, mplak_bicgstab_state* dst,
                              mplak_stream_t s) {
  if (!src || !dst) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(copy_bicg_state_kernel, dim3(1), dim3(kWave), 0, as_stream(s), src, dst);
  return launch_status();
}

}  // extern "C"

extern "C" {

size_t mplak_bcg_workspace_bytes(int64_t n, int64_t nrhs) {
  return (size_t)2 * vec_blocks(n > 0 ? n : 1) * (size_t)(nrhs > 0 ? nrhs : 1) * sizeof(double);
}

static int bcg_grid(int64_t n, int64_t nrhs, dim3* grid) {
  if (nrhs <= 0 || nrhs > 65535) return MPLAK_ERR_ARG;
  *grid = dim3(vec_blocks(n), (unsigned)nrhs);
  return MPLAK_OK;
}

int mplak_bcg_start(int64_t n, int64_t nrhs, const double* B, const double* Z, double* R, double* D,
                    mplak_cg_state* states, mplak_bcg_group* group, int rule, double eps, int32_t iter_max,
                    double n_global, void* work, size_t work_bytes, mplak_stream_t s) {
  dim3 grid;
  if (n <= 0 || !B || !Z || !R || !D || !states || !group || bcg_grid(n, nrhs, &grid) != MPLAK_OK)
    return MPLAK_ERR_ARG;
  if (rule != MPLAK_STOP_RMS && rule != MPLAK_STOP_RELATIVE) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bcg_workspace_bytes(n, nrhs)) return MPLAK_ERR_WORKSPACE;
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bcg_start_kernel, grid, dim3(kBlock), 0, hs, n, B, Z, R, D, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bcg_start_final_kernel, dim3((unsigned)nrhs), dim3(kBlock), 0, hs, part, (int)grid.x, states,
                     rule, eps, (int)iter_max, n_global);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bcg_group_kernel, dim3(1), dim3(kBlock), 0, hs, states, (int)nrhs, group);
  return launch_status();
}

int mplak_bcg_iterate(int64_t n, int64_t nrhs, double* X, double* R, double* D, const double* Z,
                      const mplak_cg_state* in, mplak_cg_state* out, mplak_bcg_group* group, void* work,
                      size_t work_bytes, mplak_stream_t s) {
  dim3 grid;
  if (n <= 0 || !X || !R || !D || !Z || !in || !out || !group || in == out || bcg_grid(n, nrhs, &grid) != MPLAK_OK)
    return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_bcg_workspace_bytes(n, nrhs)) return MPLAK_ERR_WORKSPACE;
  double* part = static_cast<double*>(work);
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(bcg_dz_kernel, grid, dim3(kBlock), 0, hs, n, D, Z, in, part);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bcg_update_kernel, grid, dim3(kBlock), 0, hs, n, X, R, D, Z, in, nrhs, part);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bcg_direct_kernel, grid, dim3(kBlock), 0, hs, n, R, D, in, out, nrhs, part);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(bcg_group_kernel, dim3(1), dim3(kBlock), 0, hs, out, (int)nrhs, group);
  return launch_status();
}

int mplak_bcg_copy_states(const mplak_cg_state* src, mplak_cg_state* dst, int64_t nrhs, mplak_bcg_group* group,
                          mplak_stream_t s) {
  if (!src || !dst || nrhs <= 0 || nrhs > 65535) return MPLAK_ERR_ARG;
  hipStream_t hs = as_stream(s);
  hipLaunchKernelGGL(copy_states_kernel, dim3(1), dim3(kBlock), 0, hs, src, dst, (int)nrhs);
  int st = launch_status();
  if (st != MPLAK_OK || group == nullptr) return st;
  hipLaunchKernelGGL(bcg_group_kernel, dim3(1), dim3(kBlock), 0, hs, dst, (int)nrhs, group);
  return launch_status();
}

}  // extern "C"
