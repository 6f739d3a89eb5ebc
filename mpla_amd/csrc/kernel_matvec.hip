// Matrix-free kernel-matrix products for gfx950.
//
// y_i = sum_j k(|p_i - p_j|) x_j (+ nugget x_i on the global diagonal,
// + coincident_shift x_j where |p_i - p_j| < 1e-14), entries never stored.
//
// Tiling: a 256-thread workgroup owns a tile of 512 target points (two per
// thread, held in registers) and one slab of source columns.  The slab is
// streamed through LDS in chunks of 256 points (coordinates pre-scaled so the
// kernel argument is a plain sum of squares) together with the matching x
// values; every lane reads the same LDS word at a time (broadcast, no bank
// conflicts).  Slab partial sums go to a workspace and a second kernel adds
// them in slab order, so the result is deterministic for a given shape.
//
// Entry cost is VALU-bound: FP64 pays a software exp (~20 FP64 ops); FP32 uses
// the hardware v_exp_f32 / v_sqrt_f32 and folds each chunk's FP32 partial sum
// into an FP64 accumulator.

#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <type_traits>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "entry.hpp"

namespace mplak {
namespace {

constexpr int kKmvTpt = 2;                     // target points per thread
constexpr int kKmvTile = kBlock * kKmvTpt;     // 512 targets per workgroup
constexpr int kKmvChunk = kBlock;              // sources per LDS chunk
constexpr int kKmvMaxDim = 16;
constexpr int kKmvTargetBlocks = 4096;

struct KmvPlan {
  int64_t row_tiles, slabs, slab_cols;
};

inline KmvPlan kmv_plan(int64_t rows, int64_t cols) {
  KmvPlan p;
  p.row_tiles = (rows + kKmvTile - 1) / kKmvTile;
  int64_t ks = (kKmvTargetBlocks + p.row_tiles - 1) / p.row_tiles;
  int64_t max_ks = (cols + kKmvChunk - 1) / kKmvChunk;  // at least one chunk per slab
  if (ks > max_ks) ks = max_ks;
  if (ks < 1) ks = 1;
  if (ks > 65535) ks = 65535;
  // slab width is a whole number of chunks
  int64_t chunks = (cols + kKmvChunk - 1) / kKmvChunk;
  int64_t chunks_per_slab = (chunks + ks - 1) / ks;
  p.slab_cols = chunks_per_slab * kKmvChunk;
  p.slabs = (cols + p.slab_cols - 1) / p.slab_cols;
  return p;
}

// DIM > 0: compile-time dimension; DIM == 0: runtime dim <= kKmvMaxDim.
template <typename T, int KIND, int DIM, bool COINC>
__global__ __launch_bounds__(kBlock) void kmv_partial_kernel(
    const double* const* __restrict__ points, int dim_rt, int64_t row_off, int64_t rows,
    int64_t col_off, int64_t cols, int64_t slab_cols, double scale, double coinc_thr,
    double coinc_shift, const double* __restrict__ x, double* __restrict__ out,
    const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  constexpr int MD = DIM > 0 ? DIM : kKmvMaxDim;
  const int dim = DIM > 0 ? DIM : dim_rt;
  __shared__ T lp[MD][kKmvChunk];
  __shared__ T lx[kKmvChunk];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the chunk barrier below

  const int tid = threadIdx.x;
  const int64_t tile0 = (int64_t)blockIdx.x * kKmvTile;
  const T sc = (T)scale;
  const T thr = (T)coinc_thr;
  const T shift = (T)coinc_shift;

  // Target coordinates in registers (rows past the end are computed on a
  // copy of the last row and never stored).
  T pi[kKmvTpt][MD];
#pragma unroll
  for (int t = 0; t < kKmvTpt; ++t) {
    int64_t i = tile0 + tid + t * kBlock;
    if (i >= rows) i = rows - 1;
#pragma unroll
    for (int d = 0; d < MD; ++d) pi[t][d] = d < dim ? (T)(points[d][row_off + i] * scale) : T(0);
  }

  const int64_t c0 = (int64_t)blockIdx.y * slab_cols;
  const int64_t c1 = (c0 + slab_cols < cols) ? c0 + slab_cols : cols;

  double acc[kKmvTpt];
#pragma unroll
  for (int t = 0; t < kKmvTpt; ++t) acc[t] = 0.0;

  for (int64_t cb = c0; cb < c1; cb += kKmvChunk) {
    __syncthreads();
    {
      const int64_t j = cb + tid;
      if (j < c1) {
#pragma unroll
        for (int d = 0; d < MD; ++d) lp[d][tid] = d < dim ? (T)(points[d][col_off + j] * scale) : T(0);
        lx[tid] = (T)x[j];
      } else {
#pragma unroll
        for (int d = 0; d < MD; ++d) lp[d][tid] = T(0);
        lx[tid] = T(0);
      }
    }
    __syncthreads();
    T cacc[kKmvTpt];
#pragma unroll
    for (int t = 0; t < kKmvTpt; ++t) cacc[t] = T(0);
#pragma unroll 4
    for (int jj = 0; jj < kKmvChunk; ++jj) {
      T pj[MD];
#pragma unroll
      for (int d = 0; d < MD; ++d) pj[d] = d < dim ? lp[d][jj] : T(0);
      const T xj = lx[jj];
#pragma unroll
      for (int t = 0; t < kKmvTpt; ++t) {
        T s = T(0);
#pragma unroll
        for (int d = 0; d < MD; ++d) {
          if (d < dim) {
            const T df = pi[t][d] - pj[d];
            s = fma(df, df, s);
          }
        }
        T k = kernel_entry<T, KIND>(s, exp_tab);
        if (COINC) k += (s < thr) ? shift : T(0);
        cacc[t] = fma(k, xj, cacc[t]);
      }
    }
#pragma unroll
    for (int t = 0; t < kKmvTpt; ++t) acc[t] += (double)cacc[t];
  }
  (void)sc;

#pragma unroll
  for (int t = 0; t < kKmvTpt; ++t) {
    const int64_t i = tile0 + tid + t * kBlock;
    if (i < rows) out[(int64_t)blockIdx.y * rows + i] = acc[t];
  }
}

// y_i = sum over slabs (slab order) + nugget * x at the global diagonal.
__global__ __launch_bounds__(kBlock) void kmv_finish_kernel(
    int64_t rows, int64_t slabs, const double* __restrict__ part, int64_t row_off,
    int64_t col_off, int64_t cols, double nugget, const double* __restrict__ x,
    double* __restrict__ y, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += stride) {
    double s = part[i];
    for (int64_t k = 1; k < slabs; ++k) s += part[k * rows + i];
    const int64_t jd = row_off + i - col_off;
    if (nugget != 0.0 && jd >= 0 && jd < cols) s = fma(nugget, x[jd], s);
    y[i] = s;
  }
}

__global__ __launch_bounds__(kBlock) void gaussian_rhs_kernel(const double* const* __restrict__ points,
                                                              int dim, int64_t row_off, int64_t rows,
                                                              double* __restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double half_sqrt_pi = 0.88622692545275801365;  // sqrt(pi) / 2
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows; i += stride) {
    double v = 1.0;
    for (int d = 0; d < dim; ++d) {
      const double p = points[d][row_off + i];
      v = v * half_sqrt_pi * (erf(1.0 - p) - erf(-p));
    }
    b[i] = v;
  }
}

template <typename T, int KIND, int DIM>
void launch_partial(dim3 grid, hipStream_t st, bool coinc, const mplak_kernel_desc* d,
                    int64_t slab_cols, double scale, double thr, const double* x, double* out,
                    const int32_t* skip) {
  if (coinc)
    hipLaunchKernelGGL((kmv_partial_kernel<T, KIND, DIM, true>), grid, dim3(kBlock), 0, st,
                       d->points, d->dim, d->row_offset, d->row_count, d->col_offset, d->col_count,
                       slab_cols, scale, thr, d->coincident_shift, x, out, skip);
  else
    hipLaunchKernelGGL((kmv_partial_kernel<T, KIND, DIM, false>), grid, dim3(kBlock), 0, st,
                       d->points, d->dim, d->row_offset, d->row_count, d->col_offset, d->col_count,
                       slab_cols, scale, thr, d->coincident_shift, x, out, skip);
}

template <typename T, int KIND>
void dispatch_dim(dim3 grid, hipStream_t st, bool coinc, const mplak_kernel_desc* d,
                  int64_t slab_cols, double scale, double thr, const double* x, double* out,
                  const int32_t* skip) {
  switch (d->dim) {
    case 1: launch_partial<T, KIND, 1>(grid, st, coinc, d, slab_cols, scale, thr, x, out, skip); break;
    case 2: launch_partial<T, KIND, 2>(grid, st, coinc, d, slab_cols, scale, thr, x, out, skip); break;
    case 3: launch_partial<T, KIND, 3>(grid, st, coinc, d, slab_cols, scale, thr, x, out, skip); break;
    case 4: launch_partial<T, KIND, 4>(grid, st, coinc, d, slab_cols, scale, thr, x, out, skip); break;
    default: launch_partial<T, KIND, 0>(grid, st, coinc, d, slab_cols, scale, thr, x, out, skip); break;
  }
}


// ---------------------------------------------------------------------------
// Symmetric product for a square block whose rows and columns are the same
// points (row range == column range): K_ij = K_ji bitwise, so every
// off-diagonal entry is evaluated once and used for both y_i += K_ij x_j and
// y_j += K_ij x_i.
//
// Workgroup p owns the tile pair (I, J), I <= J, tiles of 512 points
// (256 threads x 2 targets).  The I-side sum is the plain loop over the 512
// sources of J.  For the J-side, wave w handles its 128 targets against the
// sources in 8 groups of 64: at step s lane l takes source (l + s) mod 64 of
// the group (from LDS, conflict-free) and adds its two entries times x_i to a
// register t that then moves one lane down the wave (DPP wave_rol:1), so t
// always belongs to the source the lane works on next; after 64 steps lane l
// holds the whole wave's sum for source l.  The four waves' sums are added in
// wave order.  Diagonal pairs (I == J) use the plain loop only.  Partials
// go to part_i[p][512] and part_j[p][512]; the finish kernel adds, for every
// row, the partials of its 512-tile in partner order.  All orders are fixed,
// so the product is deterministic.
//
// In FP32 each group of 64 sources is stored twice in a row in LDS, so the
// rotated read (l + s) mod 64 becomes group[l + s], an immediate offset from
// one per-lane address (no index arithmetic in the loop).  In FP64 the
// exponential dominates and the group is stored once.
// ---------------------------------------------------------------------------

constexpr int kSymTile = 2 * kBlock;  // 512 points per tile
constexpr int kSymMinN = 12288;       // crossover measured on MI355X (plain kernel faster below)

// MPLAK_KMV_SYMMETRIC=0 turns the symmetric product off, =1 uses it for every
// eligible block regardless of size; unset: blocks of at least kSymMinN points.
inline int sym_mode() {
  const char* v = getenv("MPLAK_KMV_SYMMETRIC");
  if (v == nullptr || *v == '\0') return -1;
  return atoi(v) != 0 ? 1 : 0;
}

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

__device__ __forceinline__ float wave_rotate(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134 /* wave_rol:1 */, 0xF, 0xF, false));
}
__device__ __forceinline__ double wave_rotate(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int64_t sym_pair_index(int64_t I, int64_t J) { return J * (J + 1) / 2 + I; }

// Tile pair p = J (J + 1) / 2 + I, I <= J.
__device__ __forceinline__ void decode_pair(int64_t p, int64_t& I, int64_t& J) {
  J = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (J * (J + 1) / 2 > p) --J;
  while ((J + 1) * (J + 2) / 2 <= p) ++J;
  I = p - J * (J + 1) / 2;
}

// Squared distance between two axis-aligned boxes (lo at [0..2], hi at [3..5]).
template <typename T, int DIM>
__device__ __forceinline__ T box_gap2(const T* a, const T* b) {
  T d2 = T(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T gap = fmax(T(0), fmax(b[d] - a[3 + d], a[d] - b[3 + d]));
    d2 = fma(gap, gap, d2);
  }
  return d2;
}

enum SymMode { kSymExact = 0, kSymScreened = 1, kSymPlanned = 2 };

// kSymScreened: a wave skips a 64-source group when the bounding boxes of its
// 128 targets and of the group are further apart (squared, in the scaled
// coordinates) than drop_s, i.e. when every entry of that 128 x 64 block is
// below the caller's drop level.  The test uses wave-reduced boxes, so the
// branch is uniform across the wave.
// kSymPlanned: as kSymScreened, over the tile pairs list[0 .. *count) of a
// screening plan (pairs with at least one kept block, heaviest first); the
// partials of the other pairs were zeroed when the plan was built.
template <typename T, int KIND, int DIM, bool COINC, int MODE>
__global__ __launch_bounds__(kBlock) void kmv_sym_kernel(
    const double* const* __restrict__ points, int64_t off, int64_t n, double scale, double coinc_thr,
    double coinc_shift, double drop_s, const double* __restrict__ x, double* __restrict__ part_i,
    double* __restrict__ part_j, const uint32_t* __restrict__ list, const int64_t* __restrict__ count,
    const int32_t* __restrict__ skip) {
  constexpr bool SCREEN = MODE != kSymExact;
  // CENTER: on compact target tiles (spatially ordered points) the screened
  // FP32 product evaluates |p - q|^2 as n_p + n_q - 2 p.q in coordinates
  // centred on the target tile (n = squared norm; one add and DIM FMAs
  // instead of 2 DIM subtractions and DIM products).  The tile-centred norms
  // are small, so the cancellation error stays at the level of the FP32
  // coordinates themselves; tiles wider than kCenterRadius2 (unordered
  // points) and the coincidence test (which needs s == 0 exactly) use the
  // difference form.
  constexpr bool CENTER = SCREEN && !COINC && sizeof(T) == 4;
  constexpr float kCenterRadius2 = 49.0f;
  if (skip != nullptr && *skip != 0) return;
  if (MODE == kSymPlanned && (int64_t)blockIdx.x >= *count) return;
  typedef typename Quad<T>::type Q;
  constexpr int kDup = sizeof(T) == 4 ? 2 : 1;  // copies of each 64-source group
  constexpr int kGroups = kSymTile / kWave;
  // Column-side sums of the four waves, red[w * kSymTile + k].  In FP32 each
  // wave holds its 8 group sums in registers and they reuse the source buffer
  // once every wave is done with it (16 KB of LDS per workgroup, 8 resident
  // workgroups per CU); in FP64 registers are the scarcer resource, and the
  // sums go to a buffer of their own as each group completes.
  constexpr bool kHoldSums = sizeof(T) == 4;
  static_assert(sizeof(Q) * kGroups * kDup * kWave >= sizeof(T) * (kBlock / kWave) * kSymTile, "LDS reuse");
  __shared__ Q src[kGroups][kDup * kWave];
  __shared__ T gbox[kGroups][6];
  __shared__ T own_red[kHoldSums ? 1 : (kBlock / kWave) * kSymTile];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the staging barrier
  __shared__ T nsrc[CENTER ? kGroups : 1][CENTER ? kDup * kWave : 1];  // centred source norms
  __shared__ T ibox[CENTER ? kBlock / kWave : 1][6];                 // target-tile box per wave
  T* red = kHoldSums ? reinterpret_cast<T*>(&src[0][0]) : own_red;

  const int64_t p = MODE == kSymPlanned ? (int64_t)list[blockIdx.x] : (int64_t)blockIdx.x;
  int64_t I, J;
  decode_pair(p, I, J);

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const T thr = (T)coinc_thr, shift = (T)coinc_shift;

  // two targets per lane: tile offsets w*128 + lane and w*128 + 64 + lane
  T pi[2][DIM], xi[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = I * kSymTile + w * 2 * kWave + h * kWave + lane;
    const int64_t ic = i < n ? i : n - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) pi[h][d] = (T)(points[d][off + ic] * scale);
    xi[h] = i < n ? (T)x[i] : T(0);
  }
  T tbox[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (SCREEN) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      tbox[d] = wave_min(fmin(pi[0][d], pi[1][d]));
      tbox[3 + d] = wave_max(fmax(pi[0][d], pi[1][d]));
    }
  }
  // centring: the target tile's box from the four wave boxes
  T ctr[3] = {T(0), T(0), T(0)};
  bool centered = false;
  if (CENTER) {
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 6; ++q) ibox[w][q] = tbox[q];
    }
    __syncthreads();
    T r2 = T(0);
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      T lo = ibox[0][d], hi = ibox[0][3 + d];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) {
        lo = fmin(lo, ibox[q][d]);
        hi = fmax(hi, ibox[q][3 + d]);
      }
      ctr[d] = T(0.5) * (lo + hi);
      const T half = T(0.5) * (hi - lo);
      r2 = fma(half, half, r2);
    }
    centered = __builtin_amdgcn_readfirstlane((int)(r2 <= (T)kCenterRadius2)) != 0;
  }
  for (int k = tid; k < kSymTile; k += kBlock) {
    const int64_t j = J * kSymTile + k;
    const int64_t jc = j < n ? j : n - 1;
    Q q;
    q.x = (T)(points[0][off + jc] * scale);
    q.y = DIM > 1 ? (T)(points[1][off + jc] * scale) : T(0);
    q.z = DIM > 2 ? (T)(points[2][off + jc] * scale) : T(0);
    q.w = j < n ? (T)x[j] : T(0);
    if (SCREEN) {  // k / kWave is uniform across the wave: one group per wave and pass
      const T lo0 = wave_min(q.x), lo1 = wave_min(q.y), lo2 = wave_min(q.z);
      const T hi0 = wave_max(q.x), hi1 = wave_max(q.y), hi2 = wave_max(q.z);
      if (lane == 0) {
        T* b = gbox[k / kWave];
        b[0] = lo0, b[1] = lo1, b[2] = lo2, b[3] = hi0, b[4] = hi1, b[5] = hi2;
      }
    }
    if (CENTER && centered) {
      q.x -= ctr[0];
      if (DIM > 1) q.y -= ctr[1];
      if (DIM > 2) q.z -= ctr[2];
      T nq = q.x * q.x;
      if (DIM > 1) nq = fma(q.y, q.y, nq);
      if (DIM > 2) nq = fma(q.z, q.z, nq);
#pragma unroll
      for (int c = 0; c < kDup; ++c) nsrc[k / kWave][c * kWave + (k & (kWave - 1))] = nq;
    }
#pragma unroll
    for (int c = 0; c < kDup; ++c) src[k / kWave][c * kWave + (k & (kWave - 1))] = q;
  }
  // centred targets: pa = -2 (p - c), np = |p - c|^2
  T pa[2][DIM], np_[2] = {T(0), T(0)};
  if (CENTER && centered) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const T v = pi[h][d] - ctr[d];
        pa[h][d] = T(-2) * v;
        np_[h] = d == 0 ? v * v : fma(v, v, np_[h]);
      }
    }
  }
  // true when every entry between this wave's targets and group g is dropped
  auto far_group = [&](int g) -> bool {
    if (!SCREEN) return false;
    const T d2 = box_gap2<T, DIM>(tbox, gbox[g]);
    return __builtin_amdgcn_readfirstlane((int)(d2 > (T)drop_s)) != 0;
  };
  __syncthreads();

  T accI[2] = {T(0), T(0)};
  // centred form: s = (n_p + n_q) + pa . q
  auto entry2c = [&](const Q& q, T nq, T (&e)[2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      T s = np_[h] + nq;
      s = fma(pa[h][0], q.x, s);
      if (DIM > 1) s = fma(pa[h][1], q.y, s);
      if (DIM > 2) s = fma(pa[h][2], q.z, s);
      if (KIND == MPLAK_KERNEL_MATERN32) s = fmax(s, T(0));  // sqrt of a rounding-negative s
      e[h] = kernel_entry<T, KIND>(s, exp_tab);
    }
  };
  auto entry2 = [&](const Q& q, T (&e)[2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T dx = pi[h][0] - q.x;
      T s = dx * dx;
      if (DIM > 1) {
        const T dy = pi[h][1] - q.y;
        s = fma(dy, dy, s);
      }
      if (DIM > 2) {
        const T dz = pi[h][2] - q.z;
        s = fma(dz, dz, s);
      }
      T k = kernel_entry<T, KIND>(s, exp_tab);
      if (COINC) k += (s < thr) ? shift : T(0);
      e[h] = k;
    }
  };

  if (I == J) {
    auto diag_group = [&](int g, auto centred) {
#pragma unroll 4
      for (int s = 0; s < kWave; ++s) {
        const Q q = src[g][s];
        T e[2];
        if constexpr (decltype(centred)::value)
          entry2c(q, nsrc[g][s], e);
        else
          entry2(q, e);
        accI[0] = fma(e[0], q.w, accI[0]);
        accI[1] = fma(e[1], q.w, accI[1]);
      }
    };
#pragma unroll 1
    for (int g = 0; g < kGroups; ++g) {
      if (far_group(g)) continue;
      if (CENTER && centered)
        diag_group(g, std::true_type{});
      else
        diag_group(g, std::false_type{});
    }
  } else {
    auto rotate_group = [&](int g, auto centred) -> T {
      T t = T(0);
#pragma unroll 4
      for (int s = 0; s < kWave; ++s) {
        const int jj = kDup == 2 ? lane + s : (lane + s) & (kWave - 1);
        const Q q = src[g][jj];
        T e[2];
        if constexpr (decltype(centred)::value)
          entry2c(q, nsrc[g][jj], e);
        else
          entry2(q, e);
        accI[0] = fma(e[0], q.w, accI[0]);
        accI[1] = fma(e[1], q.w, accI[1]);
        t = fma(e[0], xi[0], t);
        t = fma(e[1], xi[1], t);
        t = wave_rotate(t);
      }
      return t;
    };
    auto group_sum = [&](int g) -> T {
      if (far_group(g)) return T(0);
      if (CENTER && centered) return rotate_group(g, std::true_type{});
      return rotate_group(g, std::false_type{});
    };
    if constexpr (kHoldSums) {
      T tg[kGroups];
#pragma unroll
      for (int g = 0; g < kGroups; ++g) tg[g] = group_sum(g);
      __syncthreads();  // every wave is done reading src
#pragma unroll
      for (int g = 0; g < kGroups; ++g) red[w * kSymTile + g * kWave + lane] = tg[g];
    } else {
#pragma unroll 1
      for (int g = 0; g < kGroups; ++g) red[w * kSymTile + g * kWave + lane] = group_sum(g);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = w * 2 * kWave + h * kWave + lane;
    part_i[p * kSymTile + k] = (double)accI[h];
  }
  if (I != J) {
    __syncthreads();
    for (int k = tid; k < kSymTile; k += kBlock) {
      double s = (double)red[k];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) s += (double)red[q * kSymTile + k];
      part_j[p * kSymTile + k] = s;
    }
  }
}

// y_i = sum over partner tiles P of row i's tile X of that pair's partial,
// plus the nugget term.  A workgroup takes 64 consecutive rows (one tile) and
// four partner slices: wave k sums the partners P = k (mod 4) in ascending
// order, then the four slice sums are added in slice order.  With a partner
// list (planned product) a wave visits only the listed partners of its slice,
// in the same order; the partials it leaves out are zero, so the result is
// the same.
constexpr int kFinishRows = kWave;
constexpr int kFinishSlices = kBlock / kWave;

__device__ __forceinline__ double sym_partial(const double* __restrict__ part_i, const double* __restrict__ part_j,
                                              int64_t X, int64_t P, int64_t t) {
  return P < X ? part_j[sym_pair_index(P, X) * kSymTile + t] : part_i[sym_pair_index(X, P) * kSymTile + t];
}

template <bool LIST>
__global__ __launch_bounds__(kBlock) void kmv_sym_finish_kernel(
    int64_t n, int64_t ntiles, const double* __restrict__ part_i, const double* __restrict__ part_j, double nugget,
    const double* __restrict__ x, double* __restrict__ y, const uint32_t* __restrict__ partners,
    const int32_t* __restrict__ pcount, int64_t pcap, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  __shared__ double red[kFinishSlices][kFinishRows];
  const int r = threadIdx.x % kFinishRows, k = threadIdx.x / kFinishRows;
  for (int64_t row0 = (int64_t)blockIdx.x * kFinishRows; row0 < n; row0 += (int64_t)gridDim.x * kFinishRows) {
    const int64_t i = row0 + r;
    const int64_t X = row0 / kSymTile, t = i % kSymTile;
    double s = 0.0;
    if (i < n) {
      if (LIST) {
        const uint32_t* lst = partners + (X * kFinishSlices + k) * pcap;
        const int cnt = pcount[X * kFinishSlices + k];
        for (int q = 0; q < cnt; ++q) s += sym_partial(part_i, part_j, X, (int64_t)lst[q], t);
      } else {
        for (int64_t P = k; P < ntiles; P += kFinishSlices) s += sym_partial(part_i, part_j, X, P, t);
      }
    }
    __syncthreads();
    red[k][r] = s;
    __syncthreads();
    if (k == 0 && i < n) {
      double v = red[0][r];
#pragma unroll
      for (int q = 1; q < kFinishSlices; ++q) v += red[q][r];
      if (nugget != 0.0) v = fma(nugget, x[i], v);
      y[i] = v;
    }
  }
}

inline int64_t sym_tiles(int64_t n);
bool sym_eligible(const mplak_kernel_desc* d);

inline bool use_symmetric(const mplak_kernel_desc* d) {
  if (!sym_eligible(d)) return false;
  const int mode = sym_mode();
  return mode < 0 ? d->row_count >= kSymMinN : mode == 1;
}

inline int64_t sym_tiles(int64_t n) { return (n + kSymTile - 1) / kSymTile; }

struct SymArgs {
  double scale, thr, drop_s;
  const double* x;
  double *pi, *pj;
  const uint32_t* list;
  const int64_t* count;
  const int32_t* skip;
};

template <typename T, int KIND, int DIM, int MODE>
void launch_sym_dim(hipStream_t st, bool coinc, const mplak_kernel_desc* d, const SymArgs& a) {
  const int64_t nt = sym_tiles(d->row_count);
  const dim3 grid((unsigned)(nt * (nt + 1) / 2));
  if (coinc)
    hipLaunchKernelGGL((kmv_sym_kernel<T, KIND, DIM, true, MODE>), grid, dim3(kBlock), 0, st, d->points,
                       d->row_offset, d->row_count, a.scale, a.thr, d->coincident_shift, a.drop_s, a.x, a.pi, a.pj,
                       a.list, a.count, a.skip);
  else
    hipLaunchKernelGGL((kmv_sym_kernel<T, KIND, DIM, false, MODE>), grid, dim3(kBlock), 0, st, d->points,
                       d->row_offset, d->row_count, a.scale, a.thr, d->coincident_shift, a.drop_s, a.x, a.pi, a.pj,
                       a.list, a.count, a.skip);
}

template <typename T, int KIND, int MODE>
void launch_sym(hipStream_t st, bool coinc, const mplak_kernel_desc* d, const SymArgs& a) {
  switch (d->dim) {
    case 1: launch_sym_dim<T, KIND, 1, MODE>(st, coinc, d, a); break;
    case 2: launch_sym_dim<T, KIND, 2, MODE>(st, coinc, d, a); break;
    default: launch_sym_dim<T, KIND, 3, MODE>(st, coinc, d, a); break;
  }
}

template <int MODE>
void launch_sym_any(hipStream_t st, int precision, bool coinc, const mplak_kernel_desc* d, const SymArgs& a) {
  const bool gauss = d->kind == MPLAK_KERNEL_GAUSSIAN;
  if (precision == MPLAK_PREC_F64) {
    if (gauss)
      launch_sym<double, MPLAK_KERNEL_GAUSSIAN, MODE>(st, coinc, d, a);
    else
      launch_sym<double, MPLAK_KERNEL_MATERN32, MODE>(st, coinc, d, a);
  } else {
    if (gauss)
      launch_sym<float, MPLAK_KERNEL_GAUSSIAN, MODE>(st, coinc, d, a);
    else
      launch_sym<float, MPLAK_KERNEL_MATERN32, MODE>(st, coinc, d, a);
  }
}

bool sym_eligible(const mplak_kernel_desc* d) {
  return d->row_offset == d->col_offset && d->row_count == d->col_count && d->dim <= 3 &&
         sym_tiles(d->row_count) < 65536;  // pair count fits a 1-D grid
}

size_t sym_workspace_bytes(int64_t n) {
  const int64_t nt = sym_tiles(n);
  return (size_t)2 * (size_t)(nt * (nt + 1) / 2) * kSymTile * sizeof(double);
}

// Symmetric product: tile-pair kernel (exact, screened or planned) and the
// finish.  `work` holds part_i then part_j ([pairs][512] doubles each).
int run_sym(const mplak_kernel_desc* d, int precision, int mode, double drop_s, const uint32_t* list,
            const int64_t* count, const uint32_t* partners, const int32_t* pcount, int64_t pcap, const double* x,
            double* y, void* work, const int32_t* skip_if, mplak_stream_t s) {
  SymArgs a;
  a.scale = coord_scale(d->kind, d->length_scale);
  const bool coinc = d->coincident_shift != 0.0;
  a.thr = (1e-14 * a.scale) * (1e-14 * a.scale);
  a.drop_s = drop_s;
  a.x = x;
  a.list = list;
  a.count = count;
  a.skip = skip_if;
  const int64_t nt = sym_tiles(d->row_count);
  a.pi = static_cast<double*>(work);
  a.pj = a.pi + (size_t)(nt * (nt + 1) / 2) * kSymTile;
  double* pi = a.pi;
  double* pj = a.pj;
  hipStream_t hs = as_stream(s);
  if (mode == kSymPlanned)
    launch_sym_any<kSymPlanned>(hs, precision, coinc, d, a);
  else if (mode == kSymScreened)
    launch_sym_any<kSymScreened>(hs, precision, coinc, d, a);
  else
    launch_sym_any<kSymExact>(hs, precision, coinc, d, a);
  int st = launch_status();
  if (st != MPLAK_OK) return st;
  int64_t fb = (d->row_count + kFinishRows - 1) / kFinishRows;
  if (fb > 65536) fb = 65536;
  if (partners != nullptr)
    hipLaunchKernelGGL(kmv_sym_finish_kernel<true>, dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count, nt, pi,
                       pj, d->nugget, x, y, partners, pcount, pcap, skip_if);
  else
    hipLaunchKernelGGL(kmv_sym_finish_kernel<false>, dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count, nt, pi,
                       pj, d->nugget, x, y, nullptr, nullptr, 0, skip_if);
  return launch_status();
}



// ---------------------------------------------------------------------------
// Screening plan: which tile pairs of the symmetric product have at least one
// kept 128 x 64 block (same test as the screened kernel, on the same
// precision-cast coordinates), ordered by kept block count, heaviest first
// (ties by pair index), so that the planned product launches no all-dropped
// pair and starts the longest workgroups first.  The partials of the dropped
// pairs are zeroed once at build time and never written again.
// ---------------------------------------------------------------------------

constexpr int kSymGroups = kSymTile / kWave;
constexpr uint64_t kPlanDropped = 63;  // key class of pairs with no kept block

struct PlanLayout {
  size_t count, boxes, list, keys_in, keys_out, tmp, tmp_bytes, pcount, partners, part, total;
  int64_t pcap;  // partner list capacity per (tile, slice)
};

PlanLayout plan_layout(int64_t n) {
  PlanLayout L{};
  const int64_t nt = sym_tiles(n);
  const size_t np = (size_t)(nt * (nt + 1) / 2);
  L.count = 0;
  L.boxes = 256;
  L.list = L.boxes + align_up(sizeof(double) * 6 * kSymGroups * (size_t)nt, 256);
  L.keys_in = L.list + align_up(sizeof(uint32_t) * np, 256);
  L.keys_out = L.keys_in + align_up(sizeof(uint64_t) * np, 256);
  L.tmp = L.keys_out + align_up(sizeof(uint64_t) * np, 256);
  size_t tb = 0;
  if (rocprim::radix_sort_keys(nullptr, tb, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                               (unsigned)np, 32, 38) != hipSuccess)
    return PlanLayout{};
  L.tmp_bytes = tb;
  L.pcap = (nt + kFinishSlices - 1) / kFinishSlices;
  L.pcount = L.tmp + align_up(tb, 256);
  L.partners = L.pcount + align_up(sizeof(int32_t) * kFinishSlices * (size_t)nt, 256);
  L.part = L.partners + align_up(sizeof(uint32_t) * (size_t)nt * kFinishSlices * (size_t)L.pcap, 256);
  L.total = L.part + 2 * np * kSymTile * sizeof(double);
  return L;
}

// Bounding box of every 64-point group of every tile, from the same
// precision-cast scaled coordinates the product kernel uses (padding points
// are clamped to the last point, as there).
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_box_kernel(const double* const* __restrict__ points, int64_t off,
                                                          int64_t n, double scale, double* __restrict__ boxes) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int k = threadIdx.x; k < kSymTile; k += kBlock) {
    const int64_t j = (int64_t)blockIdx.x * kSymTile + k;
    const int64_t jc = j < n ? j : n - 1;
    T lo[3] = {T(0), T(0), T(0)}, hi[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T v = (T)(points[d][off + jc] * scale);
      lo[d] = wave_min(v);
      hi[d] = wave_max(v);
    }
    if (lane == 0) {
      double* b = boxes + ((size_t)blockIdx.x * kSymGroups + k / kWave) * 6;
      for (int d = 0; d < 3; ++d) {
        b[d] = (double)lo[d];
        b[3 + d] = (double)hi[d];
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void load_box(const double* b, T* out) {
#pragma unroll
  for (int d = 0; d < 6; ++d) out[d] = (T)b[d];
}

// key[p] = (32 - kept blocks) << 32 | p, or kPlanDropped << 32 | p.
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_pair_kernel(const double* __restrict__ boxes, int64_t np,
                                                           double drop_s, uint64_t* __restrict__ keys) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  int64_t I, J;
  decode_pair(p, I, J);
  int kept = 0;
  for (int w = 0; w < kBlock / kWave; ++w) {
    T a[6], b[6], t[6];
    load_box(boxes + ((size_t)I * kSymGroups + 2 * w) * 6, a);
    load_box(boxes + ((size_t)I * kSymGroups + 2 * w + 1) * 6, b);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      t[d] = fmin(a[d], b[d]);
      t[3 + d] = fmax(a[3 + d], b[3 + d]);
    }
    for (int g = 0; g < kSymGroups; ++g) {
      T gb[6];
      load_box(boxes + ((size_t)J * kSymGroups + g) * 6, gb);
      kept += box_gap2<T, DIM>(t, gb) > (T)drop_s ? 0 : 1;
    }
  }
  keys[p] = ((kept > 0 ? (uint64_t)(32 - kept) : kPlanDropped) << 32) | (uint64_t)p;
}

__global__ __launch_bounds__(kBlock) void plan_zero_kernel(const uint64_t* __restrict__ keys, int64_t np,
                                                           double* __restrict__ part_i, double* __restrict__ part_j) {
  for (int64_t p = blockIdx.x; p < np; p += gridDim.x) {
    if ((keys[p] >> 32) != kPlanDropped) continue;
    for (int k = threadIdx.x; k < kSymTile; k += kBlock) {
      part_i[p * kSymTile + k] = 0.0;
      part_j[p * kSymTile + k] = 0.0;
    }
  }
}

__global__ __launch_bounds__(kBlock) void plan_list_kernel(const uint64_t* __restrict__ keys, int64_t np,
                                                           uint32_t* __restrict__ list, int64_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= np) return;
  list[i] = (uint32_t)(keys[i] & 0xffffffffull);
  const bool kept = (keys[i] >> 32) != kPlanDropped;
  if (kept && (i + 1 == np || (keys[i + 1] >> 32) == kPlanDropped)) *count = i + 1;
  if (i == 0 && !kept) *count = 0;
}

// counts[0] = tile pairs with a kept block, counts[1] = kept 128 x 64 blocks,
// counts[2] = kept blocks on diagonal pairs (I == J) (one workgroup, fixed order).
__global__ __launch_bounds__(kBlock) void plan_count_kernel(const uint64_t* __restrict__ keys, int64_t np,
                                                            int64_t* __restrict__ counts) {
  __shared__ int64_t red[3][kBlock];
  int64_t pairs = 0, blocks = 0, diag = 0;
  for (int64_t i = threadIdx.x; i < np; i += kBlock) {
    const uint64_t c = keys[i] >> 32;
    if (c != kPlanDropped) {
      int64_t I, J;
      decode_pair((int64_t)(keys[i] & 0xffffffffull), I, J);
      pairs += 1;
      blocks += 32 - (int64_t)c;
      if (I == J) diag += 32 - (int64_t)c;
    }
  }
  red[0][threadIdx.x] = pairs;
  red[1][threadIdx.x] = blocks;
  red[2][threadIdx.x] = diag;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kBlock; ++k) {
      pairs += red[0][k];
      blocks += red[1][k];
      diag += red[2][k];
    }
    counts[0] = pairs;
    counts[1] = blocks;
    counts[2] = diag;
  }
}

// Partner lists of the finish: for tile X and slice k, the tiles P = k (mod 4)
// in ascending order whose pair with X holds a kept block (one thread per
// (tile, slice); keys are the unsorted per-pair keys).
__global__ __launch_bounds__(kBlock) void plan_partner_kernel(const uint64_t* __restrict__ keys, int64_t nt,
                                                              int64_t pcap, uint32_t* __restrict__ partners,
                                                              int32_t* __restrict__ pcount) {
  const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= nt * kFinishSlices) return;
  const int64_t X = id / kFinishSlices, k = id % kFinishSlices;
  uint32_t* lst = partners + id * pcap;
  int32_t c = 0;
  for (int64_t P = k; P < nt; P += kFinishSlices) {
    const int64_t p = P < X ? sym_pair_index(P, X) : sym_pair_index(X, P);
    if ((keys[p] >> 32) != kPlanDropped) lst[c++] = (uint32_t)P;
  }
  pcount[id] = c;
}

template <typename T>
void launch_plan_geometry(hipStream_t st, const mplak_kernel_desc* d, double scale, double drop_s, double* boxes,
                          uint64_t* keys) {
  const int64_t nt = sym_tiles(d->row_count);
  const int64_t np = nt * (nt + 1) / 2;
  const unsigned pg = (unsigned)((np + kBlock - 1) / kBlock);
  switch (d->dim) {
    case 1:
      hipLaunchKernelGGL((plan_box_kernel<T, 1>), dim3((unsigned)nt), dim3(kBlock), 0, st, d->points, d->row_offset,
                         d->row_count, scale, boxes);
      hipLaunchKernelGGL((plan_pair_kernel<T, 1>), dim3(pg), dim3(kBlock), 0, st, boxes, np, drop_s, keys);
      break;
    case 2:
      hipLaunchKernelGGL((plan_box_kernel<T, 2>), dim3((unsigned)nt), dim3(kBlock), 0, st, d->points, d->row_offset,
                         d->row_count, scale, boxes);
      hipLaunchKernelGGL((plan_pair_kernel<T, 2>), dim3(pg), dim3(kBlock), 0, st, boxes, np, drop_s, keys);
      break;
    default:
      hipLaunchKernelGGL((plan_box_kernel<T, 3>), dim3((unsigned)nt), dim3(kBlock), 0, st, d->points, d->row_offset,
                         d->row_count, scale, boxes);
      hipLaunchKernelGGL((plan_pair_kernel<T, 3>), dim3(pg), dim3(kBlock), 0, st, boxes, np, drop_s, keys);
      break;
  }
}

int check_desc(const mplak_kernel_desc* d) {
  if (d == nullptr || d->points == nullptr) return MPLAK_ERR_ARG;
  if (d->dim < 1 || d->dim > kKmvMaxDim) return MPLAK_ERR_ARG;
  if (d->kind != MPLAK_KERNEL_GAUSSIAN && d->kind != MPLAK_KERNEL_MATERN32) return MPLAK_ERR_ARG;
  if (!(d->length_scale > 0.0)) return MPLAK_ERR_ARG;
  if (d->row_count < 0 || d->col_count < 0 || d->row_offset < 0 || d->col_offset < 0)
    return MPLAK_ERR_ARG;
  return MPLAK_OK;
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

size_t mplak_kernel_matvec_workspace_bytes(const mplak_kernel_desc* d) {
  if (d == nullptr || d->row_count <= 0 || d->col_count <= 0) return 0;
  if (use_symmetric(d)) return sym_workspace_bytes(d->row_count);
  const KmvPlan p = kmv_plan(d->row_count, d->col_count);
  return (size_t)p.slabs * (size_t)d->row_count * sizeof(double);
}

size_t mplak_kernel_matvec_screened_workspace_bytes(const mplak_kernel_desc* d) {
  if (d == nullptr || d->row_count <= 0 || d->col_count <= 0) return 0;
  if (sym_eligible(d)) return sym_workspace_bytes(d->row_count);
  return mplak_kernel_matvec_workspace_bytes(d);
}

int mplak_kernel_matvec_screened(const mplak_kernel_desc* d, int precision, double drop_log2, const double* x,
                                 double* y, void* work, size_t work_bytes, const int32_t* skip_if,
                                 mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision != MPLAK_PREC_F64 && precision != MPLAK_PREC_F32) return MPLAK_ERR_ARG;
  if (!(drop_log2 > 0.0) || drop_log2 > 2000.0) return MPLAK_ERR_ARG;
  if (!sym_eligible(d)) return mplak_kernel_matvec(d, precision, x, y, work, work_bytes, skip_if, s);
  if (d->row_count == 0) return MPLAK_OK;
  if (x == nullptr || y == nullptr) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < sym_workspace_bytes(d->row_count)) return MPLAK_ERR_WORKSPACE;
  return run_sym(d, precision, kSymScreened, drop_distance2(d->kind, drop_log2), nullptr, nullptr,
                 nullptr, nullptr, 0, x, y, work, skip_if, s);
}

size_t mplak_screen_plan_bytes(const mplak_kernel_desc* d) {
  if (d == nullptr || d->row_count <= 0 || !sym_eligible(d)) return 0;
  return plan_layout(d->row_count).total;
}

int mplak_screen_plan_build(const mplak_kernel_desc* d, int precision, double drop_log2, void* plan,
                            size_t plan_bytes, mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision != MPLAK_PREC_F64 && precision != MPLAK_PREC_F32) return MPLAK_ERR_ARG;
  if (!(drop_log2 > 0.0) || drop_log2 > 2000.0) return MPLAK_ERR_ARG;
  if (!sym_eligible(d) || d->row_count == 0) return MPLAK_ERR_ARG;
  const PlanLayout L = plan_layout(d->row_count);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (plan == nullptr || plan_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  char* base = static_cast<char*>(plan);
  const int64_t nt = sym_tiles(d->row_count);
  const int64_t np = nt * (nt + 1) / 2;
  double* boxes = reinterpret_cast<double*>(base + L.boxes);
  uint64_t* kin = reinterpret_cast<uint64_t*>(base + L.keys_in);
  uint64_t* kout = reinterpret_cast<uint64_t*>(base + L.keys_out);
  double* pi = reinterpret_cast<double*>(base + L.part);
  double* pj = pi + (size_t)np * kSymTile;
  const double scale = coord_scale(d->kind, d->length_scale);
  const double drop_s = drop_distance2(d->kind, drop_log2);
  hipStream_t hs = as_stream(s);
  if (precision == MPLAK_PREC_F64)
    launch_plan_geometry<double>(hs, d, scale, drop_s, boxes, kin);
  else
    launch_plan_geometry<float>(hs, d, scale, drop_s, boxes, kin);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(plan_zero_kernel, dim3((unsigned)std::min<int64_t>(np, 4096)), dim3(kBlock), 0, hs, kin, np, pi,
                     pj);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(plan_partner_kernel, dim3((unsigned)((nt * kFinishSlices + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, hs, kin, nt, L.pcap, reinterpret_cast<uint32_t*>(base + L.partners),
                     reinterpret_cast<int32_t*>(base + L.pcount));
  st = launch_status();
  if (st != MPLAK_OK) return st;
  size_t tb = L.tmp_bytes;
  if (rocprim::radix_sort_keys(base + L.tmp, tb, kin, kout, (unsigned)np, 32, 38, hs) != hipSuccess)
    return MPLAK_ERR_HIP;
  hipLaunchKernelGGL(plan_list_kernel, dim3((unsigned)((np + kBlock - 1) / kBlock)), dim3(kBlock), 0, hs, kout, np,
                     reinterpret_cast<uint32_t*>(base + L.list), reinterpret_cast<int64_t*>(base + L.count));
  return launch_status();
}

int mplak_screen_plan_counts(const mplak_kernel_desc* d, const void* plan, size_t plan_bytes, int64_t* counts,
                             mplak_stream_t s) {
  if (d == nullptr || plan == nullptr || counts == nullptr || d->row_count <= 0 || !sym_eligible(d))
    return MPLAK_ERR_ARG;
  const PlanLayout L = plan_layout(d->row_count);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (plan_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  const int64_t nt = sym_tiles(d->row_count);
  const char* base = static_cast<const char*>(plan);
  hipLaunchKernelGGL(plan_count_kernel, dim3(1), dim3(kBlock), 0, as_stream(s),
                     reinterpret_cast<const uint64_t*>(base + L.keys_out), nt * (nt + 1) / 2, counts);
  return launch_status();
}

int mplak_kernel_matvec_planned(const mplak_kernel_desc* d, int precision, double drop_log2, void* plan,
                                size_t plan_bytes, const double* x, double* y, const int32_t* skip_if,
                                mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision != MPLAK_PREC_F64 && precision != MPLAK_PREC_F32) return MPLAK_ERR_ARG;
  if (!(drop_log2 > 0.0) || drop_log2 > 2000.0) return MPLAK_ERR_ARG;
  if (!sym_eligible(d) || d->row_count == 0) return MPLAK_ERR_ARG;
  if (x == nullptr || y == nullptr) return MPLAK_ERR_ARG;
  const PlanLayout L = plan_layout(d->row_count);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (plan == nullptr || plan_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  char* base = static_cast<char*>(plan);
  return run_sym(d, precision, kSymPlanned, drop_distance2(d->kind, drop_log2),
                 reinterpret_cast<const uint32_t*>(base + L.list), reinterpret_cast<const int64_t*>(base + L.count),
                 reinterpret_cast<const uint32_t*>(base + L.partners), reinterpret_cast<const int32_t*>(base + L.pcount),
                 L.pcap, x, y, base + L.part, skip_if, s);
}

int mplak_kernel_matvec(const mplak_kernel_desc* d, int precision, const double* x, double* y,
                        void* work, size_t work_bytes, const int32_t* skip_if, mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision != MPLAK_PREC_F64 && precision != MPLAK_PREC_F32) return MPLAK_ERR_ARG;
  if (d->row_count == 0) return MPLAK_OK;
  if (y == nullptr) return MPLAK_ERR_ARG;
  if (d->col_count == 0) return mplak_fill_f64(y, d->row_count, 0.0, s);
  if (x == nullptr) return MPLAK_ERR_ARG;
  const size_t need = mplak_kernel_matvec_workspace_bytes(d);
  if (work == nullptr || work_bytes < need) return MPLAK_ERR_WORKSPACE;
  if (use_symmetric(d))
    return run_sym(d, precision, kSymExact, 0.0, nullptr, nullptr, nullptr, nullptr, 0, x, y, work, skip_if, s);

  const double scale = coord_scale(d->kind, d->length_scale);
  const bool coinc = d->coincident_shift != 0.0;
  const double thr = (1e-14 * scale) * (1e-14 * scale);
  double* part = static_cast<double*>(work);
  const KmvPlan p = kmv_plan(d->row_count, d->col_count);
  const dim3 grid((unsigned)p.row_tiles, (unsigned)p.slabs);
  hipStream_t hs = as_stream(s);
  if (precision == MPLAK_PREC_F64) {
    if (d->kind == MPLAK_KERNEL_GAUSSIAN)
      dispatch_dim<double, MPLAK_KERNEL_GAUSSIAN>(grid, hs, coinc, d, p.slab_cols, scale, thr, x, part, skip_if);
    else
      dispatch_dim<double, MPLAK_KERNEL_MATERN32>(grid, hs, coinc, d, p.slab_cols, scale, thr, x, part, skip_if);
  } else {
    if (d->kind == MPLAK_KERNEL_GAUSSIAN)
      dispatch_dim<float, MPLAK_KERNEL_GAUSSIAN>(grid, hs, coinc, d, p.slab_cols, scale, thr, x, part, skip_if);
    else
      dispatch_dim<float, MPLAK_KERNEL_MATERN32>(grid, hs, coinc, d, p.slab_cols, scale, thr, x, part, skip_if);
  }
  st = launch_status();
  if (st != MPLAK_OK) return st;
  int64_t fb = (d->row_count + kBlock - 1) / kBlock;
  if (fb > 4096) fb = 4096;
  hipLaunchKernelGGL(kmv_finish_kernel, dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count,
                     p.slabs, part, d->row_offset, d->col_offset, d->col_count, d->nugget, x, y,
                     skip_if);
  return launch_status();
}

int mplak_gaussian_kernel_rhs(const double* const* points, int32_t dim, int64_t row_offset,
                              int64_t row_count, double* b, mplak_stream_t s) {
  if (points == nullptr || dim < 1 || row_offset < 0 || row_count < 0) return MPLAK_ERR_ARG;
  if (row_count == 0) return MPLAK_OK;
  if (b == nullptr) return MPLAK_ERR_ARG;
  int64_t g = (row_count + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(gaussian_rhs_kernel, dim3((unsigned)g), dim3(kBlock), 0, as_stream(s), points,
                     dim, row_offset, row_count, b);
  return launch_status();
}

}  // extern "C"
