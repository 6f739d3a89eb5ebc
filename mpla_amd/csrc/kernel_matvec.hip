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
#include <mutex>
#include <type_traits>

#include <rocprim/device/device_radix_sort.hpp>

#include "common.hpp"
#include "entry.hpp"

// Scratch of mplak_kernel_matvec_noalloc: a fixed device array, so that the
// entry point allocates nothing.  80 MiB holds the symmetric product's tile
// partials up to n = 65536 (67.6 MB) and the plain product's slab partials
// of the blocks that fit.  Each device that runs the library holds one copy
// until the process ends; calls that overlap on one device share it and race.
constexpr size_t kNoallocScratchBytes = size_t(80) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace mplak {
namespace {

// Device address of `scratch` on the current device, looked up once per
// device with that device current.  The host-side name `scratch` is not a
// device address and is never passed to a kernel or a copy.
int noalloc_scratch(void** out) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* ptrs[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return MPLAK_ERR_HIP;
  std::lock_guard<std::mutex> lock(mu);
  if (ptrs[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return MPLAK_ERR_HIP;
    ptrs[dev] = p;
  }
  *out = ptrs[dev];
  return MPLAK_OK;
}

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

// The finish: 64 rows per workgroup, one partner slice per wave.
constexpr int kFinishRows = kWave;
constexpr int kFinishSlices = kBlock / kWave;
constexpr int kFinishBatch = 8;  // independent partial loads in flight per wave

// Partial rows of a plan's listed pairs ("slotted" layout): tile X has one
// block of pcap rows of 512 per slice k = P mod kFinishSlices, holding the
// rows of X of its listed pairs (X, P) of that slice in ascending P, so the
// finish reads a slice's partials knowing only their count.  slots[2 p] is
// the row of pair p = (I, J) in I's block, slots[2 p + 1] its row in J's
// block (I < J).
template <typename P>
__device__ __forceinline__ P* slot_row(P* part, int64_t X, int64_t partner, int32_t slot, int64_t pcap) {
  return part + ((X * kFinishSlices + (partner & (kFinishSlices - 1))) * pcap + slot) * kSymTile;
}

constexpr int kSymMinN = 12288;       // crossover measured on MI355X (plain kernel faster below)

// MPLAK_KMV_SYMMETRIC=0 turns the symmetric product off, =1 uses it for every
// eligible block regardless of size; unset: blocks of at least kSymMinN points.
inline int sym_mode() {
  const char* v = getenv("MPLAK_KMV_SYMMETRIC");
  if (v == nullptr || *v == '\0') return -1;
  return atoi(v) != 0 ? 1 : 0;
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

// Load through the global address space (pointers read from memory are
// generic to the compiler, and flat loads also count against the LDS
// counter, which forces waits on LDS traffic).
__device__ __forceinline__ double global_load(const double* ptr) {
  typedef const __attribute__((address_space(1))) double* global_ptr;
  return *(global_ptr)ptr;
}

enum SymMode { kSymExact = 0, kSymScreened = 1, kSymPlanned = 2 };

// One tile pair (I, J) of the symmetric product on the vector unit.
// Screening (SCREEN): a wave skips a 64-source group when every entry of
// that 128 x 64 block is below the drop level.  MASKED takes that decision
// from bit 8 w + g of `wmask` (a screening plan); otherwise it compares the
// wave-reduced bounding boxes of the wave's 128 targets and of the group
// (squared gap in scaled coordinates > drop_s), so the branch is uniform
// across the wave.  src, gbox, own_red and exp_tab are the caller's LDS.
// One staged 64-source group: stored twice in a row in FP32 (the rotated
// read (l + s) mod 64 becomes an immediate offset), once in FP64.
template <typename T>
using SymSrc = typename Quad<T>::type[(sizeof(T) == 4 ? 2 : 1) * kWave];

template <typename T, int KIND, int DIM, bool COINC, bool SCREEN, bool MASKED, typename P>
__device__ __forceinline__ void sym_valu_pair(const double* const* __restrict__ points, int64_t off, int64_t n,
                                              double scale, double coinc_thr, double coinc_shift, double drop_s,
                                              const double* __restrict__ x, P* __restrict__ rows_i,
                                              P* __restrict__ rows_j, int64_t p, uint32_t pmask,
                                              SymSrc<T>* src, T (*gbox)[6], T* own_red, const double* exp_tab) {
  typedef typename Quad<T>::type Q;
  constexpr int kDup = sizeof(T) == 4 ? 2 : 1;  // copies of each 64-source group
  constexpr int kGroups = kSymTile / kWave;
  // Column-side sums of the four waves, red[w * kSymTile + k].  In FP32 each
  // wave holds its 8 group sums in registers and they reuse the source buffer
  // once every wave is done with it; in FP64 registers are the scarcer
  // resource, and the sums go to a buffer of their own as each group
  // completes.
  constexpr bool kHoldSums = sizeof(T) == 4;
  T* red = kHoldSums ? reinterpret_cast<T*>(&src[0][0]) : own_red;
  int64_t I, J;
  decode_pair(p, I, J);

  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int w = tid / kWave;
  const T thr = (T)coinc_thr, shift = (T)coinc_shift;
  const uint32_t wmask = MASKED ? (uint32_t)__builtin_amdgcn_readfirstlane((int)((pmask >> (8 * w)) & 0xffu)) : 0u;

  // two targets per lane: tile offsets w*128 + lane and w*128 + 64 + lane;
  // every global load of the staging is issued before the first use
  constexpr int kPer = kSymTile / kBlock;
  double ti[2][DIM], tv[2], sc_[kPer][DIM], sv[kPer];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = I * kSymTile + w * 2 * kWave + h * kWave + lane;
    const int64_t ic = i < n ? i : n - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) ti[h][d] = global_load(points[d] + off + ic);
    tv[h] = global_load(x + ic);
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int64_t j = J * kSymTile + tid + it * kBlock;
    const int64_t jc = j < n ? j : n - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) sc_[it][d] = global_load(points[d] + off + jc);
    sv[it] = global_load(x + jc);
  }
  T pi[2][DIM], xi[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = I * kSymTile + w * 2 * kWave + h * kWave + lane;
#pragma unroll
    for (int d = 0; d < DIM; ++d) pi[h][d] = (T)(ti[h][d] * scale);
    xi[h] = i < n ? (T)tv[h] : T(0);
  }
  constexpr bool kBoxes = SCREEN && !MASKED;
  T tbox[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
  if (kBoxes) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      tbox[d] = wave_min(fmin(pi[0][d], pi[1][d]));
      tbox[3 + d] = wave_max(fmax(pi[0][d], pi[1][d]));
    }
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int k = tid + it * kBlock;
    const int64_t j = J * kSymTile + k;
    Q q;
    q.x = (T)(sc_[it][0] * scale);
    q.y = T(0);
    q.z = T(0);
    if constexpr (DIM > 1) q.y = (T)(sc_[it][1] * scale);
    if constexpr (DIM > 2) q.z = (T)(sc_[it][2] * scale);
    q.w = j < n ? (T)sv[it] : T(0);
    if (kBoxes) {  // k / kWave is uniform across the wave: one group per wave and pass
      const T lo0 = wave_min(q.x), lo1 = wave_min(q.y), lo2 = wave_min(q.z);
      const T hi0 = wave_max(q.x), hi1 = wave_max(q.y), hi2 = wave_max(q.z);
      if (lane == 0) {
        T* b = gbox[k / kWave];
        b[0] = lo0, b[1] = lo1, b[2] = lo2, b[3] = hi0, b[4] = hi1, b[5] = hi2;
      }
    }
#pragma unroll
    for (int c = 0; c < kDup; ++c) src[k / kWave][c * kWave + (k & (kWave - 1))] = q;
  }
  // true when every entry between this wave's targets and group g is dropped
  auto far_group = [&](int g) -> bool {
    if (!SCREEN) return false;
    if (MASKED) return ((wmask >> g) & 1u) == 0;
    const T d2 = box_gap2<T, DIM>(tbox, gbox[g]);
    return __builtin_amdgcn_readfirstlane((int)(d2 > (T)drop_s)) != 0;
  };
  __syncthreads();

  T accI[2] = {T(0), T(0)};
  auto entry2 = [&](const Q& q, T (&e)[2]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const T dx = pi[h][0] - q.x;
      T s = dx * dx;
      if constexpr (DIM > 1) {
        const T dy = pi[h][1] - q.y;
        s = fma(dy, dy, s);
      }
      if constexpr (DIM > 2) {
        const T dz = pi[h][2] - q.z;
        s = fma(dz, dz, s);
      }
      T k = kernel_entry<T, KIND>(s, exp_tab);
      if (COINC) k += (s < thr) ? shift : T(0);
      e[h] = k;
    }
  };

  if (I == J) {
#pragma unroll 1
    for (int g = 0; g < kGroups; ++g) {
      if (far_group(g)) continue;
#pragma unroll 4
      for (int s = 0; s < kWave; ++s) {
        const Q q = src[g][s];
        T e[2];
        entry2(q, e);
        accI[0] = fma(e[0], q.w, accI[0]);
        accI[1] = fma(e[1], q.w, accI[1]);
      }
    }
  } else {
    auto group_sum = [&](int g) -> T {
      if (far_group(g)) return T(0);
      T t = T(0);
#pragma unroll 4
      for (int s = 0; s < kWave; ++s) {
        const int jj = kDup == 2 ? lane + s : (lane + s) & (kWave - 1);
        const Q q = src[g][jj];
        T e[2];
        entry2(q, e);
        accI[0] = fma(e[0], q.w, accI[0]);
        accI[1] = fma(e[1], q.w, accI[1]);
        t = fma(e[0], xi[0], t);
        t = fma(e[1], xi[1], t);
        t = wave_rotate(t);
      }
      return t;
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
    rows_i[k] = (P)accI[h];
  }
  if (I != J) {
    __syncthreads();
    for (int k = tid; k < kSymTile; k += kBlock) {
      double s = (double)red[k];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) s += (double)red[q * kSymTile + k];
      rows_j[k] = (P)s;
    }
  }
}

// Symmetric product kernel on the vector unit: kSymExact / kSymScreened over
// all tile pairs (workgroup p = pair p, partial rows by pair),
// kSymPlanned over the tile pairs list[0 .. *count) of a screening plan
// (pairs with a kept block; rows in the plan's slotted layout, which holds
// no row for the others).
template <typename T, int KIND, int DIM, bool COINC, int MODE>
__global__ __launch_bounds__(kBlock) void kmv_sym_kernel(
    const double* const* __restrict__ points, int64_t off, int64_t n, double scale, double coinc_thr,
    double coinc_shift, double drop_s, const double* __restrict__ x, double* __restrict__ part_i,
    double* __restrict__ part_j, const uint32_t* __restrict__ list, const int64_t* __restrict__ count,
    const int32_t* __restrict__ slots, int64_t pcap, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  if (MODE == kSymPlanned && (int64_t)blockIdx.x >= *count) return;
  typedef typename Quad<T>::type Q;
  constexpr int kDup = sizeof(T) == 4 ? 2 : 1;
  constexpr int kGroups = kSymTile / kWave;
  constexpr bool kHoldSums = sizeof(T) == 4;
  static_assert(sizeof(Q) * kGroups * kDup * kWave >= sizeof(T) * (kBlock / kWave) * kSymTile, "LDS reuse");
  __shared__ Q src[kGroups][kDup * kWave];
  static_assert(sizeof(SymSrc<T>) == sizeof(src[0]), "staged group layout");
  __shared__ T gbox[kGroups][6];
  __shared__ T own_red[kHoldSums ? 1 : (kBlock / kWave) * kSymTile];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the staging barrier
  const int64_t p = MODE == kSymPlanned ? (int64_t)list[blockIdx.x] : (int64_t)blockIdx.x;
  // partial rows: by pair (part_i for tile I, part_j for tile J), or with a
  // plan in its slotted layout (part_i)
  double* rows_i = part_i + p * kSymTile;
  double* rows_j = part_j + p * kSymTile;
  if (MODE == kSymPlanned) {
    int64_t I, J;
    decode_pair(p, I, J);
    rows_i = slot_row(part_i, I, J, slots[2 * p], pcap);
    rows_j = I != J ? slot_row(part_i, J, I, slots[2 * p + 1], pcap) : rows_i;
  }
  sym_valu_pair<T, KIND, DIM, COINC, MODE != kSymExact, false, double>(
      points, off, n, scale, coinc_thr, coinc_shift, drop_s, x, rows_i, rows_j, p, 0u,
      reinterpret_cast<SymSrc<T>*>(&src[0][0]), gbox, own_red, exp_tab);
}

// ---------------------------------------------------------------------------
// Planned FP32 product (plan built for FP32, no coincidence shift): one
// workgroup per listed tile pair, heaviest first.
//
// Compact target tile: matrix-core form.  In coordinates centred on the
// target tile, s = |p - q|^2 = n_p + n_q - 2 p.q (n = squared norm), and
// v_mfma_f32_16x16x4f32 with A = [p, n_p] (16 targets), B = [-2 q, 1] (16
// sources) and C = n_q (one value per lane) gives s for a 16 x 16 tile in
// the accumulator layout: lane 16 r + c holds targets 4 r + v (v < 4)
// against source c.  The vector unit is left with the exponential and the
// two accumulations.  The tile-centred norms are small (half-diagonal^2 <=
// kCenterRadius2), so the cancellation error stays at the level of the FP32
// coordinates themselves.  Both tiles are staged in LDS in operand layout.
// Wave w takes the target slabs 4 a + w (a < 8; interleaved, which balances
// the waves' kept tiles on spatially ordered points) against the 16 x 16
// tiles the plan keeps (tile mask of (w, group g): bit 8 b + a for the
// wave's slab a and source slab b of the group); the branches are on scalar
// masks.  Row sums accR[a][v] collect over the sources of the lane's column
// and are reduce-scattered over the 16 lanes of a DPP row once per tile pair;
// column sums collect over the lane's 4 targets of every slab and are
// reduce-scattered over the four lane rows once per group, then added over
// the four waves in wave order.
//
// Other target tiles: the vector unit's difference form over the plan's kept
// 128 x 64 blocks.  All orders are fixed, so the product is deterministic.
// ---------------------------------------------------------------------------
constexpr float kCenterRadius2 = 49.0f;
constexpr int kSymGroups = kSymTile / kWave;
constexpr int kSlabs = kSymTile / 16;                      // 16-point slabs per tile
constexpr int kTmaskWords = (kBlock / kWave) * kSymGroups;  // tile masks per pair

typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Reduce-scatter over the 16 lanes of a DPP row: v[q] (q < 32) of every lane
// in, out[h] = sum over the row's lanes of v[16 h + c] for lane c.  Round k
// pairs lane c with c ^ 15, c ^ 7, c ^ 2, c ^ 1 (row_mirror,
// row_half_mirror, quad_perm) and keeps the values whose index bit 3, 2, 1,
// 0 equals the lane's, adding the partner's copy of them.  Fixed order.
__device__ __forceinline__ void row16_reduce_scatter(const float (&v)[32], int c16, float (&out)[2]) {
  float r1[16], r2[8], r3[4];
  const bool b3 = (c16 & 8) != 0, b2 = (c16 & 4) != 0, b1 = (c16 & 2) != 0, b0 = (c16 & 1) != 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {  // index 16 h + q (bit 3 clear) pairs with + 8
      const float lo = v[16 * h + q], hi = v[16 * h + q + 8];
      r1[8 * h + q] = (b3 ? hi : lo) + dpp_f32<0x140>(b3 ? lo : hi);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float lo = r1[8 * h + q], hi = r1[8 * h + q + 4];
      r2[4 * h + q] = (b2 ? hi : lo) + dpp_f32<0x141>(b2 ? lo : hi);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const float lo = r2[4 * h + q], hi = r2[4 * h + q + 2];
      r3[2 * h + q] = (b1 ? hi : lo) + dpp_f32<0x4E>(b1 ? lo : hi);
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const float lo = r3[2 * h], hi = r3[2 * h + 1];
    out[h] = (b0 ? hi : lo) + dpp_f32<0xB1>(b0 ? lo : hi);
  }
}

// Reduce-scatter over the four 16-lane rows of the wave: v[b] of every lane
// in; lane 16 r + c gets the sum over r' of v[r] of lane 16 r' + c.  Fixed
// order: rows paired by r ^ 1 (lane ^ 16), then r ^ 2 (lane ^ 32).
__device__ __forceinline__ float rows4_reduce_scatter(const float (&v)[4], int r4) {
  const bool b0 = (r4 & 1) != 0, b1 = (r4 & 2) != 0;
  float t[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const float lo = v[2 * q], hi = v[2 * q + 1];
    t[q] = (b0 ? hi : lo) + __shfl_xor(b0 ? lo : hi, 16);
  }
  return (b1 ? t[1] : t[0]) + __shfl_xor(b1 ? t[0] : t[1], 32);
}

// LDS layout of the matrix-core form (floats): operands [32 slabs][64 lanes]
// of the targets (lane 16 k + c: coordinate k of target 16 slab + c for
// k < 3, at k = 3 its norm (Matern) or 1 (Gaussian)) and of the sources
// (-2 x coordinate k; at k = 3: 1 (Matern) or the source norm (Gaussian)),
// then source norms (Matern), source values, target values (Gaussian: times
// 2^-n_p) and, Gaussian, the target row scales 2^-n_p.
constexpr int kMfOp = kSlabs * kWave;
constexpr int kMfB = kMfOp;
constexpr int kMfNorm = 2 * kMfOp;
constexpr int kMfXs = kMfNorm + kSymTile;
constexpr int kMfXt = kMfXs + kSymTile;
constexpr int kMfScale = kMfXt + kSymTile;
constexpr int kMfFloats = kMfScale + kSymTile;

// Gaussian entries factor as 2^-s = 2^-n_p 2^-t with t = n_q - 2 p.q, so the
// matrix cores give t with no C operand (A = [p, 1], B = [-2 q, n_q]); the
// factor 2^-n_p goes into the target values (column side) and into the row
// sums once per target.  t >= -n_p >= -kCenterRadius2 on compact tiles, so
// 2^-t stays in range.  Matern entries are not a product of that form and
// take s = t + n_p with C = n_q.
template <int KIND, int DIM, typename P>
__device__ __forceinline__ void sym_mfma_pair(const double* const* __restrict__ points, int64_t off, int64_t n,
                                              double scale, const double* __restrict__ x, P* __restrict__ rows_i,
                                              P* __restrict__ rows_j, int64_t p, int64_t I, int64_t J,
                                              float4 centre, const uint32_t* __restrict__ tmask, float* mf) {
  constexpr bool FACTOR = KIND == MPLAK_KERNEL_GAUSSIAN;
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int c16 = lane & 15, r4 = lane >> 4;
  const float ctr[3] = {centre.x, centre.y, centre.z};

  // staging: centred coordinates c, norm |c|^2, value x (0 past the end).
  // Thread t stages points t and t + 256 of both tiles; every global load is
  // issued first (clamped indices, no branches), then the LDS writes.
  constexpr int kPer = kSymTile / kBlock;
  double pc[2][kPer][3], pv[2][kPer];
#pragma unroll
  for (int side = 0; side < 2; ++side) {  // 0: sources (tile J), 1: targets (tile I)
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int64_t j = (side == 0 ? J : I) * kSymTile + tid + it * kBlock;
      const int64_t jc = j < n ? j : n - 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) pc[side][it][d] = global_load(points[d] + off + jc);
      pv[side][it] = global_load(x + jc);
    }
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int k = tid + it * kBlock;
      const int64_t j = (side == 0 ? J : I) * kSymTile + k;
      float c[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int d = 0; d < DIM; ++d) c[d] = (float)(pc[side][it][d] * scale) - ctr[d];
      float nc = c[0] * c[0];
      if (DIM > 1) nc = fmaf(c[1], c[1], nc);
      if (DIM > 2) nc = fmaf(c[2], c[2], nc);
      const float xv = j < n ? (float)pv[side][it] : 0.0f;
      const int o = (k >> 4) * kWave + (k & 15);
      if (side == 0) {
        mf[kMfB + o] = -2.0f * c[0];
        mf[kMfB + o + 16] = -2.0f * c[1];
        mf[kMfB + o + 32] = -2.0f * c[2];
        mf[kMfB + o + 48] = FACTOR ? nc : 1.0f;
        if (!FACTOR) mf[kMfNorm + k] = nc;
        mf[kMfXs + k] = xv;
      } else {
        mf[o] = c[0];
        mf[o + 16] = c[1];
        mf[o + 32] = c[2];
        mf[o + 48] = FACTOR ? 1.0f : nc;
        const float sc = FACTOR ? __builtin_amdgcn_exp2f(-nc) : 1.0f;
        mf[kMfXt + k] = xv * sc;
        if (FACTOR) mf[kMfScale + k] = sc;
      }
    }
  }
  // the wave's 8 group masks (scalar loads, before the barrier)
  uint32_t gmask[kSymGroups];
  const uint32_t* tm = tmask + p * kTmaskWords + w * kSymGroups;
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) gmask[g] = (uint32_t)__builtin_amdgcn_readfirstlane((int)tm[g]);
  __syncthreads();

  float A[8], accR[8][4];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    A[a] = mf[(4 * a + w) * kWave + lane];
#pragma unroll
    for (int v = 0; v < 4; ++v) accR[a][v] = 0.0f;
  }
  const float* xt = &mf[kMfXt + 16 * w + 4 * r4];  // xt[64 a + v]: target 16 (4 a + w) + 4 r + v
  // One 64-source group: target slabs a (outer) against the group's four
  // source slabs b; the kept MFMAs of slab a are issued before their results
  // are used.  accR[a][v] collects over b, accC[b] over a and v, in the same
  // order whatever the mask.
  auto mfma_group = [&](int g, uint32_t gm, auto diagonal) -> float {
    constexpr bool DIAG = decltype(diagonal)::value;
    float bv[4], xn[4];
    f32x4 cq[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      bv[b] = mf[kMfB + (g * 4 + b) * kWave + lane];
      xn[b] = mf[kMfXs + g * kWave + 16 * b + c16];
      cq[b] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
      if (!FACTOR) {
        const float nq = mf[kMfNorm + g * kWave + 16 * b + c16];
        cq[b] = f32x4{nq, nq, nq, nq};
      }
    }
    float accC[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      // bit b: tile (a, b) kept (bit 8 b + a of gm)
      const uint32_t cm = ((gm >> a) & 1u) | ((gm >> (7 + a)) & 2u) | ((gm >> (14 + a)) & 4u) | ((gm >> (21 + a)) & 8u);
      if (cm == 0) continue;
      const f32x4 xm = *reinterpret_cast<const f32x4*>(&xt[64 * a]);
      // diagonal pair: the tile of target slab 4 a + w against source slab
      // 4 g + b is on the diagonal for a = g, b = w (rows only); the plan
      // keeps no tile below it
      const int dtile = DIAG && a == g ? w : -1;
      f32x4 s[4];
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((cm >> b) & 1u) s[b] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[a], bv[b], cq[b], 0, 0, 0);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        if (((cm >> b) & 1u) == 0) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float sv = s[b][v];
          if (KIND == MPLAK_KERNEL_MATERN32) sv = fmaxf(sv, 0.0f);  // sqrt of a rounding-negative s
          const float e = kernel_entry<float, KIND>(sv, nullptr);
          accR[a][v] = fmaf(e, xn[b], accR[a][v]);
          if (b != dtile) accC[b] = fmaf(e, xm[v], accC[b]);
        }
      }
    }
    // lane 16 r + c keeps the sum of source 16 r + c of the group
    return rows4_reduce_scatter(accC, r4);
  };
  float tg[kSymGroups];
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) tg[g] = 0.0f;
  // group g's mask: the uniform select over gmask keeps it in scalar registers
  auto group_mask = [&](int g) -> uint32_t {
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < kSymGroups; ++q)
      if (q == g) m = gmask[q];
    return m;
  };
  if (I == J) {
#pragma unroll 1
    for (int g = 0; g < kSymGroups; ++g) {
      const uint32_t gm = group_mask(g);
      if (gm == 0) continue;
      const float v = mfma_group(g, gm, std::true_type{});
#pragma unroll
      for (int q = 0; q < kSymGroups; ++q)
        if (q == g) tg[q] = v;
    }
  } else {
#pragma unroll 1
    for (int g = 0; g < kSymGroups; ++g) {
      const uint32_t gm = group_mask(g);
      if (gm == 0) continue;
      const float v = mfma_group(g, gm, std::false_type{});
#pragma unroll
      for (int q = 0; q < kSymGroups; ++q)
        if (q == g) tg[q] = v;
    }
  }
  // row sums: lane 16 r + c writes rows q = c and q = c + 16 of its lane
  // row's 32 (q = 4 a + v, tile row 16 (4 a + w) + 4 r + v)
  float rv[32], out[2];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
#pragma unroll
    for (int v = 0; v < 4; ++v) rv[4 * a + v] = accR[a][v];
  }
  row16_reduce_scatter(rv, c16, out);
  if (I == J) {
    // the column sums of the upper tiles belong to the same rows: the four
    // waves' sums in wave order (FP64), then added to each row's sum
    __syncthreads();  // every wave is done reading the staged tiles
#pragma unroll
    for (int g = 0; g < kSymGroups; ++g) mf[w * kSymTile + g * kWave + lane] = tg[g];
    __syncthreads();
    static_assert((kBlock / kWave) * kSymTile + 2 * kSymTile <= kMfScale, "column sums clear of the row scales");
    double* cs = reinterpret_cast<double*>(mf + (kBlock / kWave) * kSymTile);
    for (int k = tid; k < kSymTile; k += kBlock) {
      double sum = (double)mf[k];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) sum += (double)mf[q * kSymTile + k];
      cs[k] = sum;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = 256 * h + 64 * (c16 >> 2) + 16 * w + 4 * r4 + (c16 & 3);
      const float v = FACTOR ? out[h] * mf[kMfScale + k] : out[h];
      rows_i[k] = (P)((double)v + cs[k]);
    }
    return;
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = 256 * h + 64 * (c16 >> 2) + 16 * w + 4 * r4 + (c16 & 3);
    const float v = FACTOR ? out[h] * mf[kMfScale + k] : out[h];
    rows_i[k] = (P)v;
  }
  if (I != J) {
    __syncthreads();  // every wave is done reading the staged tiles
#pragma unroll
    for (int g = 0; g < kSymGroups; ++g) mf[w * kSymTile + g * kWave + lane] = tg[g];
    __syncthreads();
    for (int k = tid; k < kSymTile; k += kBlock) {
      double s = (double)mf[k];
#pragma unroll
      for (int q = 1; q < kBlock / kWave; ++q) s += (double)mf[q * kSymTile + k];
      rows_j[k] = (P)s;
    }
  }
}

// Screening plans of one operator at up to kMaxLevels drop levels (level 0
// the tightest), for relaxed products in inexact inner solves: level l > 0 is
// used while the value at `sel` is <= below[l] (thresholds decreasing), level
// 0 otherwise (also for a NaN).  The kernel and its finish read the same
// value, so they use the same level; every level writes the partial sums of
// level 0's plan in its own slotted layout (below), which is all the finish
// reads.  A single plan is the case nlev = 1.
constexpr int kMaxLevels = 3;
struct PlanLevels {
  const uint32_t* list[kMaxLevels];
  const int64_t* count[kMaxLevels];
  const float4* tinfo[kMaxLevels];
  const uint32_t* bmask[kMaxLevels];
  const uint32_t* tmask[kMaxLevels];
  const int32_t* pcount[kMaxLevels];
  const int32_t* slots[kMaxLevels];
  double below[kMaxLevels];
  const double* sel;
  int64_t pcap;
  int nlev;
};


// Element l of a kernel-argument array by selects (no dynamic indexing, so
// the array stays in the argument segment).
template <typename P>
__device__ __forceinline__ P level_pick(const P (&a)[kMaxLevels], int l) {
  return l == 0 ? a[0] : (l == 1 ? a[1] : a[2]);
}

__device__ __forceinline__ int plan_level(const PlanLevels& lv) {
  if (lv.sel == nullptr || lv.nlev <= 1) return 0;
  const double v = *lv.sel;
  int l = 0;
  if (v <= lv.below[1]) l = 1;
  if (lv.nlev > 2 && v <= lv.below[2]) l = 2;
  return l;
}

template <int KIND, int DIM>
__global__ __launch_bounds__(kBlock) void kmv_sym_planned_kernel(
    const double* const* __restrict__ points, int64_t off, int64_t n, double scale, const double* __restrict__ x,
    float* __restrict__ part, const PlanLevels lv, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const int lev = plan_level(lv);
  if ((int64_t)blockIdx.x >= *level_pick(lv.count, lev)) return;
  const uint32_t* __restrict__ list = level_pick(lv.list, lev);
  const float4* __restrict__ tinfo = level_pick(lv.tinfo, lev);
  const uint32_t* __restrict__ bmask = level_pick(lv.bmask, lev);
  const uint32_t* __restrict__ tmask = level_pick(lv.tmask, lev);
  __shared__ __attribute__((aligned(16))) float mf[kMfFloats];
  static_assert((kBlock / kWave) * kSymTile <= kMfFloats, "column sums reuse the staging buffer");
  static_assert(sizeof(SymSrc<float>) * kSymGroups <= sizeof(mf), "vector-unit staging fits");
  const int32_t* __restrict__ slots = level_pick(lv.slots, lev);
  const int64_t p = (int64_t)list[blockIdx.x];
  int64_t I, J;
  decode_pair(p, I, J);
  float* rows_i = slot_row(part, I, J, slots[2 * p], lv.pcap);
  float* rows_j = I != J ? slot_row(part, J, I, slots[2 * p + 1], lv.pcap) : rows_i;
  const float4 centre = tinfo[I];
  if (centre.w != 0.0f) {
    sym_mfma_pair<KIND, DIM, float>(points, off, n, scale, x, rows_i, rows_j, p, I, J, centre, tmask, mf);
  } else {
    sym_valu_pair<float, KIND, DIM, false, true, true, float>(points, off, n, scale, 0.0, 0.0, 0.0, x, rows_i, rows_j, p,
                                                       bmask[p], reinterpret_cast<SymSrc<float>*>(mf), nullptr,
                                                       nullptr, nullptr);
  }
}

// ---------------------------------------------------------------------------
// Planned FP64 product (plan built for FP64, no coincidence shift): the FP32
// kernel's 16 x 16 tile screen and wave layout, with the entries from FP64
// coordinate differences and the table-driven FP64 exponential on the vector
// unit.  Lane 16 r + c of wave w evaluates, for each kept tile (a, b) of
// source group g, targets 16 (4 a + w) + 4 r + v (v < 4) against source
// 64 g + 16 b + c: the row sums stay per lane (accR[a][v]) and are reduced
// over the 16 lanes of the lane row at the end, the column sums of a group
// over the wave's 4 lane rows after the group, then over the 4 waves in wave
// order.  Compared with the 128 x 64 block screen of the vector-unit form,
// the tile screen evaluates about a quarter fewer pairs at FP64 drop levels
// on spatially ordered points.  The difference form needs no centring, so
// every tile takes it, compact or not.

// LDS layout (doubles): target coordinates [3][512], source coordinates
// [3][512], target values [512], source values [512], the exp2 table; the
// column sums of the four waves reuse the coordinates.
constexpr int kT64Tc = 0;
constexpr int kT64Sc = 3 * kSymTile;
constexpr int kT64Xt = 6 * kSymTile;
constexpr int kT64Xs = 7 * kSymTile;
constexpr int kT64Tab = 8 * kSymTile;
constexpr int kT64Doubles = 8 * kSymTile + kExp2Table;

// Reduce-scatter over the 16 lanes of a lane row (FP64): v[q] (q < 16) of
// every lane in, the row's sum of v[c] out for lane c.  Round k pairs lane c
// with c ^ 8, c ^ 4, c ^ 2, c ^ 1 and keeps the half of the values selected
// by the lane's bit 3, 2, 1, 0, adding the partner's copy.  Fixed order.
__device__ __forceinline__ double row16_reduce_scatter_f64(const double (&v)[16], int c16) {
  const bool b3 = (c16 & 8) != 0, b2 = (c16 & 4) != 0, b1 = (c16 & 2) != 0, b0 = (c16 & 1) != 0;
  double r1[8], r2[4], r3[2];
#pragma unroll
  for (int q = 0; q < 8; ++q) r1[q] = (b3 ? v[q + 8] : v[q]) + __shfl_xor(b3 ? v[q] : v[q + 8], 8);
#pragma unroll
  for (int q = 0; q < 4; ++q) r2[q] = (b2 ? r1[q + 4] : r1[q]) + __shfl_xor(b2 ? r1[q] : r1[q + 4], 4);
#pragma unroll
  for (int q = 0; q < 2; ++q) r3[q] = (b1 ? r2[q + 2] : r2[q]) + __shfl_xor(b1 ? r2[q] : r2[q + 2], 2);
  return (b0 ? r3[1] : r3[0]) + __shfl_xor(b0 ? r3[0] : r3[1], 1);
}

// Reduce-scatter over the four 16-lane rows of the wave (FP64): lane 16 r + c
// gets the sum over r' of v[r] of lane 16 r' + c; rows paired by r ^ 1, then
// r ^ 2.
__device__ __forceinline__ double rows4_reduce_scatter_f64(const double (&v)[4], int r4) {
  const bool b0 = (r4 & 1) != 0, b1 = (r4 & 2) != 0;
  double t[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) t[q] = (b0 ? v[2 * q + 1] : v[2 * q]) + __shfl_xor(b0 ? v[2 * q] : v[2 * q + 1], 16);
  return (b1 ? t[1] : t[0]) + __shfl_xor(b1 ? t[0] : t[1], 32);
}

template <int KIND, int DIM>
__device__ __forceinline__ void sym_tile64_pair(const double* const* __restrict__ points, int64_t off, int64_t n,
                                                double scale, const double* __restrict__ x,
                                                double* __restrict__ rows_i, double* __restrict__ rows_j, int64_t p,
                                                int64_t I, int64_t J, const uint32_t* __restrict__ tmask,
                                                double* lds) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int c16 = lane & 15, r4 = lane >> 4;
  double* tc = lds + kT64Tc;
  double* sc = lds + kT64Sc;
  double* xt = lds + kT64Xt;
  double* xs = lds + kT64Xs;
  double* tab = lds + kT64Tab;
  // staging: thread t takes points t and t + 256 of both tiles; every global
  // load is issued before the LDS writes (clamped indices, no branches)
  constexpr int kPer = kSymTile / kBlock;
  double pc[2][kPer][3], pv[2][kPer];
#pragma unroll
  for (int side = 0; side < 2; ++side) {  // 0: sources (tile J), 1: targets (tile I)
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int64_t j = (side == 0 ? J : I) * kSymTile + tid + it * kBlock;
      const int64_t jc = j < n ? j : n - 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) pc[side][it][d] = global_load(points[d] + off + jc);
      pv[side][it] = global_load(x + jc);
    }
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int k = tid + it * kBlock;
      const int64_t j = (side == 0 ? J : I) * kSymTile + k;
      double* cd = side == 0 ? sc : tc;
#pragma unroll
      for (int d = 0; d < 3; ++d) cd[d * kSymTile + k] = d < DIM ? pc[side][it][d] * scale : 0.0;
      (side == 0 ? xs : xt)[k] = j < n ? pv[side][it] : 0.0;
    }
  }
  fill_exp2_table(tab);
  uint32_t gmask[kSymGroups];
  const uint32_t* tm = tmask + p * kTmaskWords + w * kSymGroups;
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) gmask[g] = (uint32_t)__builtin_amdgcn_readfirstlane((int)tm[g]);
  __syncthreads();

  // Two passes over the source groups, each for half of the wave's target
  // slabs (a = 4 h + a'), so that only 16 row sums stay in registers; the
  // column sums of a group add the two passes in order.
  double tg[kSymGroups];
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) tg[g] = 0.0;
  auto group_mask = [&](int g) -> uint32_t {
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < kSymGroups; ++q)
      if (q == g) m = gmask[q];
    return m;
  };
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    double accR[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) accR[a][v] = 0.0;
    // bits 8 b + a of a group mask for this pass's slabs a = 4 h .. 4 h + 3
    const uint32_t half_bits = 0x0f0f0f0fu << (4 * h);
#pragma unroll 1
    for (int g = 0; g < kSymGroups; ++g) {
      const uint32_t gm = group_mask(g);
      if ((gm & half_bits) == 0) continue;
      double accC[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int a2 = 0; a2 < 4; ++a2) {
        const int a = 4 * h + a2;
        // bit b: tile (a, b) kept (bit 8 b + a of gm)
        const uint32_t cm =
            ((gm >> a) & 1u) | ((gm >> (7 + a)) & 2u) | ((gm >> (14 + a)) & 4u) | ((gm >> (21 + a)) & 8u);
        if (cm == 0) continue;
        const int t0 = 16 * (4 * a + w) + 4 * r4;
        double pt[4][3], xp[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
          for (int d = 0; d < 3; ++d) pt[v][d] = tc[d * kSymTile + t0 + v];
          xp[v] = xt[t0 + v];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (((cm >> b) & 1u) == 0) continue;
          const int k = g * kWave + 16 * b + c16;
          double q[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) q[d] = sc[d * kSymTile + k];
          const double xq = xs[k];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double d0 = pt[v][0] - q[0];
            double sq = d0 * d0;
            if (DIM > 1) {
              const double d1 = pt[v][1] - q[1];
              sq = fma(d1, d1, sq);
            }
            if (DIM > 2) {
              const double d2 = pt[v][2] - q[2];
              sq = fma(d2, d2, sq);
            }
            const double e = kernel_entry<double, KIND>(sq, tab);
            accR[a2][v] = fma(e, xq, accR[a2][v]);
            accC[b] = fma(e, xp[v], accC[b]);
          }
        }
      }
      const double cs = rows4_reduce_scatter_f64(accC, r4);  // source 16 r + c of the group
#pragma unroll
      for (int q2 = 0; q2 < kSymGroups; ++q2)
        if (q2 == g) tg[q2] += cs;
    }
    // row sums of the pass: lane c of lane row r ends with v[c] = accR[a'][v]
    // for 4 a' + v = c, target 16 (4 (4 h + a') + w) + 4 r + v
    double rv[16];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) rv[4 * a + v] = accR[a][v];
    const double out = row16_reduce_scatter_f64(rv, c16);
    const int k = 16 * (4 * (4 * h + (c16 >> 2)) + w) + 4 * r4 + (c16 & 3);
    rows_i[k] = out;
  }
  if (I != J) {
    __syncthreads();  // every wave is done reading the staged tiles
#pragma unroll
    for (int g = 0; g < kSymGroups; ++g) lds[w * kSymTile + g * kWave + lane] = tg[g];
    __syncthreads();
    for (int k = tid; k < kSymTile; k += kBlock) {
      double sum = lds[k];
#pragma unroll
      for (int q2 = 1; q2 < kBlock / kWave; ++q2) sum += lds[q2 * kSymTile + k];
      rows_j[k] = sum;
    }
  }
}

// FP64 matrix-core form of the planned FP64 tile pair: Gaussian kernel,
// compact target tile.  Coordinates are centred on tile I's box centre;
// v_mfma_f64_16x16x4f64 gives t = |c_s|^2 - 2 c_t.c_s for a 16 x 16 tile
// (A = [c_t, 1], B = [-2 c_s, |c_s|^2]; A's rows permuted, below) in the
// lane layout of the difference form (lane 16 r + c: targets
// 16 (4 a + w) + 4 r + v, v < 4, source 64 g + 16 b + c).  The entry 2^-(|c_t|^2 + t) = 2^-|c_t|^2 2^-t: the factor
// 2^-|c_t|^2 goes into the target values (column sums) and into each row sum
// once, so the loop takes 2^-t only (t >= -|c_t|^2 >= -kCenterRadius2).  The
// rounding of t is at the level of |c|^2 ulps, which the compact target tile
// bounds for every entry that matters.  Same 16 x 16 tile screen, passes and
// reduction order as the difference form.
// LDS layout (doubles): target operand [32 slabs][48] (lane 16 k + c of a
// slab: coordinate k of its centred target c, k < 3), source operand
// [32 slabs][64] (k < 3: -2 x coordinate k; k = 3: |c_s|^2), target values
// times 2^-|c_t|^2, source values, the exp2 table; the column sums of the
// four waves reuse the operands.
constexpr int kM64Tgt = 0;
constexpr int kM64Src = kSlabs * 48;
constexpr int kM64Xt = kM64Src + kSlabs * kWave;
constexpr int kM64Xs = kM64Xt + kSymTile;
constexpr int kM64Tab = kM64Xs + kSymTile;
constexpr int kM64Doubles = kM64Tab + kExp2Table;
static_assert((kBlock / kWave) * kSymTile <= kM64Xt, "column sums reuse the operands");

typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int DIM>
__device__ __forceinline__ void sym_mfma64_pair(const double* const* __restrict__ points, int64_t off, int64_t n,
                                                double scale, const double* __restrict__ x,
                                                double* __restrict__ rows_i, double* __restrict__ rows_j, int64_t p,
                                                int64_t I, int64_t J, float4 centre,
                                                const uint32_t* __restrict__ tmask, double* lds) {
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  const int w = __builtin_amdgcn_readfirstlane(tid / kWave);
  const int c16 = lane & 15, r4 = lane >> 4;
  const double ctr[3] = {(double)centre.x, (double)centre.y, (double)centre.z};
  double* tgt = lds + kM64Tgt;
  double* src = lds + kM64Src;
  double* xt = lds + kM64Xt;
  double* xs = lds + kM64Xs;
  double* tab = lds + kM64Tab;
  // staging: thread t takes points t and t + 256 of both tiles; every global
  // load is issued before the LDS writes (clamped indices, no branches)
  constexpr int kPer = kSymTile / kBlock;
  double pc[2][kPer][3], pv[2][kPer];
#pragma unroll
  for (int side = 0; side < 2; ++side) {  // 0: sources (tile J), 1: targets (tile I)
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int64_t j = (side == 0 ? J : I) * kSymTile + tid + it * kBlock;
      const int64_t jc = j < n ? j : n - 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) pc[side][it][d] = global_load(points[d] + off + jc);
      pv[side][it] = global_load(x + jc);
    }
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
#pragma unroll
    for (int it = 0; it < kPer; ++it) {
      const int k = tid + it * kBlock;
      const int64_t j = (side == 0 ? J : I) * kSymTile + k;
      double c[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int d = 0; d < DIM; ++d) c[d] = pc[side][it][d] * scale - ctr[d];
      const double nc = fma(c[2], c[2], fma(c[1], c[1], c[0] * c[0]));
      const double xv = j < n ? pv[side][it] : 0.0;
      const int slab = k >> 4, cc = k & 15;
      if (side == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) src[slab * kWave + 16 * d + cc] = -2.0 * c[d];
        src[slab * kWave + 48 + cc] = nc;
        xs[k] = xv;
      } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) tgt[slab * 48 + 16 * d + cc] = c[d];
        xt[k] = xv * exp2(-nc);
      }
    }
  }
  fill_exp2_table(tab);
  uint32_t gmask[kSymGroups];
  const uint32_t* tm = tmask + p * kTmaskWords + w * kSymGroups;
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) gmask[g] = (uint32_t)__builtin_amdgcn_readfirstlane((int)tm[g]);
  __syncthreads();

  double tg[kSymGroups];
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) tg[g] = 0.0;
  auto group_mask = [&](int g) -> uint32_t {
    uint32_t m = 0;
#pragma unroll
    for (int q = 0; q < kSymGroups; ++q)
      if (q == g) m = gmask[q];
    return m;
  };
  const f64x4 zero4 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
  for (int h = 0; h < 2; ++h) {
    double accR[4][4], Aop[4];
#pragma unroll
    for (int a2 = 0; a2 < 4; ++a2) {
      // Result i of the FP64 MFMA in lane 16 r + c is row 4 i + r of the
      // tile (column c), so row 4 v + r of A takes target 4 r + v of the
      // slab: lane 16 k + c holds coordinate k of target 4 (c mod 4) + c / 4
      // of slab 4 (4 h + a2) + w (1 at k = 3), and result v of lane 16 r + c
      // is target 4 r + v, as in the difference form.
      const int st = 4 * (4 * h + a2) + w;
      const int tc = 4 * (c16 & 3) + (c16 >> 2);
      const double av = tgt[st * 48 + (r4 < 3 ? 16 * r4 + tc : 0)];
      Aop[a2] = r4 < 3 ? av : 1.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) accR[a2][v] = 0.0;
    }
    const uint32_t half_bits = 0x0f0f0f0fu << (4 * h);
#pragma unroll 1
    for (int g = 0; g < kSymGroups; ++g) {
      const uint32_t gm = group_mask(g);
      if ((gm & half_bits) == 0) continue;
      double accC[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int a2 = 0; a2 < 4; ++a2) {
        const int a = 4 * h + a2;
        const uint32_t cm =
            ((gm >> a) & 1u) | ((gm >> (7 + a)) & 2u) | ((gm >> (14 + a)) & 4u) | ((gm >> (21 + a)) & 8u);
        if (cm == 0) continue;
        const int t0 = 16 * (4 * a + w) + 4 * r4;
        double xp[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) xp[v] = xt[t0 + v];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          if (((cm >> b) & 1u) == 0) continue;
          const double bq = src[(4 * g + b) * kWave + lane];
          const double xq = xs[g * kWave + 16 * b + c16];
          const f64x4 t = __builtin_amdgcn_mfma_f64_16x16x4f64(Aop[a2], bq, zero4, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const double e = exp2_neg(t[v], tab);
            accR[a2][v] = fma(e, xq, accR[a2][v]);
            accC[b] = fma(e, xp[v], accC[b]);
          }
        }
      }
      const double cs = rows4_reduce_scatter_f64(accC, r4);  // source 16 r + c of the group
#pragma unroll
      for (int q2 = 0; q2 < kSymGroups; ++q2)
        if (q2 == g) tg[q2] += cs;
    }
    double rv[16];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int v = 0; v < 4; ++v) rv[4 * a + v] = accR[a][v];
    const double out = row16_reduce_scatter_f64(rv, c16);
    const int k = 16 * (4 * (4 * h + (c16 >> 2)) + w) + 4 * r4 + (c16 & 3);
    // the row's factor 2^-|c_t|^2, from the staged coordinates as in staging
    const int slab = k >> 4, cc = k & 15;
    const double t0c = tgt[slab * 48 + cc], t1c = tgt[slab * 48 + 16 + cc], t2c = tgt[slab * 48 + 32 + cc];
    const double nt = fma(t2c, t2c, fma(t1c, t1c, t0c * t0c));
    rows_i[k] = out * exp2(-nt);
  }
  if (I != J) {
    __syncthreads();  // every wave is done reading the staged tiles
#pragma unroll
    for (int g = 0; g < kSymGroups; ++g) lds[w * kSymTile + g * kWave + lane] = tg[g];
    __syncthreads();
    for (int k = tid; k < kSymTile; k += kBlock) {
      double sum = lds[k];
#pragma unroll
      for (int q2 = 1; q2 < kBlock / kWave; ++q2) sum += lds[q2 * kSymTile + k];
      rows_j[k] = sum;
    }
  }
}

// Four waves per SIMD (128 registers): left to itself the compiler gives the
// Gaussian kernel, which holds both tile forms, 146 and three waves.
template <int KIND, int DIM>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void kmv_sym_planned64_kernel(
    const double* const* __restrict__ points, int64_t off, int64_t n, double scale, const double* __restrict__ x,
    double* __restrict__ part, const uint32_t* __restrict__ list, const int64_t* __restrict__ count,
    const uint32_t* __restrict__ tmask, const float4* __restrict__ tinfo, const int32_t* __restrict__ slots,
    int64_t pcap, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  if ((int64_t)blockIdx.x >= *count) return;
  constexpr int kBufDoubles = kM64Doubles > kT64Doubles ? kM64Doubles : kT64Doubles;
  __shared__ __attribute__((aligned(16))) double buf[kBufDoubles];
  const int64_t p = (int64_t)list[blockIdx.x];
  int64_t I, J;
  decode_pair(p, I, J);
  double* rows_i = slot_row(part, I, J, slots[2 * p], pcap);
  double* rows_j = I != J ? slot_row(part, J, I, slots[2 * p + 1], pcap) : rows_i;
  // FP64 plans mark compact tiles 2, the others 1 (both take the tile screen)
  const float4 centre = tinfo[I];
  if (KIND == MPLAK_KERNEL_GAUSSIAN && centre.w == 2.0f)
    sym_mfma64_pair<DIM>(points, off, n, scale, x, rows_i, rows_j, p, I, J, centre, tmask, buf);
  else
    sym_tile64_pair<KIND, DIM>(points, off, n, scale, x, rows_i, rows_j, p, I, J, tmask, buf);
}

// y_i = sum over partner tiles P of row i's tile X of that pair's partial,
// plus the nugget term.  A workgroup takes 64 consecutive rows (one tile) and
// four partner slices: wave k sums the partners P = k (mod 4) in ascending
// order, then the four slice sums are added in slice order.  With a plan
// (planned product) a wave adds only the listed partners of its slice, in the
// same order; the partials it leaves out are zero, so the result is the same.
// Their rows are contiguous in the plan's slotted layout, so the wave needs
// only their count; kFinishBatch independent partial loads go out at a
// time.  With `dots` the finish also leaves, per workgroup, the partial sums
// of <x,x> and <x,y> over its rows (wave butterfly, so a fixed order):
// dots[b] and dots[gridDim.x + b].

template <typename PT>
__device__ __forceinline__ double sym_partial(const PT* __restrict__ part_i, const PT* __restrict__ part_j,
                                              int64_t X, int64_t P, int64_t t) {
  return P < X ? (double)part_j[sym_pair_index(P, X) * kSymTile + t] : (double)part_i[sym_pair_index(X, P) * kSymTile + t];
}

// Partials in FP64, or in FP32 where the tile sums are FP32 sums anyway (the
// planned FP32 kernel); the finish adds them in FP64 either way.
template <bool LIST, typename PT>
__global__ __launch_bounds__(kBlock) void kmv_sym_finish_kernel(
    int64_t n, int64_t ntiles, const PT* __restrict__ part_i, const PT* __restrict__ part_j, double nugget,
    const double* __restrict__ x, double* __restrict__ y, const PlanLevels lv, int64_t pcap,
    const int32_t* __restrict__ skip, double* __restrict__ dots) {
  if (skip != nullptr && *skip != 0) return;
  const int lev = LIST ? plan_level(lv) : 0;
  const int32_t* __restrict__ pcount = level_pick(lv.pcount, lev);
  __shared__ double red[kFinishSlices][kFinishRows];
  const int r = threadIdx.x % kFinishRows, k = threadIdx.x / kFinishRows;
  double dxx = 0.0, dxy = 0.0;
  for (int64_t row0 = (int64_t)blockIdx.x * kFinishRows; row0 < n; row0 += (int64_t)gridDim.x * kFinishRows) {
    const int64_t i = row0 + r;
    const int64_t X = row0 / kSymTile, t = i % kSymTile;  // rows past n read their tile's padding partials
    // The slice's partner count, and its first kFinishBatch partial rows
    // (slotted layout: the partners' rows in list order; clamped to the
    // slice's pcap rows, no dependence on the count) with x of the row: all in
    // flight together.  A batch past the count adds +0.0, which leaves a sum
    // that started at +0.0 unchanged, so the bits are those of adding one by
    // one in list order.
    const PT* rows = LIST ? part_i + (X * kFinishSlices + k) * pcap * kSymTile + t : nullptr;
    const int craw = LIST ? pcount[X * kFinishSlices + k] : 0;
    double v0[kFinishBatch];
    if (LIST) {
#pragma unroll
      for (int u = 0; u < kFinishBatch; ++u) v0[u] = (double)rows[(u < pcap ? u : pcap - 1) * kSymTile];
    }
    const double xi = (k == 0 && i < n) ? x[i] : 0.0;
    double s = 0.0;
    if (LIST) {
      const int cnt = __builtin_amdgcn_readfirstlane(craw);
#pragma unroll
      for (int u = 0; u < kFinishBatch; ++u) s += u < cnt ? v0[u] : 0.0;
      for (int q = kFinishBatch; q < cnt; q += kFinishBatch) {
        double v[kFinishBatch];
#pragma unroll
        for (int u = 0; u < kFinishBatch; ++u) {
          const int qu = q + u < cnt ? q + u : cnt - 1;
          const double pv = (double)rows[qu * kSymTile];
          v[u] = q + u < cnt ? pv : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kFinishBatch; ++u) s += v[u];
      }
    } else {
      for (int64_t P = k; P < ntiles; P += kFinishSlices) s += sym_partial(part_i, part_j, X, P, t);
    }
    __syncthreads();
    red[k][r] = s;
    __syncthreads();
    if (k == 0 && i < n) {
      double v = red[0][r];
#pragma unroll
      for (int q = 1; q < kFinishSlices; ++q) v += red[q][r];
      if (nugget != 0.0) v = fma(nugget, xi, v);
      y[i] = v;
      dxx = fma(xi, xi, dxx);
      dxy = fma(xi, v, dxy);
    }
  }
  if (dots != nullptr && k == 0) {
    dxx = wave_sum(dxx);
    dxy = wave_sum(dxy);
    if (r == 0) {
      dots[blockIdx.x] = dxx;
      dots[gridDim.x + blockIdx.x] = dxy;
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

// Workgroups of the symmetric finish (one per 64 rows, grid-strided past 2^16).
inline int64_t sym_finish_blocks(int64_t n) {
  const int64_t fb = (n + kFinishRows - 1) / kFinishRows;
  return fb > 65536 ? 65536 : fb;
}

struct SymArgs {
  double scale, thr, drop_s;
  const double* x;
  double *pi, *pj;
  const uint32_t* list;
  const int64_t* count;
  const int32_t* slots;  // planned: the plan's slotted partial rows (level 0)
  int64_t pcap;
  // planned FP32 without coincidence shift: the plans' tile data, for the
  // planned kernel (matrix-core form on compact target tiles); null otherwise
  const PlanLevels* lv;
  const int32_t* skip;
};

template <typename T, int KIND, int DIM, int MODE>
void launch_sym_dim(hipStream_t st, bool coinc, const mplak_kernel_desc* d, const SymArgs& a) {
  const int64_t nt = sym_tiles(d->row_count);
  const dim3 grid((unsigned)(nt * (nt + 1) / 2));
  if constexpr (MODE == kSymPlanned && sizeof(T) == 8) {
    if (a.lv != nullptr && !coinc) {
      hipLaunchKernelGGL((kmv_sym_planned64_kernel<KIND, DIM>), grid, dim3(kBlock), 0, st, d->points, d->row_offset,
                         d->row_count, a.scale, a.x, a.pi, a.lv->list[0], a.lv->count[0], a.lv->tmask[0],
                         a.lv->tinfo[0], a.lv->slots[0], a.lv->pcap, a.skip);
      return;
    }
  }
  if constexpr (MODE == kSymPlanned && sizeof(T) == 4) {
    if (a.lv != nullptr && !coinc) {
      hipLaunchKernelGGL((kmv_sym_planned_kernel<KIND, DIM>), grid, dim3(kBlock), 0, st, d->points, d->row_offset,
                         d->row_count, a.scale, a.x, reinterpret_cast<float*>(a.pi), *a.lv, a.skip);
      return;
    }
  }
  if (coinc)
    hipLaunchKernelGGL((kmv_sym_kernel<T, KIND, DIM, true, MODE>), grid, dim3(kBlock), 0, st, d->points,
                       d->row_offset, d->row_count, a.scale, a.thr, d->coincident_shift, a.drop_s, a.x, a.pi, a.pj,
                       a.list, a.count, a.slots, a.pcap, a.skip);
  else
    hipLaunchKernelGGL((kmv_sym_kernel<T, KIND, DIM, false, MODE>), grid, dim3(kBlock), 0, st, d->points,
                       d->row_offset, d->row_count, a.scale, a.thr, d->coincident_shift, a.drop_s, a.x, a.pi, a.pj,
                       a.list, a.count, a.slots, a.pcap, a.skip);
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
// Planned: `lv` holds the plan(s); more than one level only with FP32 tiles.
int run_sym(const mplak_kernel_desc* d, int precision, int mode, double drop_s, const PlanLevels* lv, int64_t pcap,
            const double* x, double* y, void* work, const int32_t* skip_if, mplak_stream_t s,
            double* dots = nullptr) {
  SymArgs a;
  const bool coinc = d->coincident_shift != 0.0;
  // planned FP32 without coincidence shift: the planned kernel, FP32 partials
  const bool tiles = mode == kSymPlanned && precision == MPLAK_PREC_F32 && lv != nullptr && !coinc;
  // planned FP64 without coincidence shift: the FP64 tile form
  const bool tiles64 = mode == kSymPlanned && precision == MPLAK_PREC_F64 && lv != nullptr && !coinc;
  if (mode == kSymPlanned && lv == nullptr) return MPLAK_ERR_ARG;
  if (!tiles && lv != nullptr && lv->nlev > 1) return MPLAK_ERR_ARG;
  a.lv = tiles || tiles64 ? lv : nullptr;
  a.scale = coord_scale(d->kind, d->length_scale);
  a.thr = (1e-14 * a.scale) * (1e-14 * a.scale);
  a.drop_s = drop_s;
  a.x = x;
  a.list = lv != nullptr ? lv->list[0] : nullptr;
  a.count = lv != nullptr ? lv->count[0] : nullptr;
  a.slots = lv != nullptr ? lv->slots[0] : nullptr;
  a.pcap = lv != nullptr ? lv->pcap : 0;
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
  const int64_t fb = sym_finish_blocks(d->row_count);
  if (tiles) {
    const float* fi = reinterpret_cast<const float*>(pi);
    const float* fj = fi + (size_t)(nt * (nt + 1) / 2) * kSymTile;
    hipLaunchKernelGGL((kmv_sym_finish_kernel<true, float>), dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count,
                       nt, fi, fj, d->nugget, x, y, *lv, pcap, skip_if, dots);
  } else if (lv != nullptr) {
    hipLaunchKernelGGL((kmv_sym_finish_kernel<true, double>), dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count,
                       nt, pi, pj, d->nugget, x, y, *lv, pcap, skip_if, dots);
  } else {
    const PlanLevels none{};
    hipLaunchKernelGGL((kmv_sym_finish_kernel<false, double>), dim3((unsigned)fb), dim3(kBlock), 0, hs,
                       d->row_count, nt, pi, pj, d->nugget, x, y, none, 0, skip_if, dots);
  }
  return launch_status();
}



// ---------------------------------------------------------------------------
// Screening plan.  For every tile pair p = (I, J) of the symmetric product:
//  - bmask[p]: bit 8 w + g set when the 128 x 64 block of wave w's targets
//    of tile I and source group g of tile J is kept (the screened kernel's
//    test, on the same precision-cast coordinates);
//  - tmask[p][w][g]: bit 8 b + a set when the 16 x 16 tile of target slab
//    4 a + w of tile I and source slab 4 g + b of tile J is kept (the
//    matrix-core form's granularity and wave assignment);
// and for every tile the centre of its box and whether it is compact
// (half-diagonal^2 <= kCenterRadius2, FP32 plans only).  Pairs are listed by
// estimated cost, heaviest first (ties by pair index): kept 16 x 16 tiles /
// 32 on compact target tiles, 3 x kept blocks on the others (the vector-unit
// form costs about three times as much per entry).  Pairs with nothing kept
// are left out: no workgroup writes their partials, and the finish's partner
// lists never read them.
// ---------------------------------------------------------------------------

constexpr uint64_t kPlanDropped = 127;  // key class of pairs with nothing kept
constexpr int kPlanKeyBits = 7;         // key classes: 96 - cost (cost 1 .. 96) and kPlanDropped
constexpr int kPlanCounts = 5;          // entries written by mplak_screen_plan_counts

struct PlanLayout {
  size_t count, tinfo, boxes, sboxes, list, bmask, tmask, keys_in, keys_out, tmp, tmp_bytes, pcount, slots,
      part, total;
  int64_t pcap;  // partner slots per (tile, slice)
};

PlanLayout plan_layout(int64_t n) {
  PlanLayout L{};
  const int64_t nt = sym_tiles(n);
  const size_t np = (size_t)(nt * (nt + 1) / 2);
  L.count = 0;
  L.tinfo = 256;
  L.boxes = L.tinfo + align_up(sizeof(float4) * (size_t)nt, 256);
  L.sboxes = L.boxes + align_up(sizeof(double) * 6 * kSymGroups * (size_t)nt, 256);
  L.list = L.sboxes + align_up(sizeof(double) * 6 * kSlabs * (size_t)nt, 256);
  L.bmask = L.list + align_up(sizeof(uint32_t) * np, 256);
  L.tmask = L.bmask + align_up(sizeof(uint32_t) * np, 256);
  L.keys_in = L.tmask + align_up(sizeof(uint32_t) * kTmaskWords * np, 256);
  L.keys_out = L.keys_in + align_up(sizeof(uint64_t) * np, 256);
  L.tmp = L.keys_out + align_up(sizeof(uint64_t) * np, 256);
  size_t tb = 0;
  if (rocprim::radix_sort_keys(nullptr, tb, static_cast<uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                               (unsigned)np, 32, 32 + kPlanKeyBits) != hipSuccess)
    return PlanLayout{};
  L.tmp_bytes = tb;
  L.pcap = (nt + kFinishSlices - 1) / kFinishSlices;
  L.pcount = L.tmp + align_up(tb, 256);
  L.slots = L.pcount + align_up(sizeof(int32_t) * kFinishSlices * (size_t)nt, 256);
  L.part = L.slots + align_up(sizeof(int32_t) * 2 * np, 256);
  // partial rows: by pair (2 per pair) without a plan, slotted with one
  const size_t rows = std::max((size_t)2 * np, (size_t)nt * kFinishSlices * (size_t)L.pcap);
  L.total = L.part + rows * kSymTile * sizeof(double);
  return L;
}

// Level l of `lv` from the plan at `base`.
void plan_level_view(const PlanLayout& L, char* base, int l, PlanLevels& lv) {
  lv.list[l] = reinterpret_cast<const uint32_t*>(base + L.list);
  lv.count[l] = reinterpret_cast<const int64_t*>(base + L.count);
  lv.tinfo[l] = reinterpret_cast<const float4*>(base + L.tinfo);
  lv.bmask[l] = reinterpret_cast<const uint32_t*>(base + L.bmask);
  lv.tmask[l] = reinterpret_cast<const uint32_t*>(base + L.tmask);
  lv.pcount[l] = reinterpret_cast<const int32_t*>(base + L.pcount);
  lv.slots[l] = reinterpret_cast<const int32_t*>(base + L.slots);
  lv.pcap = L.pcap;
}

template <typename T>
__device__ __forceinline__ T min_over_lanes(T v, int from, int to) {
  for (int m = from; m < to; m <<= 1) v = fmin(v, __shfl_xor(v, m));
  return v;
}
template <typename T>
__device__ __forceinline__ T max_over_lanes(T v, int from, int to) {
  for (int m = from; m < to; m <<= 1) v = fmax(v, __shfl_xor(v, m));
  return v;
}

// Bounding boxes of every 16-point slab and 64-point group of every tile,
// from the same precision-cast scaled coordinates the product kernels use
// (padding points are clamped to the last point, as there).
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_box_kernel(const double* const* __restrict__ points, int64_t off,
                                                          int64_t n, double scale, double* __restrict__ boxes,
                                                          double* __restrict__ sboxes) {
  const int lane = threadIdx.x & (kWave - 1);
  for (int k = threadIdx.x; k < kSymTile; k += kBlock) {
    const int64_t j = (int64_t)blockIdx.x * kSymTile + k;
    const int64_t jc = j < n ? j : n - 1;
    T lo[3] = {T(0), T(0), T(0)}, hi[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const T v = (T)(points[d][off + jc] * scale);
      lo[d] = min_over_lanes(v, 1, 16);
      hi[d] = max_over_lanes(v, 1, 16);
    }
    if ((lane & 15) == 0) {
      double* b = sboxes + ((size_t)blockIdx.x * kSlabs + k / 16) * 6;
      for (int d = 0; d < 3; ++d) {
        b[d] = (double)lo[d];
        b[3 + d] = (double)hi[d];
      }
    }
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      lo[d] = min_over_lanes(lo[d], 16, kWave);
      hi[d] = max_over_lanes(hi[d], 16, kWave);
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

// tinfo[t] = (centre of tile t's box, compactness flag: see below).
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_tile_kernel(const double* __restrict__ boxes, int64_t nt,
                                                           float4* __restrict__ tinfo) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= nt) return;
  T lo[3], hi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    lo[d] = (T)boxes[(size_t)t * kSymGroups * 6 + d];
    hi[d] = (T)boxes[(size_t)t * kSymGroups * 6 + 3 + d];
  }
  for (int g = 1; g < kSymGroups; ++g) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      lo[d] = fmin(lo[d], (T)boxes[((size_t)t * kSymGroups + g) * 6 + d]);
      hi[d] = fmax(hi[d], (T)boxes[((size_t)t * kSymGroups + g) * 6 + 3 + d]);
    }
  }
  T r2 = T(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const T half = T(0.5) * (hi[d] - lo[d]);
    r2 = fma(half, half, r2);
  }
  // the matrix-core forms' centring bound.  FP32 plans: 1 compact, 0 not
  // (vector form, block screen); FP64 plans: 2 compact, 1 not (every tile
  // takes the tile screen; the compact ones the matrix-core form)
  const bool compact = r2 <= (T)kCenterRadius2;
  const float flag = sizeof(T) == 8 ? (compact ? 2.0f : 1.0f) : (compact ? 1.0f : 0.0f);
  tinfo[t] = make_float4((float)(T(0.5) * (lo[0] + hi[0])), (float)(T(0.5) * (lo[1] + hi[1])),
                         (float)(T(0.5) * (lo[2] + hi[2])), flag);
}

// tmask[p][w][g] (one thread each).
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_tmask_kernel(const double* __restrict__ sboxes, int64_t np,
                                                            double drop_s, uint32_t* __restrict__ tmask) {
  const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= np * kTmaskWords) return;
  const int64_t p = id / kTmaskWords;
  const int w = (int)(id % kTmaskWords) / kSymGroups, g = (int)(id % kSymGroups);
  int64_t I, J;
  decode_pair(p, I, J);
  // the thread's 8 target slabs and 4 source slabs, each loaded once
  T tb[8][6], sb[4][6];
#pragma unroll
  for (int a = 0; a < 8; ++a) load_box(sboxes + ((size_t)I * kSlabs + 4 * a + w) * 6, tb[a]);
#pragma unroll
  for (int b = 0; b < 4; ++b) load_box(sboxes + ((size_t)J * kSlabs + 4 * g + b) * 6, sb[b]);
  // FP32 plans, diagonal pairs: only tiles with target slab <= source slab;
  // the matrix-core form adds each strictly upper tile to both its rows and
  // its columns (the lower tile is its transpose)
  const bool upper_only = sizeof(T) == 4 && I == J;
  uint32_t bits = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int a = 0; a < 8; ++a)
      if (!(upper_only && 4 * a + w > 4 * g + b) && box_gap2<T, DIM>(tb[a], sb[b]) <= (T)drop_s)
        bits |= 1u << (8 * b + a);
  tmask[id] = bits;
}

// key[p] = class << 32 | p and bmask[p].
template <typename T, int DIM>
__global__ __launch_bounds__(kBlock) void plan_pair_kernel(const double* __restrict__ boxes,
                                                           const float4* __restrict__ tinfo,
                                                           const uint32_t* __restrict__ tmask, int64_t np,
                                                           double drop_s, uint64_t* __restrict__ keys,
                                                           uint32_t* __restrict__ bmask) {
  const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (p >= np) return;
  int64_t I, J;
  decode_pair(p, I, J);
  // the 8 group boxes of each tile, each loaded once
  T ib[kSymGroups][6], jb[kSymGroups][6];
#pragma unroll
  for (int g = 0; g < kSymGroups; ++g) {
    load_box(boxes + ((size_t)I * kSymGroups + g) * 6, ib[g]);
    load_box(boxes + ((size_t)J * kSymGroups + g) * 6, jb[g]);
  }
  int kept = 0;
  uint32_t bits = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    T t[6];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      t[d] = fmin(ib[2 * w][d], ib[2 * w + 1][d]);
      t[3 + d] = fmax(ib[2 * w][3 + d], ib[2 * w + 1][3 + d]);
    }
#pragma unroll
    for (int g = 0; g < kSymGroups; ++g) {
      if (box_gap2<T, DIM>(t, jb[g]) <= (T)drop_s) {
        kept += 1;
        bits |= 1u << (8 * w + g);
      }
    }
  }
  int cost;
  if (tinfo[I].w != 0.0f) {
    int tiles = 0;
    for (int q = 0; q < kTmaskWords; ++q) tiles += __popc(tmask[p * kTmaskWords + q]);
    cost = (tiles + 31) / 32;
  } else {
    cost = 3 * kept;
  }
  const uint64_t cls = cost == 0 ? kPlanDropped : (uint64_t)(96 - cost);
  keys[p] = (cls << 32) | (uint64_t)p;
  bmask[p] = bits;
}

// list = pair indices in key order; *count = pairs with something kept.
__global__ __launch_bounds__(kBlock) void plan_list_kernel(const uint64_t* __restrict__ keys, int64_t np,
                                                           uint32_t* __restrict__ list, int64_t* __restrict__ count) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= np) return;
  list[i] = (uint32_t)(keys[i] & 0xffffffffull);
  const bool kept = (keys[i] >> 32) != kPlanDropped;
  if (kept && (i + 1 == np || (keys[i + 1] >> 32) == kPlanDropped)) *count = i + 1;
  if (i == 0 && !kept) *count = 0;
}

// counts[0] = listed tile pairs, counts[1] = their kept 128 x 64 blocks,
// counts[2] = kept blocks on diagonal pairs (I == J), counts[3] and
// counts[4] = kernel entries the planned FP32 product evaluates off and on
// the diagonal pairs (256 per kept 16 x 16 tile on compact target tiles,
// 8192 per kept block on the others).  One workgroup, fixed order.
__global__ __launch_bounds__(kBlock) void plan_count_kernel(const uint64_t* __restrict__ keys,
                                                            const float4* __restrict__ tinfo,
                                                            const uint32_t* __restrict__ bmask,
                                                            const uint32_t* __restrict__ tmask, int64_t np,
                                                            int64_t* __restrict__ counts) {
  __shared__ int64_t red[kPlanCounts][kBlock];
  int64_t c[kPlanCounts] = {0, 0, 0, 0, 0};
  for (int64_t i = threadIdx.x; i < np; i += kBlock) {
    if ((keys[i] >> 32) == kPlanDropped) continue;
    const int64_t p = (int64_t)(keys[i] & 0xffffffffull);
    int64_t I, J;
    decode_pair(p, I, J);
    const int64_t blocks = __popc(bmask[p]);
    int64_t evals = 8192 * blocks;
    if (tinfo[I].w != 0.0f) {
      evals = 0;
      for (int q = 0; q < kTmaskWords; ++q) evals += 256 * __popc(tmask[p * kTmaskWords + q]);
    }
    c[0] += 1;
    c[1] += blocks;
    c[2] += I == J ? blocks : 0;
    c[I == J ? 4 : 3] += evals;
  }
  for (int q = 0; q < kPlanCounts; ++q) red[q][threadIdx.x] = c[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < kPlanCounts; ++q) {
      int64_t s = 0;
      for (int k = 0; k < kBlock; ++k) s += red[q][k];
      counts[q] = s;
    }
  }
}

// Partner slots of the finish: for tile X and slice k, the tiles P = k (mod 4)
// in ascending order whose pair with X is listed take slots 0, 1, ... of that
// (tile, slice): the slot of each kept pair's two sides and the count per
// (tile, slice) (one wave per (tile, slice); keys are the unsorted per-pair
// keys).
__global__ __launch_bounds__(kBlock) void plan_partner_kernel(const uint64_t* __restrict__ keys, int64_t nt,
                                                              int32_t* __restrict__ pcount,
                                                              int32_t* __restrict__ slots) {
  // one wave per (tile X, slice k): lane l tests P = k + 4 (q0 + l), and the
  // kept ones are ranked in P order (ballot, then each lane's rank)
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t id = ((int64_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
  if (id >= nt * kFinishSlices) return;
  const int64_t X = id / kFinishSlices, k = id % kFinishSlices;
  int32_t c = 0;
  for (int64_t P0 = k; P0 < nt; P0 += kFinishSlices * kWave) {
    const int64_t P = P0 + kFinishSlices * lane;
    bool kept = false;
    if (P < nt) {
      const int64_t p = P < X ? sym_pair_index(P, X) : sym_pair_index(X, P);
      kept = (keys[p] >> 32) != kPlanDropped;
    }
    const uint64_t m = __ballot(kept);
    if (kept) {
      // the pair's partial rows of X go to row `pos` of this list's slots
      const int32_t pos = c + __popcll(m & ((1ull << lane) - 1ull));
      if (P >= X) slots[2 * sym_pair_index(X, P)] = pos;                 // X is tile I of the pair
      if (P <= X) slots[2 * sym_pair_index(P, X) + 1] = P == X ? 0 : pos;  // X is tile J (none on the diagonal)
    }
    c += __popcll(m);
  }
  if (lane == 0) pcount[id] = c;
}

struct PlanPtrs {
  double *boxes, *sboxes;
  float4* tinfo;
  uint32_t *bmask, *tmask;
  uint64_t* keys;
};

template <typename T, int DIM>
void launch_plan_geometry_dim(hipStream_t st, const mplak_kernel_desc* d, double scale, double drop_s,
                              const PlanPtrs& q) {
  const int64_t nt = sym_tiles(d->row_count);
  const int64_t np = nt * (nt + 1) / 2;
  auto groups = [](int64_t threads) { return dim3((unsigned)((threads + kBlock - 1) / kBlock)); };
  hipLaunchKernelGGL((plan_box_kernel<T, DIM>), dim3((unsigned)nt), dim3(kBlock), 0, st, d->points, d->row_offset,
                     d->row_count, scale, q.boxes, q.sboxes);
  hipLaunchKernelGGL((plan_tile_kernel<T, DIM>), groups(nt), dim3(kBlock), 0, st, q.boxes, nt, q.tinfo);
  hipLaunchKernelGGL((plan_tmask_kernel<T, DIM>), groups(np * kTmaskWords), dim3(kBlock), 0, st, q.sboxes, np,
                     drop_s, q.tmask);
  hipLaunchKernelGGL((plan_pair_kernel<T, DIM>), groups(np), dim3(kBlock), 0, st, q.boxes, q.tinfo, q.tmask, np,
                     drop_s, q.keys, q.bmask);
}

template <typename T>
void launch_plan_geometry(hipStream_t st, const mplak_kernel_desc* d, double scale, double drop_s,
                          const PlanPtrs& q) {
  switch (d->dim) {
    case 1: launch_plan_geometry_dim<T, 1>(st, d, scale, drop_s, q); break;
    case 2: launch_plan_geometry_dim<T, 2>(st, d, scale, drop_s, q); break;
    default: launch_plan_geometry_dim<T, 3>(st, d, scale, drop_s, q); break;
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
  return run_sym(d, precision, kSymScreened, drop_distance2(d->kind, drop_log2), nullptr, 0, x, y, work, skip_if, s);
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
  const double scale = coord_scale(d->kind, d->length_scale);
  const double drop_s = drop_distance2(d->kind, drop_log2);
  hipStream_t hs = as_stream(s);
  PlanPtrs q;
  q.boxes = boxes;
  q.sboxes = reinterpret_cast<double*>(base + L.sboxes);
  q.tinfo = reinterpret_cast<float4*>(base + L.tinfo);
  q.bmask = reinterpret_cast<uint32_t*>(base + L.bmask);
  q.tmask = reinterpret_cast<uint32_t*>(base + L.tmask);
  q.keys = kin;
  if (precision == MPLAK_PREC_F64)
    launch_plan_geometry<double>(hs, d, scale, drop_s, q);
  else
    launch_plan_geometry<float>(hs, d, scale, drop_s, q);
  st = launch_status();
  if (st != MPLAK_OK) return st;
  hipLaunchKernelGGL(plan_partner_kernel, dim3((unsigned)((nt * kFinishSlices * kWave + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, hs, kin, nt,
                     reinterpret_cast<int32_t*>(base + L.pcount), reinterpret_cast<int32_t*>(base + L.slots));
  st = launch_status();
  if (st != MPLAK_OK) return st;
  size_t tb = L.tmp_bytes;
  if (rocprim::radix_sort_keys(base + L.tmp, tb, kin, kout, (unsigned)np, 32, 32 + kPlanKeyBits, hs) != hipSuccess)
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
                     reinterpret_cast<const uint64_t*>(base + L.keys_out),
                     reinterpret_cast<const float4*>(base + L.tinfo), reinterpret_cast<const uint32_t*>(base + L.bmask),
                     reinterpret_cast<const uint32_t*>(base + L.tmask), nt * (nt + 1) / 2, counts);
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
  PlanLevels lv{};
  plan_level_view(L, base, 0, lv);
  lv.nlev = 1;
  return run_sym(d, precision, kSymPlanned, drop_distance2(d->kind, drop_log2), &lv, L.pcap, x, y, base + L.part,
                 skip_if, s);
}

int64_t mplak_planned_dots_count(int64_t n) { return n > 0 ? mplak::sym_finish_blocks(n) : 0; }

int mplak_kernel_matvec_planned_dots(const mplak_kernel_desc* d, int precision, double drop_log2, void* plan,
                                     size_t plan_bytes, const double* x, double* y, double* dots,
                                     const int32_t* skip_if, mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision != MPLAK_PREC_F64 && precision != MPLAK_PREC_F32) return MPLAK_ERR_ARG;
  if (!(drop_log2 > 0.0) || drop_log2 > 2000.0) return MPLAK_ERR_ARG;
  if (!sym_eligible(d) || d->row_count == 0) return MPLAK_ERR_ARG;
  if (x == nullptr || y == nullptr || dots == nullptr) return MPLAK_ERR_ARG;
  const PlanLayout L = plan_layout(d->row_count);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (plan == nullptr || plan_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  char* base = static_cast<char*>(plan);
  PlanLevels lv{};
  plan_level_view(L, base, 0, lv);
  lv.nlev = 1;
  return run_sym(d, precision, kSymPlanned, drop_distance2(d->kind, drop_log2), &lv, L.pcap, x, y, base + L.part,
                 skip_if, s, dots);
}

int mplak_kernel_matvec_planned_relaxed(const mplak_kernel_desc* d, int nlev, const double* drop_log2,
                                        void* const* plans, size_t plan_bytes, const double* sel,
                                        const double* sel_below, const double* x, double* y, double* dots,
                                        const int32_t* skip_if, mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (nlev < 1 || nlev > kMaxLevels || drop_log2 == nullptr || plans == nullptr) return MPLAK_ERR_ARG;
  if (nlev > 1 && (sel == nullptr || sel_below == nullptr)) return MPLAK_ERR_ARG;
  if (!sym_eligible(d) || d->row_count == 0 || d->coincident_shift != 0.0) return MPLAK_ERR_ARG;
  if (x == nullptr || y == nullptr) return MPLAK_ERR_ARG;
  const PlanLayout L = plan_layout(d->row_count);
  if (L.total == 0) return MPLAK_ERR_HIP;
  if (plan_bytes < L.total) return MPLAK_ERR_WORKSPACE;
  PlanLevels lv{};
  for (int l = 0; l < nlev; ++l) {
    if (!(drop_log2[l] > 0.0) || drop_log2[l] > 2000.0 || plans[l] == nullptr) return MPLAK_ERR_ARG;
    if (l > 0 && !(drop_log2[l] < drop_log2[l - 1] && sel_below[l - 1] < (l > 1 ? sel_below[l - 2] : INFINITY)))
      return MPLAK_ERR_ARG;
    plan_level_view(L, static_cast<char*>(plans[l]), l, lv);
    lv.below[l] = l > 0 ? sel_below[l - 1] : INFINITY;
  }
  lv.sel = nlev > 1 ? sel : nullptr;
  lv.nlev = nlev;
  return run_sym(d, MPLAK_PREC_F32, kSymPlanned, drop_distance2(d->kind, drop_log2[0]), &lv, L.pcap, x, y,
                 static_cast<char*>(plans[0]) + L.part, skip_if, s, dots);
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
    return run_sym(d, precision, kSymExact, 0.0, nullptr, 0, x, y, work, skip_if, s);

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

int64_t mplak_kernel_matvec_noalloc_max_bytes(void) { return (int64_t)kNoallocScratchBytes; }

int mplak_kernel_matvec_noalloc(const mplak_kernel_desc* d, int precision, const double* x, double* y,
                                const int32_t* skip_if, mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (mplak_kernel_matvec_workspace_bytes(d) > kNoallocScratchBytes) return MPLAK_ERR_TOO_LARGE;
  void* work = nullptr;
  st = noalloc_scratch(&work);
  if (st != MPLAK_OK) return st;
  // the same kernels and launch shapes as mplak_kernel_matvec, on the
  // device's fixed scratch: every partial the finish reads is written first,
  // so nothing needs clearing between calls
  return mplak_kernel_matvec(d, precision, x, y, work, kNoallocScratchBytes, skip_if, s);
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
