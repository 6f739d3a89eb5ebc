// Multi-right-hand-side kernel-matrix products Y = K V (V is N x R) on gfx950.
//
// No counterpart in the reference (its only product is single-vector,
// src/kernel.cu:87-171).  Two kernels:
//
// * VALU kernel (FP64, and FP32 for narrow V): one target per thread, a tile
//   of RT columns of V accumulated per thread, sources streamed through LDS.
//   Each kernel entry is computed once per column tile.
//
// * MFMA kernel (BF16, split BF16 "bf16x3", FP32): a wave owns 64 targets
//   (two 32-row MFMA blocks) and all R columns.  For every step of 16 sources
//   each lane evaluates its 16 kernel entries in FP32 directly in the
//   A-operand layout of v_mfma_f32_32x32x16_bf16 (lane l: target l%32 of the
//   row block, sources 8*(l/32) .. +7 of the step), so entries never leave
//   registers.  V is packed once per call into an MFMA-native layout (one
//   16-byte B operand per lane per step and 32-column block, contiguous per
//   wave).  bf16x3 splits entries and V into hi + lo BF16 parts and adds the
//   three leading products (K_hi V_hi + K_hi V_lo + K_lo V_hi).  FP32 uses
//   v_mfma_f32_32x32x2_f32 with eight K=2 products per step.
//
// Both kernels split the sources into slabs (FP32 accumulation within a slab,
// FP64 partials per slab) and a finish kernel sums the slabs in order and adds
// the nugget term: deterministic for a given shape.

#include <math.h>

#include <algorithm>
#include <type_traits>

#include "common.hpp"
#include "entry.hpp"

namespace mplak {
namespace {

constexpr int kChunk = 256;          // sources per LDS chunk
constexpr int kMaxDim = 16;
constexpr int kTargetBlocks = 2048;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct Slabs {
  int64_t tiles, slabs, slab_cols;
};

inline Slabs slab_plan(int64_t rows, int64_t cols, int tile_rows, int64_t col_tiles) {
  Slabs p;
  p.tiles = (rows + tile_rows - 1) / tile_rows;
  int64_t ks = (kTargetBlocks + p.tiles * col_tiles - 1) / (p.tiles * col_tiles);
  const int64_t chunks = (cols + kChunk - 1) / kChunk;
  if (ks > chunks) ks = chunks;
  if (ks < 1) ks = 1;
  if (ks > 65535) ks = 65535;
  const int64_t per = (chunks + ks - 1) / ks;
  p.slab_cols = per * kChunk;
  p.slabs = (cols + p.slab_cols - 1) / p.slab_cols;
  return p;
}

// ---------------------------------------------------------------------------
// VALU kernel.  grid = (target tiles of 256, slabs, column tiles of RT).
// part layout: part[(slab * R + r) * rows + i].

template <typename T, int KIND, int DIM, int RT>
__global__ __launch_bounds__(kBlock) void mrhs_valu_kernel(
    const double* const* __restrict__ points, int dim_rt, int64_t row_off, int64_t rows, int64_t col_off,
    int64_t cols, int64_t slab_cols, double scale, int64_t R, const double* __restrict__ V, int64_t ldv,
    double* __restrict__ part, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  constexpr int MD = DIM > 0 ? DIM : kMaxDim;
  const int dim = DIM > 0 ? DIM : dim_rt;
  __shared__ T lp[MD][kChunk];
  __shared__ T lv[RT][kChunk];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the chunk barrier below
  const int tid = threadIdx.x;
  int64_t i = (int64_t)blockIdx.x * kBlock + tid;
  const bool active = i < rows;
  if (!active) i = rows - 1;
  const int64_t r0 = (int64_t)blockIdx.z * RT;
  T pi[MD];
#pragma unroll
  for (int d = 0; d < MD; ++d) pi[d] = d < dim ? (T)(points[d][row_off + i] * scale) : T(0);
  const int64_t c0 = (int64_t)blockIdx.y * slab_cols;
  const int64_t c1 = (c0 + slab_cols < cols) ? c0 + slab_cols : cols;
  double acc[RT];
#pragma unroll
  for (int q = 0; q < RT; ++q) acc[q] = 0.0;
  for (int64_t cb = c0; cb < c1; cb += kChunk) {
    __syncthreads();
    const int64_t j = cb + tid;
#pragma unroll
    for (int d = 0; d < MD; ++d) lp[d][tid] = (j < c1 && d < dim) ? (T)(points[d][col_off + j] * scale) : T(0);
#pragma unroll
    for (int q = 0; q < RT; ++q) lv[q][tid] = (j < c1 && r0 + q < R) ? (T)V[(r0 + q) * ldv + j] : T(0);
    __syncthreads();
    T cacc[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) cacc[q] = T(0);
#pragma unroll 2
    for (int jj = 0; jj < kChunk; ++jj) {
      T s = T(0);
#pragma unroll
      for (int d = 0; d < MD; ++d)
        if (d < dim) {
          const T df = pi[d] - lp[d][jj];
          s = fma(df, df, s);
        }
      const T k = kernel_entry<T, KIND>(s, exp_tab);
#pragma unroll
      for (int q = 0; q < RT; ++q) cacc[q] = fma(k, lv[q][jj], cacc[q]);
    }
#pragma unroll
    for (int q = 0; q < RT; ++q) acc[q] += (double)cacc[q];
  }
  if (active) {
#pragma unroll
    for (int q = 0; q < RT; ++q)
      if (r0 + q < R) part[((int64_t)blockIdx.y * R + r0 + q) * rows + i] = acc[q];
  }
}

// ---------------------------------------------------------------------------
// MFMA kernel.
//
// Packed V (built by pack_v_kernel), per 16-source step s and 32-column block
// cb, 64 lanes x one operand each:
//   BF16 / BF16X3 : 8 bf16 per lane, lane l holds V[16 s + 8 (l/32) + t][32 cb + l%32]
//   FP32          : 8 floats per lane, lane l holds V[16 s + 2 t + l/32][32 cb + l%32]
// (t = 0..7).  BF16X3 stores the hi part and then, in a second array of the
// same size, the lo part.

enum MfmaMode { kModeBf16 = 0, kModeBf16x3 = 1, kModeF32 = 2 };

constexpr int kWaveTargets = 64;                    // two 32-row MFMA blocks
constexpr int kMfmaTile = kWaveTargets * (kBlock / kWave);  // 256 targets per workgroup
constexpr int64_t kMfmaGroup = 64;  // columns per MFMA launch (two 32-column blocks)

__global__ __launch_bounds__(kBlock) void pack_v_kernel(const double* __restrict__ V, int64_t ldv, int64_t R,
                                                        int64_t cols, int64_t steps, int64_t cblocks, int mode,
                                                        void* __restrict__ out) {
  // one thread per (step, cblock, lane)
  const int64_t total = steps * cblocks * kWave;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int lane = (int)(g % kWave);
    const int64_t sc = g / kWave;
    const int64_t cb = sc % cblocks;
    const int64_t s = sc / cblocks;
    const int64_t r = cb * 32 + (lane % 32);
    if (mode == kModeF32) {
      float* o = static_cast<float*>(out) + g * 8;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int64_t j = 16 * s + 2 * t + lane / 32;
        o[t] = (j < cols && r < R) ? (float)V[r * ldv + j] : 0.0f;
      }
    } else {
      __bf16* hi = static_cast<__bf16*>(out) + g * 8;
      __bf16* lo = hi + total * 8;
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int64_t j = 16 * s + 8 * (lane / 32) + t;
        const float v = (j < cols && r < R) ? (float)V[r * ldv + j] : 0.0f;
        const __bf16 h = (__bf16)v;
        hi[t] = h;
        if (mode == kModeBf16x3) lo[t] = (__bf16)(v - (float)h);
      }
    }
  }
}

// DIM: the points' dimension (1..4); coordinates past it are not evaluated.
// Squared distance over the first DIM coordinates.
template <int DIM>
__device__ __forceinline__ float dist2(const float (&p)[4], const float4& q) {
  const float dx = p[0] - q.x;
  float s = dx * dx;
  if (DIM > 1) {
    const float dy = p[1] - q.y;
    s = fmaf(dy, dy, s);
  }
  if (DIM > 2) {
    const float dz = p[2] - q.z;
    s = fmaf(dz, dz, s);
  }
  if (DIM > 3) {
    const float dw = p[3] - q.w;
    s = fmaf(dw, dw, s);
  }
  return s;
}

template <int KIND, int MODE, int CB, int DIM>
__global__ __launch_bounds__(kBlock) void mrhs_mfma_kernel(
    const double* const* __restrict__ points, int dim, int64_t row_off, int64_t rows, int64_t col_off,
    int64_t cols, int64_t slab_cols, double scale, int64_t R, const void* __restrict__ vpk, int64_t vpk_total,
    double* __restrict__ part, const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  __shared__ float4 lp[kChunk];  // scaled source coordinates (x, y, z, w) for dim <= 4
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid / kWave;
  const int64_t base = (int64_t)blockIdx.x * kMfmaTile + wave * kWaveTargets;

  // target coordinates for the two row blocks of this lane
  float pa[2][4];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    int64_t i = base + h * 32 + (lane % 32);
    if (i >= rows) i = rows - 1;
#pragma unroll
    for (int d = 0; d < 4; ++d) pa[h][d] = d < DIM ? (float)(points[d][row_off + i] * scale) : 0.0f;
  }

  f32x16 acc[2][CB];
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[h][c][v] = 0.0f;

  const int64_t c0 = (int64_t)blockIdx.y * slab_cols;
  const int64_t c1 = (c0 + slab_cols < cols) ? c0 + slab_cols : cols;
  for (int64_t cbase = c0; cbase < c1; cbase += kChunk) {
    __syncthreads();
    {
      const int64_t j = cbase + tid;
      float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
      if (j < c1) {
        q.x = (float)(points[0][col_off + j] * scale);
        if (DIM > 1) q.y = (float)(points[1][col_off + j] * scale);
        if (DIM > 2) q.z = (float)(points[2][col_off + j] * scale);
        if (DIM > 3) q.w = (float)(points[3][col_off + j] * scale);
      } else {
        q.x = 1.0e4f;  // far away: entries of padded sources underflow to 0
      }
      lp[tid] = q;
    }
    __syncthreads();
    const int64_t step0 = cbase / 16;
#pragma unroll 1
    for (int st = 0; st < kChunk / 16; ++st) {
      const int64_t step = step0 + st;
      // ---- kernel entries in A-operand layout
      float e[2][8];
      if (MODE == kModeF32) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float4 q = lp[st * 16 + 2 * t + lane / 32];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            e[h][t] = kernel_entry<float, KIND>(dist2<DIM>(pa[h], q), nullptr);
          }
        }
      } else {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float4 q = lp[st * 16 + 8 * (lane / 32) + t];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            e[h][t] = kernel_entry<float, KIND>(dist2<DIM>(pa[h], q), nullptr);
          }
        }
      }
      // ---- contraction with the packed V operands
      if (MODE == kModeF32) {
        const float* vp = static_cast<const float*>(vpk);
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const float4* src = reinterpret_cast<const float4*>(vp + ((step * CB + c) * kWave + lane) * 8);
          const float4 b0 = src[0], b1 = src[1];
          const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
          for (int t = 0; t < 8; ++t) {
#pragma unroll
            for (int h = 0; h < 2; ++h)
              acc[h][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(e[h][t], bv[t], acc[h][c], 0, 0, 0);
          }
        }
      } else {
        bf16x8 ahi[2], alo[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const __bf16 hb = (__bf16)e[h][t];
            ahi[h][t] = hb;
            if (MODE == kModeBf16x3) alo[h][t] = (__bf16)(e[h][t] - (float)hb);
          }
        const __bf16* vh = static_cast<const __bf16*>(vpk);
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const int64_t off = ((step * CB + c) * kWave + lane) * 8;
          const bf16x8 bhi = *reinterpret_cast<const bf16x8*>(vh + off);
#pragma unroll
          for (int h = 0; h < 2; ++h) acc[h][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[h], bhi, acc[h][c], 0, 0, 0);
          if (MODE == kModeBf16x3) {
            const bf16x8 blo = *reinterpret_cast<const bf16x8*>(vh + vpk_total * 8 + off);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              acc[h][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[h], blo, acc[h][c], 0, 0, 0);
              acc[h][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[h], bhi, acc[h][c], 0, 0, 0);
            }
          }
        }
      }
    }
  }
  // ---- write FP64 slab partials.  32x32 accumulator layout: lane l holds
  // column l%32 and rows 8*(v/4) + 4*(l/32) + v%4, v = 0..15.
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int64_t r = c * 32 + (lane % 32);
      if (r >= R) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t i = base + h * 32 + 8 * (v / 4) + 4 * (lane / 32) + (v % 4);
        if (i < rows) part[((int64_t)blockIdx.y * R + r) * rows + i] = (double)acc[h][c][v];
      }
    }
}

// Y[i + r ldy] = sum over slabs (slab order) + nugget * V at the global diagonal.
__global__ __launch_bounds__(kBlock) void mrhs_finish_kernel(int64_t rows, int64_t R, int64_t slabs,
                                                             const double* __restrict__ part, int64_t row_off,
                                                             int64_t col_off, int64_t cols, double nugget,
                                                             const double* __restrict__ V, int64_t ldv,
                                                             double* __restrict__ Y, int64_t ldy,
                                                             const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const int64_t total = rows * R;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
    const int64_t i = g % rows, r = g / rows;
    double s = part[r * rows + i];
    for (int64_t k = 1; k < slabs; ++k) s += part[(k * R + r) * rows + i];
    const int64_t jd = row_off + i - col_off;
    if (nugget != 0.0 && jd >= 0 && jd < cols) s = fma(nugget, V[r * ldv + jd], s);
    Y[r * ldy + i] = s;
  }
}

// ---------------------------------------------------------------------------
// host-side planning and dispatch

struct MrhsPlan {
  bool mfma;
  int mode;       // MfmaMode when mfma
  int rt;         // VALU column tile
  int64_t cblocks;
  Slabs s;
  size_t vpk_bytes, part_bytes;
  int64_t steps;
};

inline int valu_rt(int64_t R) {
  if (R <= 1) return 1;
  if (R <= 2) return 2;
  if (R <= 4) return 4;
  if (R <= 8) return 8;
  return 16;
}

inline MrhsPlan make_plan(const mplak_kernel_desc* d, int precision, int64_t R) {
  MrhsPlan p{};
  const bool wide = R >= 16;
  p.mfma = (precision == MPLAK_PREC_BF16 || precision == MPLAK_PREC_BF16X3 ||
            (precision == MPLAK_PREC_F32 && wide)) && d->dim <= 4;
  if (p.mfma) {
    p.mode = precision == MPLAK_PREC_BF16 ? kModeBf16 : (precision == MPLAK_PREC_BF16X3 ? kModeBf16x3 : kModeF32);
    p.cblocks = (R + 31) / 32;
    p.s = slab_plan(d->row_count, d->col_count, kMfmaTile, 1);
    const int64_t cols_pad = p.s.slabs * p.s.slab_cols;
    p.steps = cols_pad / 16;
    const int64_t total = p.steps * p.cblocks * kWave;
    const size_t elem = p.mode == kModeF32 ? 4 : 2;
    p.vpk_bytes = align_up((size_t)total * 8 * elem * (p.mode == kModeBf16x3 ? 2 : 1), 256);
  } else {
    p.rt = valu_rt(R);
    const int64_t ctiles = (R + p.rt - 1) / p.rt;
    p.s = slab_plan(d->row_count, d->col_count, kBlock, ctiles);
    p.vpk_bytes = 0;
  }
  p.part_bytes = align_up((size_t)p.s.slabs * (size_t)R * (size_t)d->row_count * sizeof(double), 256);
  return p;
}

template <typename T, int KIND, int DIM>
void launch_valu_rt(int rt, dim3 grid, hipStream_t st, const mplak_kernel_desc* d, const MrhsPlan& p, double scale,
                    int64_t R, const double* V, int64_t ldv, double* part, const int32_t* skip) {
#define MPLAK_VALU_CASE(RTV)                                                                                \
  case RTV:                                                                                                 \
    hipLaunchKernelGGL((mrhs_valu_kernel<T, KIND, DIM, RTV>), grid, dim3(kBlock), 0, st, d->points, d->dim, \
                       d->row_offset, d->row_count, d->col_offset, d->col_count, p.s.slab_cols, scale, R, V, \
                       ldv, part, skip);                                                                    \
    break;
  switch (rt) {
    MPLAK_VALU_CASE(1)
    MPLAK_VALU_CASE(2)
    MPLAK_VALU_CASE(4)
    MPLAK_VALU_CASE(8)
    default:
      MPLAK_VALU_CASE(16)
  }
#undef MPLAK_VALU_CASE
}

template <typename T, int KIND>
void launch_valu(const mplak_kernel_desc* d, const MrhsPlan& p, hipStream_t st, double scale, int64_t R,
                 const double* V, int64_t ldv, double* part, const int32_t* skip) {
  const dim3 grid((unsigned)p.s.tiles, (unsigned)p.s.slabs, (unsigned)((R + p.rt - 1) / p.rt));
  switch (d->dim) {
    case 1: launch_valu_rt<T, KIND, 1>(p.rt, grid, st, d, p, scale, R, V, ldv, part, skip); break;
    case 2: launch_valu_rt<T, KIND, 2>(p.rt, grid, st, d, p, scale, R, V, ldv, part, skip); break;
    case 3: launch_valu_rt<T, KIND, 3>(p.rt, grid, st, d, p, scale, R, V, ldv, part, skip); break;
    default: launch_valu_rt<T, KIND, 0>(p.rt, grid, st, d, p, scale, R, V, ldv, part, skip); break;
  }
}

template <int KIND, int MODE>
void launch_mfma_mode(const mplak_kernel_desc* d, const MrhsPlan& p, hipStream_t st, double scale, int64_t R,
                      const void* vpk, double* part, const int32_t* skip) {
  const dim3 grid((unsigned)p.s.tiles, (unsigned)p.s.slabs);
  const int64_t total = p.steps * p.cblocks * kWave;
  auto go = [&](auto cb_tag, auto dim_tag) {
    constexpr int CBV = decltype(cb_tag)::value, DIMV = decltype(dim_tag)::value;
    hipLaunchKernelGGL((mrhs_mfma_kernel<KIND, MODE, CBV, DIMV>), grid, dim3(kBlock), 0, st, d->points, d->dim,
                       d->row_offset, d->row_count, d->col_offset, d->col_count, p.s.slab_cols, scale, R, vpk,
                       total, part, skip);
  };
  auto with_dim = [&](auto cb_tag) {
    switch (d->dim) {
      case 1: go(cb_tag, std::integral_constant<int, 1>{}); break;
      case 2: go(cb_tag, std::integral_constant<int, 2>{}); break;
      case 3: go(cb_tag, std::integral_constant<int, 3>{}); break;
      default: go(cb_tag, std::integral_constant<int, 4>{}); break;
    }
  };
  if (p.cblocks == 1)
    with_dim(std::integral_constant<int, 1>{});
  else
    with_dim(std::integral_constant<int, 2>{});
}

template <int KIND>
void launch_mfma(const mplak_kernel_desc* d, const MrhsPlan& p, hipStream_t st, double scale, int64_t R,
                 const void* vpk, double* part, const int32_t* skip) {
  if (p.mode == kModeBf16)
    launch_mfma_mode<KIND, kModeBf16>(d, p, st, scale, R, vpk, part, skip);
  else if (p.mode == kModeBf16x3)
    launch_mfma_mode<KIND, kModeBf16x3>(d, p, st, scale, R, vpk, part, skip);
  else
    launch_mfma_mode<KIND, kModeF32>(d, p, st, scale, R, vpk, part, skip);
}

int check_desc(const mplak_kernel_desc* d) {
  if (d == nullptr || d->points == nullptr) return MPLAK_ERR_ARG;
  if (d->dim < 1 || d->dim > kMaxDim) return MPLAK_ERR_ARG;
  if (d->kind != MPLAK_KERNEL_GAUSSIAN && d->kind != MPLAK_KERNEL_MATERN32) return MPLAK_ERR_ARG;
  if (!(d->length_scale > 0.0)) return MPLAK_ERR_ARG;
  if (d->row_count < 0 || d->col_count < 0 || d->row_offset < 0 || d->col_offset < 0) return MPLAK_ERR_ARG;
  return MPLAK_OK;
}

}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

size_t mplak_kernel_mrhs_workspace_bytes(const mplak_kernel_desc* d, int precision, int64_t nrhs) {
  if (check_desc(d) != MPLAK_OK || nrhs <= 0 || d->row_count <= 0 || d->col_count <= 0) return 0;
  // the MFMA path works on groups of at most kMfmaGroup columns
  const MrhsPlan p = make_plan(d, precision, nrhs);
  const MrhsPlan g = p.mfma ? make_plan(d, precision, std::min<int64_t>(nrhs, kMfmaGroup)) : p;
  return g.vpk_bytes + g.part_bytes;
}

int mplak_kernel_mrhs(const mplak_kernel_desc* d, int precision, int64_t nrhs, const double* V, int64_t ldv,
                      double* Y, int64_t ldy, void* work, size_t work_bytes, const int32_t* skip_if,
                      mplak_stream_t s) {
  int st = check_desc(d);
  if (st != MPLAK_OK) return st;
  if (precision < MPLAK_PREC_F64 || precision > MPLAK_PREC_BF16X3) return MPLAK_ERR_ARG;
  if (nrhs < 0) return MPLAK_ERR_ARG;
  if (nrhs == 0 || d->row_count == 0) return MPLAK_OK;
  if (Y == nullptr || ldy < d->row_count) return MPLAK_ERR_ARG;
  if (d->col_count == 0) {
    for (int64_t r = 0; r < nrhs; ++r) {
      st = mplak_fill_f64(Y + r * ldy, d->row_count, 0.0, s);
      if (st != MPLAK_OK) return st;
    }
    return MPLAK_OK;
  }
  if (V == nullptr || ldv < d->col_count) return MPLAK_ERR_ARG;
  if (work == nullptr || work_bytes < mplak_kernel_mrhs_workspace_bytes(d, precision, nrhs)) return MPLAK_ERR_WORKSPACE;
  hipStream_t hs = as_stream(s);
  const double scale = coord_scale(d->kind, d->length_scale);
  const MrhsPlan whole = make_plan(d, precision, nrhs);
  const int64_t group = whole.mfma ? kMfmaGroup : nrhs;
  for (int64_t r0 = 0; r0 < nrhs; r0 += group) {
    const int64_t R = std::min(group, nrhs - r0);
    const MrhsPlan p = make_plan(d, precision, R);
    const double* Vg = V + r0 * ldv;
    double* Yg = Y + r0 * ldy;
    char* w = static_cast<char*>(work);
    double* part = reinterpret_cast<double*>(w + p.vpk_bytes);
    if (p.mfma) {
      const int64_t total = p.steps * p.cblocks * kWave;
      int64_t pb = (total + kBlock - 1) / kBlock;
      if (pb > 8192) pb = 8192;
      hipLaunchKernelGGL(pack_v_kernel, dim3((unsigned)pb), dim3(kBlock), 0, hs, Vg, ldv, R, d->col_count, p.steps,
                         p.cblocks, p.mode, (void*)w);
      st = launch_status();
      if (st != MPLAK_OK) return st;
      if (d->kind == MPLAK_KERNEL_GAUSSIAN)
        launch_mfma<MPLAK_KERNEL_GAUSSIAN>(d, p, hs, scale, R, w, part, skip_if);
      else
        launch_mfma<MPLAK_KERNEL_MATERN32>(d, p, hs, scale, R, w, part, skip_if);
    } else {
      if (precision == MPLAK_PREC_F64) {
        if (d->kind == MPLAK_KERNEL_GAUSSIAN)
          launch_valu<double, MPLAK_KERNEL_GAUSSIAN>(d, p, hs, scale, R, Vg, ldv, part, skip_if);
        else
          launch_valu<double, MPLAK_KERNEL_MATERN32>(d, p, hs, scale, R, Vg, ldv, part, skip_if);
      } else {
        if (d->kind == MPLAK_KERNEL_GAUSSIAN)
          launch_valu<float, MPLAK_KERNEL_GAUSSIAN>(d, p, hs, scale, R, Vg, ldv, part, skip_if);
        else
          launch_valu<float, MPLAK_KERNEL_MATERN32>(d, p, hs, scale, R, Vg, ldv, part, skip_if);
      }
    }
    st = launch_status();
    if (st != MPLAK_OK) return st;
    int64_t fb = (d->row_count * R + kBlock - 1) / kBlock;
    if (fb > 8192) fb = 8192;
    hipLaunchKernelGGL(mrhs_finish_kernel, dim3((unsigned)fb), dim3(kBlock), 0, hs, d->row_count, R, p.s.slabs, part,
                       d->row_offset, d->col_offset, d->col_count, d->nugget, Vg, ldv, Yg, ldy, skip_if);
    st = launch_status();
    if (st != MPLAK_OK) return st;
  }
  return MPLAK_OK;
}

}  // extern "C"
