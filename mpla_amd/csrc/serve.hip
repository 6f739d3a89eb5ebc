// Ragged batched kernels for serving many independent kernel-interpolation
// solves at once (no reference counterpart).
//
// A batch is a device array of systems (one per request: its own points, N,
// kernel, right-hand sides and buffers) plus a device table of work items.
// One launch covers every system in the batch: workgroup b processes item b,
// whose fields name the system and the tile of work.  Each system's tiles,
// slabs and vector blocks depend only on that system's own (n, nrhs,
// precision), never on what else is in the batch, so a system's arithmetic,
// and therefore its result, is bitwise the same alone and in any mix.
//
// Products, two work splits (mplak_ragged_plan picks one per system from its
// n alone): SLABS, the VALU scheme of the multi-RHS product (one target per
// thread, RT columns per thread, sources through LDS, FP64 slab partials);
// PAIRS, the symmetric tile-pair form (each off-diagonal entry evaluated once
// for both of its rows).  The kernel function, dimension (<= 4) and scale
// come from the system record.
// CG: the batched recurrences of the multi-RHS solver with per-column
// scalars in a state pool, items (system, column, vector block).

#include <math.h>
#include <stdlib.h>

#include <type_traits>

#include "common.hpp"
#include "entry.hpp"

namespace mplak {
namespace {

constexpr int kChunk = 256;
constexpr int kRagTargetBlocks = 2048;
constexpr int kMaxRaggedDim = 4;

inline int vec_blocks_r(int64_t n) {
  int64_t b = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return (int)b;
}

inline int rt_class(int64_t R) { return R <= 1 ? 1 : (R <= 2 ? 2 : (R <= 4 ? 4 : 8)); }

// DIM > 0: every system of the launch has that dimension; DIM == 0: any
// dimension <= kMaxRaggedDim (coordinates past it are zero).
template <typename T, int RT, int DIM>
__global__ __launch_bounds__(kBlock) void ragged_product_kernel(const mplak_ragged_system* __restrict__ systems,
                                                                const mplak_ragged_item* __restrict__ items) {
  constexpr int MD = DIM > 0 ? DIM : kMaxRaggedDim;
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_system S = systems[it.sys];
  if (S.skip != nullptr && *S.skip != 0) return;
  __shared__ T lp[MD][kChunk];
  __shared__ T lv[RT][kChunk];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the chunk barrier below
  const int tid = threadIdx.x;
  const int64_t n = S.n;
  const int dim = S.dim;
  const int64_t R = S.nrhs;
  int64_t i = (int64_t)it.tile * kBlock + tid;
  const bool active = i < n;
  if (!active) i = n - 1;
  const int64_t r0 = (int64_t)it.ctile * RT;
  T pi[MD];
#pragma unroll
  for (int d = 0; d < MD; ++d) pi[d] = d < dim ? (T)(S.points[d * n + i] * S.scale) : T(0);
  const int64_t c0 = (int64_t)it.slab * S.slab_cols;
  const int64_t c1 = (c0 + S.slab_cols < n) ? c0 + S.slab_cols : n;
  // Screening (drop_s > 0): a wave skips a chunk when the bounding boxes of
  // its 64 targets and of the chunk's sources are further apart than drop_s,
  // i.e. every entry of the block is below the system's drop level, and
  // within a kept chunk each 64-source part whose box is that far.
  const bool screen = S.drop_s > 0.0;
  const int lane = tid & (kWave - 1), w = tid / kWave;
  __shared__ T cbox[kBlock / kWave][2 * MD];
  T tbox[2 * MD];
  if (screen) {
#pragma unroll
    for (int d = 0; d < MD; ++d) {
      tbox[d] = wave_min(pi[d]);
      tbox[MD + d] = wave_max(pi[d]);
    }
  }
  double acc[RT];
#pragma unroll
  for (int q = 0; q < RT; ++q) acc[q] = 0.0;
  for (int64_t cb = c0; cb < c1; cb += kChunk) {
    __syncthreads();
    const int64_t j = cb + tid;
#pragma unroll
    for (int d = 0; d < MD; ++d) lp[d][tid] = (j < c1 && d < dim) ? (T)(S.points[d * n + j] * S.scale) : T(0);
#pragma unroll
    for (int q = 0; q < RT; ++q) lv[q][tid] = (j < c1 && r0 + q < R) ? (T)S.V[(r0 + q) * n + j] : T(0);
    if (screen) {  // this wave's part of the chunk box; sources past the slab do not count
#pragma unroll
      for (int d = 0; d < MD; ++d) {
        const T lo = wave_min(j < c1 ? lp[d][tid] : (T)INFINITY);
        const T hi = wave_max(j < c1 ? lp[d][tid] : (T)-INFINITY);
        if (lane == 0) {
          cbox[w][d] = lo;
          cbox[w][MD + d] = hi;
        }
      }
    }
    __syncthreads();
    if (screen) {
      T d2 = T(0);
#pragma unroll
      for (int d = 0; d < MD; ++d) {
        T lo = cbox[0][d], hi = cbox[0][MD + d];
#pragma unroll
        for (int q = 1; q < kBlock / kWave; ++q) {
          lo = fmin(lo, cbox[q][d]);
          hi = fmax(hi, cbox[q][MD + d]);
        }
        const T gap = fmax(T(0), fmax(lo - tbox[MD + d], tbox[d] - hi));
        d2 = fma(gap, gap, d2);
      }
      if (__builtin_amdgcn_readfirstlane((int)(d2 > (T)S.drop_s)) != 0) continue;
    }
    // then per 64-source part of the chunk (wave q staged part q and its box):
    // the wave skips the parts whose box is too far from its targets
    uint32_t far_parts = 0;
    if (screen) {
#pragma unroll
      for (int q = 0; q < kBlock / kWave; ++q) {
        T d2 = T(0);
#pragma unroll
        for (int d = 0; d < MD; ++d) {
          const T gap = fmax(T(0), fmax(cbox[q][d] - tbox[MD + d], tbox[d] - cbox[q][MD + d]));
          d2 = fma(gap, gap, d2);
        }
        if (__builtin_amdgcn_readfirstlane((int)(d2 > (T)S.drop_s)) != 0) far_parts |= 1u << q;
      }
    }
    T cacc[RT];
#pragma unroll
    for (int q = 0; q < RT; ++q) cacc[q] = T(0);
#pragma unroll 1
    for (int part = 0; part < kBlock / kWave; ++part) {
      if ((far_parts >> part) & 1u) continue;
      for (int jj = part * kWave; jj < (part + 1) * kWave; ++jj) {
        T s = T(0);
#pragma unroll
        for (int d = 0; d < MD; ++d) {
          const T df = pi[d] - lp[d][jj];
          s = fma(df, df, s);
        }
        const T k = kernel_entry_rt<T>(S.kind, s, exp_tab);
#pragma unroll
        for (int q = 0; q < RT; ++q) cacc[q] = fma(k, lv[q][jj], cacc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < RT; ++q) acc[q] += (double)cacc[q];
  }
  if (active) {
#pragma unroll
    for (int q = 0; q < RT; ++q)
      if (r0 + q < R) S.part[((int64_t)it.slab * R + r0 + q) * n + i] = acc[q];
  }
}

// ---- pair scheme ------------------------------------------------------------
//
// The system's points in tiles of 512; item (pair p = (I, J), I <= J, column
// tile).  Each wave holds 128 targets of tile I (two per lane) and walks the
// eight 64-source groups of tile J, all 512 sources staged in LDS with their
// RT values.  Off the diagonal, lane l takes source (l + s) mod 64 at step s:
// one evaluation of each entry feeds the row sums of its two targets and a
// column sum per right-hand side that rotates one lane per step, so after the
// group lane l holds the column sums of source l over the wave's targets.
// The four waves' column sums of a group are added in wave order (FP64) and
// stored; the diagonal pair evaluates its whole block for the row sums.
//
// Partials (entry precision T): rows of I at part_i[(p R + r) 512 + k], rows
// of J at part_j[...] (same shape, after part_i), one flag per pair after
// both: 0 when every 128 x 64 block of the pair is screened out (nothing else
// written), 1 otherwise.  The finish adds, for a row of tile X, the pairs
// (X, P) over P = 0 .. tiles - 1 in order.

constexpr int kPairTile = 2 * kBlock;
constexpr int kPairGroups = kPairTile / kWave;
constexpr int64_t kRagPairMinN = 8192;

__host__ __device__ inline int64_t pair_count(int64_t tiles) { return tiles * (tiles + 1) / 2; }

template <typename T>
__device__ __forceinline__ T* pair_rows(const mplak_ragged_system& S) {
  return reinterpret_cast<T*>(S.part);
}
template <typename T>
__device__ __forceinline__ int32_t* pair_flags(const mplak_ragged_system& S) {
  const int64_t bytes = 2 * pair_count(S.tiles) * (int64_t)S.nrhs * kPairTile * (int64_t)sizeof(T);
  return reinterpret_cast<int32_t*>(reinterpret_cast<char*>(S.part) + (bytes + 255) / 256 * 256);
}

template <typename T, int RT, int DIM>
__global__ __launch_bounds__(kBlock) void ragged_pair_kernel(const mplak_ragged_system* __restrict__ systems,
                                                             const mplak_ragged_item* __restrict__ items) {
  typedef typename Quad<T>::type Q;
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_system S = systems[it.sys];
  // a launch specialised for DIM reads DIM coordinates: other systems are not its items
  if ((S.skip != nullptr && *S.skip != 0) || S.dim != DIM) return;
  __shared__ Q src[kPairTile];       // scaled source coordinates
  __shared__ T sval[kPairTile][RT];  // source values, column tile
  __shared__ T gbox[kPairGroups][2 * DIM];
  __shared__ T red[kBlock / kWave][RT][kWave];
  __shared__ uint32_t wkeep[kBlock / kWave];
  __shared__ double exp_tab[Exp2Lds<T>::size];
  if (sizeof(T) == 8) fill_exp2_table(exp_tab);  // first read after the staging barrier

  const int64_t n = S.n, R = S.nrhs;
  const int64_t p = it.tile;
  int64_t I, J;
  decode_pair(p, I, J);
  const int64_t r0 = (int64_t)it.ctile * RT;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  const double scale = S.scale;

  // every global load of the staging is issued before the first use
  double tc[2][DIM], tv[2][RT], sc[2][DIM], sv[2][RT];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int64_t i = I * kPairTile + w * 2 * kWave + h * kWave + lane;
    const int64_t ic = i < n ? i : n - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) tc[h][d] = S.points[d * n + ic];
#pragma unroll
    for (int q = 0; q < RT; ++q) tv[h][q] = (i < n && r0 + q < R) ? S.V[(r0 + q) * n + i] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t j = J * kPairTile + u * kBlock + tid;
    const int64_t jc = j < n ? j : n - 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) sc[u][d] = S.points[d * n + jc];
#pragma unroll
    for (int q = 0; q < RT; ++q) sv[u][q] = (j < n && r0 + q < R) ? S.V[(r0 + q) * n + j] : 0.0;
  }
  T pi[2][DIM], xi[2][RT];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) pi[h][d] = (T)(tc[h][d] * scale);
#pragma unroll
    for (int q = 0; q < RT; ++q) xi[h][q] = (T)tv[h][q];
  }
  const bool screen = S.drop_s > 0.0;
  T tbox[2 * DIM];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int k = u * kBlock + tid;  // group k / 64 = 4 u + w: one group per wave and pass
    T c[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int d = 0; d < DIM; ++d) c[d] = (T)(sc[u][d] * scale);
    Q q;
    q.x = c[0], q.y = c[1], q.z = c[2], q.w = c[3];
    src[k] = q;
#pragma unroll
    for (int r = 0; r < RT; ++r) sval[k][r] = (T)sv[u][r];
    if (screen) {  // padded sources repeat point n - 1 and do not widen the box
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const T lo = wave_min(c[d]), hi = wave_max(c[d]);
        if (lane == 0) {
          gbox[k / kWave][d] = lo;
          gbox[k / kWave][DIM + d] = hi;
        }
      }
    }
  }
  if (screen) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      tbox[d] = wave_min(fmin(pi[0][d], pi[1][d]));
      tbox[DIM + d] = wave_max(fmax(pi[0][d], pi[1][d]));
    }
  }
  __syncthreads();
  // groups this wave keeps (bit g); a pair nobody keeps writes only its flag
  uint32_t keep = (1u << kPairGroups) - 1u;
  if (screen) {
    keep = 0u;
#pragma unroll
    for (int g = 0; g < kPairGroups; ++g) {
      T d2 = T(0);
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const T gap = fmax(T(0), fmax(gbox[g][d] - tbox[DIM + d], tbox[d] - gbox[g][DIM + d]));
        d2 = fma(gap, gap, d2);
      }
      if (__builtin_amdgcn_readfirstlane((int)(d2 > (T)S.drop_s)) == 0) keep |= 1u << g;
    }
    if (lane == 0) wkeep[w] = keep;
    __syncthreads();
    const uint32_t any = wkeep[0] | wkeep[1] | wkeep[2] | wkeep[3];
    if (any == 0u) {
      if (tid == 0) pair_flags<T>(S)[p] = 0;
      return;
    }
  }
  if (tid == 0) pair_flags<T>(S)[p] = 1;

  T* part_i = pair_rows<T>(S);
  T* part_j = part_i + pair_count(S.tiles) * R * kPairTile;
  // the compute loop specialised per kernel function (the kind is uniform
  // over the workgroup: no per-entry branch)
  auto run = [&](auto kind_c) {
    constexpr int KIND = decltype(kind_c)::value;
    T acc[2][RT];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < RT; ++q) acc[h][q] = T(0);
    auto entries = [&](const Q& q, T (&e)[2]) {
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
        if constexpr (DIM > 3) {
          const T dw = pi[h][3] - q.w;
          s = fma(dw, dw, s);
        }
        e[h] = kernel_entry<T, KIND>(s, exp_tab);
      }
    };

    if (I == J) {
#pragma unroll 1
      for (int g = 0; g < kPairGroups; ++g) {
        if (((keep >> g) & 1u) == 0u) continue;
#pragma unroll 2
        for (int s = 0; s < kWave; ++s) {
          const int jj = g * kWave + s;
          T e[2];
          entries(src[jj], e);
#pragma unroll
          for (int q = 0; q < RT; ++q) {
            const T v = sval[jj][q];
            acc[0][q] = fma(e[0], v, acc[0][q]);
            acc[1][q] = fma(e[1], v, acc[1][q]);
          }
        }
      }
    } else {
#pragma unroll 1
      for (int g = 0; g < kPairGroups; ++g) {
        T t[RT];
#pragma unroll
        for (int q = 0; q < RT; ++q) t[q] = T(0);
        if ((keep >> g) & 1u) {
#pragma unroll 2
          for (int s = 0; s < kWave; ++s) {
            const int jj = g * kWave + ((lane + s) & (kWave - 1));
            T e[2];
            entries(src[jj], e);
#pragma unroll
            for (int q = 0; q < RT; ++q) {
              const T v = sval[jj][q];
              acc[0][q] = fma(e[0], v, acc[0][q]);
              acc[1][q] = fma(e[1], v, acc[1][q]);
              t[q] = fma(e[1], xi[1][q], fma(e[0], xi[0][q], t[q]));
              t[q] = wave_rotate(t[q]);
            }
          }
        }
#pragma unroll
        for (int q = 0; q < RT; ++q) red[w][q][lane] = t[q];
        __syncthreads();
        for (int v = tid; v < RT * kWave; v += kBlock) {
          const int q = v / kWave, l = v % kWave;
          double s = (double)red[0][q][l];
#pragma unroll
          for (int ww = 1; ww < kBlock / kWave; ++ww) s += (double)red[ww][q][l];
          if (r0 + q < R) part_j[(p * R + r0 + q) * kPairTile + g * kWave + l] = (T)s;
        }
        __syncthreads();
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < RT; ++q)
        if (r0 + q < R) part_i[(p * R + r0 + q) * kPairTile + w * 2 * kWave + h * kWave + lane] = acc[h][q];
  };
  if (S.kind == MPLAK_KERNEL_GAUSSIAN)
    run(std::integral_constant<int, MPLAK_KERNEL_GAUSSIAN>{});
  else
    run(std::integral_constant<int, MPLAK_KERNEL_MATERN32>{});
}

template <typename T>
__device__ __forceinline__ double pair_finish_row(const mplak_ragged_system& S, int64_t i, int64_t r) {
  const int64_t X = i / kPairTile, t = i % kPairTile, R = S.nrhs;
  const T* part_i = pair_rows<T>(S);
  const T* part_j = part_i + pair_count(S.tiles) * R * kPairTile;
  const int32_t* flags = pair_flags<T>(S);
  // the partial of pair (X, P), or 0 for a pair whose flag says it wrote
  // nothing (its block is then never used, whatever the memory holds)
  auto term = [&](int64_t P) -> double {
    const int64_t p = X <= P ? sym_pair_index(X, P) : sym_pair_index(P, X);
    const T v = (X <= P ? part_i : part_j)[(p * R + r) * kPairTile + t];
    return flags[p] != 0 ? (double)v : 0.0;
  };
  // loads batched ahead of the additions, which stay in tile order; the
  // last batch adds +0.0 past the last tile, which leaves a sum that started
  // at +0.0 unchanged, so the bits are those of adding tile by tile
  constexpr int kBatch = 8;
  double s = 0.0;
  for (int64_t P = 0; P < S.tiles; P += kBatch) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int64_t Q = P + u < S.tiles ? P + u : S.tiles - 1;
      const double t = term(Q);
      v[u] = P + u < S.tiles ? t : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) s += v[u];
  }
  return s;
}

// items: (system, block) over n * nrhs entries of Y, 256 per block
__global__ __launch_bounds__(kBlock) void ragged_finish_kernel(const mplak_ragged_system* __restrict__ systems,
                                                               const mplak_ragged_item* __restrict__ items) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_system S = systems[it.sys];
  if (S.skip != nullptr && *S.skip != 0) return;
  const int64_t g = (int64_t)it.tile * kBlock + threadIdx.x;
  const int64_t n = S.n, R = S.nrhs;
  if (g >= n * R) return;
  const int64_t i = g % n, r = g / n;
  double s;
  if (S.scheme == MPLAK_RAGGED_PAIRS) {
    s = S.precision == MPLAK_PREC_F64 ? pair_finish_row<double>(S, i, r) : pair_finish_row<float>(S, i, r);
  } else {
    s = S.part[r * n + i];
    for (int64_t k = 1; k < S.slabs; ++k) s += S.part[(k * R + r) * n + i];
  }
  if (S.nugget != 0.0) s = fma(S.nugget, S.V[r * n + i], s);
  S.Y[r * n + i] = s;
}

// ---- ragged CG --------------------------------------------------------------

__device__ __forceinline__ mplak_cg_state* cg_in(const mplak_ragged_cg& S, int parity) {
  return S.states + (int64_t)parity * S.nrhs;
}

// items: (system, column, vector block)
__global__ __launch_bounds__(kBlock) void ragged_start_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                              const mplak_ragged_item* __restrict__ items) {
  __shared__ double scratch[kBlock / kWave];
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int64_t n = S.n, off = (int64_t)it.slab * n;
  const int G = S.G;
  double rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)G * kBlock) {
    const double bi = S.B[off + i];
    const double ri = bi - S.Z[off + i];
    S.R[off + i] = ri;
    S.D[off + i] = ri;
    rr = fma(ri, ri, rr);
    bb = fma(bi, bi, bb);
  }
  rr = block_sum<kBlock>(rr, scratch);
  bb = block_sum<kBlock>(bb, scratch);
  if (threadIdx.x == 0) {
    S.part[(2 * it.slab) * G + it.tile] = rr;
    S.part[(2 * it.slab + 1) * G + it.tile] = bb;
  }
}

// items: (system, column) in fields (sys, slab); writes states[column] of buffer 0
__global__ __launch_bounds__(kBlock) void ragged_start_final_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                                    const mplak_ragged_item* __restrict__ items) {
  __shared__ double scratch[kBlock / kWave];
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int G = S.G, c = it.slab;
  double rr, bb;
  reduce_partials2<kBlock>(S.part + (2 * c) * G, S.part + (2 * c + 1) * G, G, scratch, rr, bb);
  if (threadIdx.x == 0) {
    mplak_cg_state s;
    s.rr = rr;
    s.bnorm2 = bb;
    s.eps = S.eps;
    s.n_global = (double)S.n;
    // a zero right-hand side with a zero residual is solved exactly (x = 0);
    // 0/0 would otherwise leave a NaN that never meets the tolerance
    s.rel = (bb == 0.0 && rr == 0.0) ? 0.0 : sqrt(rr) / sqrt(bb);
    s.alpha = s.beta = 0.0;
    s.k = 0;
    s.iter_max = 2;
    s.rule = MPLAK_STOP_RELATIVE;
    s.done = s.rel < S.eps ? MPLAK_CONVERGED : MPLAK_RUNNING;
    S.states[c] = s;
  }
}

// one workgroup per system (items: (system, -)); `parity` selects the buffer to read
__global__ __launch_bounds__(kBlock) void ragged_group_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                              const mplak_ragged_item* __restrict__ items, int parity) {
  __shared__ int scratch[kBlock / kWave];
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const mplak_cg_state* st = cg_in(S, parity);
  int running = 0;
  for (int c = threadIdx.x; c < S.nrhs; c += kBlock) running += st[c].done == MPLAK_RUNNING ? 1 : 0;
  running = block_sum<kBlock>(running, scratch);
  if (threadIdx.x == 0) {
    S.group->active = running;
    S.group->all_done = running == 0 ? 1 : 0;
  }
}

__global__ __launch_bounds__(kBlock) void ragged_dz_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                           const mplak_ragged_item* __restrict__ items, int parity) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int c = it.slab;
  if (cg_in(S, parity)[c].done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int64_t n = S.n, off = (int64_t)c * n;
  const int G = S.G;
  double s = 0.0;
  for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)G * kBlock)
    s = fma(S.D[off + i], S.Z[off + i], s);
  s = block_sum<kBlock>(s, scratch);
  if (threadIdx.x == 0) S.part[c * G + it.tile] = s;
}

__global__ __launch_bounds__(kBlock) void ragged_update_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                               const mplak_ragged_item* __restrict__ items,
                                                               int parity) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int c = it.slab;
  const mplak_cg_state* in = cg_in(S, parity);
  if (in[c].done != MPLAK_RUNNING) return;
  __shared__ double scratch[kBlock / kWave];
  const int G = S.G;
  const double alpha = in[c].rr / reduce_partials<kBlock>(S.part + c * G, G, scratch);
  const int64_t n = S.n, off = (int64_t)c * n;
  double rr = 0.0;
  for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)G * kBlock) {
    S.X[off + i] = fma(alpha, S.D[off + i], S.X[off + i]);
    const double ri = fma(-alpha, S.Z[off + i], S.R[off + i]);
    S.R[off + i] = ri;
    rr = fma(ri, ri, rr);
  }
  rr = block_sum<kBlock>(rr, scratch);
  if (threadIdx.x == 0) S.part[(S.nrhs + c) * G + it.tile] = rr;
}

__global__ __launch_bounds__(kBlock) void ragged_direct_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                               const mplak_ragged_item* __restrict__ items,
                                                               int parity) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int c = it.slab;
  const mplak_cg_state s_in = cg_in(S, parity)[c];
  mplak_cg_state* out = cg_in(S, 1 - parity);
  if (s_in.done != MPLAK_RUNNING) {
    if (it.tile == 0 && threadIdx.x == 0) out[c] = s_in;
    return;
  }
  __shared__ double scratch[kBlock / kWave];
  const int G = S.G;
  double dz, rr_new;
  reduce_partials2<kBlock>(S.part + c * G, S.part + (S.nrhs + c) * G, G, scratch, dz, rr_new);
  const double beta = rr_new / s_in.rr;
  const int k = s_in.k + 1;
  const double rel = sqrt(rr_new) / sqrt(s_in.bnorm2);
  const bool conv = rel < s_in.eps;
  const bool limit = k >= s_in.iter_max - 1;
  if (!conv) {
    const int64_t n = S.n, off = (int64_t)c * n;
    for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)G * kBlock)
      S.D[off + i] = fma(beta, S.D[off + i], S.R[off + i]);
  }
  if (it.tile == 0 && threadIdx.x == 0) {
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

// items: (system, column, vector block): X += E (E stored in Z's companion buffer `E`)
__global__ __launch_bounds__(kBlock) void ragged_axpy_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                             const mplak_ragged_item* __restrict__ items) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int64_t n = S.n, off = (int64_t)it.slab * n;
  for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)S.G * kBlock)
    S.Xout[off + i] += S.X[off + i];
}

__global__ __launch_bounds__(kBlock) void ragged_fill_kernel(const mplak_ragged_cg* __restrict__ systems,
                                                             const mplak_ragged_item* __restrict__ items,
                                                             double value) {
  const mplak_ragged_item it = items[blockIdx.x];
  const mplak_ragged_cg S = systems[it.sys];
  const int64_t n = S.n, off = (int64_t)it.slab * n;
  for (int64_t i = (int64_t)it.tile * kBlock + threadIdx.x; i < n; i += (int64_t)S.G * kBlock) S.X[off + i] = value;
}

template <typename T, int DIM>
void launch_product_dim(int rt, const mplak_ragged_system* sys, const mplak_ragged_item* items, int64_t count,
                        hipStream_t st) {
  const dim3 g((unsigned)count), b(kBlock);
  switch (rt) {
    case 1: hipLaunchKernelGGL((ragged_product_kernel<T, 1, DIM>), g, b, 0, st, sys, items); break;
    case 2: hipLaunchKernelGGL((ragged_product_kernel<T, 2, DIM>), g, b, 0, st, sys, items); break;
    case 4: hipLaunchKernelGGL((ragged_product_kernel<T, 4, DIM>), g, b, 0, st, sys, items); break;
    default: hipLaunchKernelGGL((ragged_product_kernel<T, 8, DIM>), g, b, 0, st, sys, items); break;
  }
}

template <typename T>
void launch_product(int rt, int dim, const mplak_ragged_system* sys, const mplak_ragged_item* items, int64_t count,
                    hipStream_t st) {
  switch (dim) {
    case 1: launch_product_dim<T, 1>(rt, sys, items, count, st); break;
    case 2: launch_product_dim<T, 2>(rt, sys, items, count, st); break;
    case 3: launch_product_dim<T, 3>(rt, sys, items, count, st); break;
    case 4: launch_product_dim<T, 4>(rt, sys, items, count, st); break;
    default: launch_product_dim<T, 0>(rt, sys, items, count, st); break;
  }
}

template <typename T, int DIM>
void launch_pairs_dim(int rt, const mplak_ragged_system* sys, const mplak_ragged_item* items, int64_t count,
                      hipStream_t st) {
  const dim3 g((unsigned)count), b(kBlock);
  switch (rt) {
    case 1: hipLaunchKernelGGL((ragged_pair_kernel<T, 1, DIM>), g, b, 0, st, sys, items); break;
    case 2: hipLaunchKernelGGL((ragged_pair_kernel<T, 2, DIM>), g, b, 0, st, sys, items); break;
    case 4: hipLaunchKernelGGL((ragged_pair_kernel<T, 4, DIM>), g, b, 0, st, sys, items); break;
    default: hipLaunchKernelGGL((ragged_pair_kernel<T, 8, DIM>), g, b, 0, st, sys, items); break;
  }
}

template <typename T>
void launch_pairs(int rt, int dim, const mplak_ragged_system* sys, const mplak_ragged_item* items, int64_t count,
                  hipStream_t st) {
  switch (dim) {
    case 1: launch_pairs_dim<T, 1>(rt, sys, items, count, st); break;
    case 2: launch_pairs_dim<T, 2>(rt, sys, items, count, st); break;
    case 3: launch_pairs_dim<T, 3>(rt, sys, items, count, st); break;
    default: launch_pairs_dim<T, 4>(rt, sys, items, count, st); break;
  }
}

// PAIRS for n >= kRagPairMinN unless MPLAK_RAGGED_PAIRS says otherwise
// (0: never, 1: every n); the pair index must fit an item's 32-bit field.
bool use_pairs(int64_t n) {
  const char* v = getenv("MPLAK_RAGGED_PAIRS");
  bool on = n >= kRagPairMinN;
  if (v != nullptr && *v != '\0') on = atoi(v) != 0;
  const int64_t tiles = (n + kPairTile - 1) / kPairTile;
  return on && pair_count(tiles) <= 0x7fffffff;
}


}  // namespace
}  // namespace mplak

using namespace mplak;

extern "C" {

size_t mplak_ragged_plan(mplak_ragged_system* s) {
  if (s == nullptr || s->n <= 0 || s->nrhs <= 0 || s->dim < 1 || s->dim > kMaxRaggedDim) return 0;
  s->rt = rt_class(s->nrhs);
  s->ctiles = (s->nrhs + s->rt - 1) / s->rt;
  s->tiles = (int32_t)((s->n + kBlock - 1) / kBlock);
  int64_t ks = (kRagTargetBlocks + (int64_t)s->tiles * s->ctiles - 1) / ((int64_t)s->tiles * s->ctiles);
  const int64_t chunks = (s->n + kChunk - 1) / kChunk;
  if (ks > chunks) ks = chunks;
  if (ks < 1) ks = 1;
  const int64_t per = (chunks + ks - 1) / ks;
  s->slab_cols = per * kChunk;
  s->slabs = (int32_t)((s->n + s->slab_cols - 1) / s->slab_cols);
  s->scale = coord_scale(s->kind, s->length_scale);
  s->drop_s = s->drop_log2 > 0.0 ? drop_distance2(s->kind, s->drop_log2) : 0.0;
  s->reserved = 0;
  if (use_pairs(s->n)) {
    s->scheme = MPLAK_RAGGED_PAIRS;
    s->tiles = (int32_t)((s->n + kPairTile - 1) / kPairTile);
    s->slabs = 0;
    s->slab_cols = 0;
    const size_t es = s->precision == MPLAK_PREC_F64 ? sizeof(double) : sizeof(float);
    const size_t pairs = (size_t)pair_count(s->tiles);
    return align_up(2 * pairs * (size_t)s->nrhs * kPairTile * es, 256) + align_up(pairs * sizeof(int32_t), 256);
  }
  s->scheme = MPLAK_RAGGED_SLABS;
  return (size_t)s->slabs * (size_t)s->nrhs * (size_t)s->n * sizeof(double);
}

int64_t mplak_ragged_product_items(const mplak_ragged_system* s, mplak_ragged_item* out, int32_t sys_index) {
  if (s == nullptr || s->tiles <= 0) return 0;
  if (s->scheme == MPLAK_RAGGED_PAIRS) {
    const int64_t pairs = pair_count(s->tiles);
    if (out) {
      int64_t k = 0;
      for (int32_t c = 0; c < s->ctiles; ++c)
        for (int64_t q = 0; q < pairs; ++q) out[k++] = mplak_ragged_item{sys_index, (int32_t)q, 0, c};
    }
    return pairs * s->ctiles;
  }
  const int64_t count = (int64_t)s->tiles * s->slabs * s->ctiles;
  if (out) {
    int64_t k = 0;
    for (int32_t c = 0; c < s->ctiles; ++c)
      for (int32_t sl = 0; sl < s->slabs; ++sl)
        for (int32_t t = 0; t < s->tiles; ++t) out[k++] = mplak_ragged_item{sys_index, t, sl, c};
  }
  return count;
}

int64_t mplak_ragged_finish_items(const mplak_ragged_system* s, mplak_ragged_item* out, int32_t sys_index) {
  if (s == nullptr || s->n <= 0) return 0;
  const int64_t count = (s->n * s->nrhs + kBlock - 1) / kBlock;
  if (out)
    for (int64_t b = 0; b < count; ++b) out[b] = mplak_ragged_item{sys_index, (int32_t)b, 0, 0};
  return count;
}

int mplak_ragged_product(const mplak_ragged_system* systems, const mplak_ragged_item* items, int64_t count, int rt,
                         int dim, int precision, int scheme, mplak_stream_t s) {
  if (count <= 0) return MPLAK_OK;
  if (!systems || !items || count > 0x7fffffff || dim < 0 || dim > kMaxRaggedDim) return MPLAK_ERR_ARG;
  if (scheme != MPLAK_RAGGED_SLABS && scheme != MPLAK_RAGGED_PAIRS) return MPLAK_ERR_ARG;
  if (scheme == MPLAK_RAGGED_PAIRS && dim == 0) return MPLAK_ERR_ARG;
  hipStream_t hs = as_stream(s);
  if (scheme == MPLAK_RAGGED_PAIRS) {
    if (precision == MPLAK_PREC_F64)
      launch_pairs<double>(rt, dim, systems, items, count, hs);
    else
      launch_pairs<float>(rt, dim, systems, items, count, hs);
  } else if (precision == MPLAK_PREC_F64) {
    launch_product<double>(rt, dim, systems, items, count, hs);
  } else {
    launch_product<float>(rt, dim, systems, items, count, hs);
  }
  return launch_status();
}

int mplak_ragged_finish(const mplak_ragged_system* systems, const mplak_ragged_item* items, int64_t count,
                        mplak_stream_t s) {
  if (count <= 0) return MPLAK_OK;
  if (!systems || !items) return MPLAK_ERR_ARG;
  hipLaunchKernelGGL(ragged_finish_kernel, dim3((unsigned)count), dim3(kBlock), 0, as_stream(s), systems, items);
  return launch_status();
}

int32_t mplak_ragged_cg_blocks(int64_t n) { return vec_blocks_r(n); }

size_t mplak_ragged_cg_part_bytes(int64_t n, int32_t nrhs) {
  return (size_t)2 * vec_blocks_r(n) * (size_t)nrhs * sizeof(double);
}

int64_t mplak_ragged_cg_items(const mplak_ragged_cg* s, mplak_ragged_item* out, int32_t sys_index) {
  if (s == nullptr || s->n <= 0) return 0;
  const int64_t count = (int64_t)s->G * s->nrhs;
  if (out) {
    int64_t k = 0;
    for (int32_t c = 0; c < s->nrhs; ++c)
      for (int32_t b = 0; b < s->G; ++b) out[k++] = mplak_ragged_item{sys_index, b, c, 0};
  }
  return count;
}

int mplak_ragged_cg_op(int op, const mplak_ragged_cg* systems, const mplak_ragged_item* items, int64_t count,
                       int parity, double value, mplak_stream_t s) {
  if (count <= 0) return MPLAK_OK;
  if (!systems || !items) return MPLAK_ERR_ARG;
  hipStream_t hs = as_stream(s);
  const dim3 g((unsigned)count), b(kBlock);
  switch (op) {
    case MPLAK_RAGGED_START: hipLaunchKernelGGL(ragged_start_kernel, g, b, 0, hs, systems, items); break;
    case MPLAK_RAGGED_START_FINAL: hipLaunchKernelGGL(ragged_start_final_kernel, g, b, 0, hs, systems, items); break;
    case MPLAK_RAGGED_GROUP: hipLaunchKernelGGL(ragged_group_kernel, g, b, 0, hs, systems, items, parity); break;
    case MPLAK_RAGGED_DZ: hipLaunchKernelGGL(ragged_dz_kernel, g, b, 0, hs, systems, items, parity); break;
    case MPLAK_RAGGED_UPDATE: hipLaunchKernelGGL(ragged_update_kernel, g, b, 0, hs, systems, items, parity); break;
    case MPLAK_RAGGED_DIRECT: hipLaunchKernelGGL(ragged_direct_kernel, g, b, 0, hs, systems, items, parity); break;
    case MPLAK_RAGGED_AXPY: hipLaunchKernelGGL(ragged_axpy_kernel, g, b, 0, hs, systems, items); break;
    case MPLAK_RAGGED_FILL: hipLaunchKernelGGL(ragged_fill_kernel, g, b, 0, hs, systems, items, value); break;
    default: return MPLAK_ERR_ARG;
  }
  return launch_status();
}

}  // extern "C"
