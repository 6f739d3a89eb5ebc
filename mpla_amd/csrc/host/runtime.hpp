// Internal runtime of the drop-in host library (libmpla.so).
//
// One Runtime per mpla_instance: the HIP stream every call of the instance is
// ordered on, device workspace arenas that grow on demand and are reused
// across calls (the reference allocates and frees temporaries in every call,
// src/mpla.cpp:723,740,776,793), and a pinned host buffer for the few scalars
// that must reach the host.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#include "../../../include/mpla.h"
#include "../../../include/mplak.h"

namespace mpla_rt {

// Separate arenas so that nested users never share memory: a solver holds
// its vectors in kSolver while the plugin core it calls uses kCore.
enum Arena { kCore = 0, kBlas = 1, kSolver = 2, kMixed = 3, kArenaCount = 4 };

// A captured chunk of solver iterations kept between calls: replayed again
// when every argument its kernels were captured with (the key) is the same.
struct GraphSlot {
  std::vector<unsigned char> key;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    key.clear();
  }
};

struct Runtime {
  hipStream_t stream = nullptr;
  int device = 0;
  void* arena[kArenaCount] = {};
  size_t arena_bytes[kArenaCount] = {};
  void* pinned = nullptr;  // kPinnedBytes of page-locked host memory
  GraphSlot mixed_chunk;   // the mixed solver's inner-iteration chunk
  // iterations each refinement step of the last mixed solve of this size
  // took to converge: a later solve launches that many before its first look
  // at the state (launches past convergence return at once)
  int64_t mixed_hint_n = -1;
  std::vector<int> mixed_hint_iters;
};

constexpr size_t kPinnedBytes = 65536;

Runtime* runtime(mpla_instance* inst);

// Device memory of at least `bytes` from the given arena (previous contents
// are not preserved when it grows).  Growing synchronises the stream and is
// refused while the stream is being captured into a graph.
void* arena(mpla_instance* inst, Arena which, size_t bytes);

// Reference error convention (src/mpla.cpp:1187-1193): print and exit.
[[noreturn]] void fail(const char* what, const char* detail);

inline void ck(int status, const char* what) {
  if (status != MPLAK_OK) fail(what, mplak_status_string(status));
}

inline void hip_ck(hipError_t e, const char* what) {
  if (e != hipSuccess) fail(what, hipGetErrorString(e));
}

inline mplak_stream_t stream_of(mpla_instance* inst) { return runtime(inst)->stream; }

// Drop level at which a screened FP64 kernel product of n points is the FP64
// product: the at most n entries below 2^-(53 + log2 n) of a row change it by
// at most 2^-53 max|x|, a unit in the last place of the largest term
// (entries are at most 1).  Between 2^-60 and 2^-80.
inline double residual_drop_log2(int64_t n) {
  const double d = 53.0 + std::ceil(std::log2((double)std::max<int64_t>(n, 1)));
  return std::min(80.0, std::max(60.0, d));
}

// Drop level of the FP64 residual products for a solve of (K + nugget I) x =
// b to relative residual epsilon (per right-hand side).  Every dropped entry
// is below 2^-drop, so a row loses at most 2^-drop ||x||_1 <= 2^-drop sqrt(n)
// ||x||_2, and the residual's error is at most n 2^-drop ||x||_2 <=
// n 2^-drop ||b||_2 / nugget (K is positive semi-definite, so ||x|| <=
// ||b|| / nugget for the solution, which the refined iterates approach).
// With drop = log2 n + log2(1/nugget) + log2(1/epsilon) + 12 that error is at
// most 2^-12 epsilon ||b||_2: the stopping test and the reported residual are
// the true ones to 2^-12 of the tolerance.  Never tighter than
// residual_drop_log2(n); that level is also taken for nugget <= 0 or
// epsilon <= 0.
inline double residual_screen_log2(int64_t n, double nugget, double epsilon) {
  const double exact = residual_drop_log2(n);
  if (!(nugget > 0.0) || !(epsilon > 0.0)) return exact;
  const double d = std::ceil(std::log2((double)std::max<int64_t>(n, 1)) - std::log2(nugget) -
                             std::log2(epsilon) + 12.0);
  return std::min(exact, std::max(24.0, d));
}

// Copy `bytes` from device to the pinned buffer and wait; returns the buffer.
void* read_back(mpla_instance* inst, const void* dev, size_t bytes);

bool is_graph_safe(mpla_dgemv_core_t core);

// b = A x through a plugin core on the 1 x 1 grid: x_redist is x itself.
void generic_apply(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x, mpla_dgemv_core_t core,
                   mpla_instance* inst);

// A vector view on caller-provided device memory with the distribution of
// `like` (no tables are copied; the view must not be freed).
mpla_vector vector_view(const mpla_vector* like, double* data);

int verbosity(mpla_instance* inst);

}  // namespace mpla_rt
