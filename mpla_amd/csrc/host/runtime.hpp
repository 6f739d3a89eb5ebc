// Internal runtime of the drop-in host library (libmpla.so).
//
// One Runtime per mpla_instance: the HIP stream every call of the instance is
// ordered on, device workspace arenas that grow on demand and are reused
// across calls (the reference allocates and frees temporaries in every call,
// src/mpla.cpp:723,740,776,793), and a pinned host buffer for the few scalars
// that must reach the host.
#pragma once

#include <hip/hip_runtime.h>

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
