// Chunked execution of solver iterations, recorded once into a HIP graph.
//
// A chunk is L consecutive iterations enqueued by a callback.  When the
// plugin core is capture-safe the chunk is captured on first use and then
// replayed with one hipGraphLaunch; otherwise it is enqueued eagerly.  The
// iteration callback receives the parity of the state buffer it reads; an odd
// L ends with `normalize`, which copies state[1] to state[0], so every chunk
// starts and ends on buffer 0.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <functional>
#include <vector>

#include "runtime.hpp"

namespace mpla_rt {

constexpr int kMaxChunk = 16;

// Largest chunk length <= kMaxChunk that divides the checkpoint interval, so
// checkpoint iterations fall on chunk ends; 1 when progress lines are printed.
inline int choose_chunk(int ckpt_every, int verbose) {
  if (verbose >= 2) return 1;
  if (ckpt_every <= 0) return kMaxChunk;
  for (int L = std::min(kMaxChunk, ckpt_every); L > 1; --L)
    if (ckpt_every % L == 0) return L;
  return 1;
}

class ChunkRunner {
 public:
  ChunkRunner(mpla_instance* inst, bool capture, int L, std::function<void(int)> iteration,
              std::function<void()> normalize)
      : inst_(inst), capture_(capture), L_(L), iteration_(std::move(iteration)),
        normalize_(std::move(normalize)) {}

  // With a slot: a graph captured by an earlier call with the same key is
  // replayed without capturing again, and a new capture is left in the slot
  // (which owns it) for later calls.  The key must hold every value the
  // iteration passes to its kernels.
  ChunkRunner(mpla_instance* inst, bool capture, int L, std::function<void(int)> iteration,
              std::function<void()> normalize, GraphSlot* slot, std::vector<unsigned char> key)
      : ChunkRunner(inst, capture, L, std::move(iteration), std::move(normalize)) {
    slot_ = slot;
    key_ = std::move(key);
    if (capture_ && slot_ != nullptr && slot_->exec != nullptr && slot_->key == key_) exec_ = slot_->exec;
  }

  ChunkRunner(const ChunkRunner&) = delete;
  ChunkRunner& operator=(const ChunkRunner&) = delete;

  ~ChunkRunner() {
    if (slot_ != nullptr) return;  // the slot owns the graph
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
  }

  void run() {
    hipStream_t s = runtime(inst_)->stream;
    if (capture_ && !exec_ && !capture_failed_) record(s);
    if (exec_) {
      hip_ck(hipGraphLaunch(exec_, s), "hipGraphLaunch");
      return;
    }
    enqueue();
  }

  bool captured() const { return exec_ != nullptr; }

 private:
  void enqueue() {
    for (int j = 0; j < L_; ++j) iteration_(j & 1);
    if (L_ & 1) normalize_();
  }

  void record(hipStream_t s) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      capture_failed_ = true;
      return;
    }
    enqueue();
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(s, &g) != hipSuccess || g == nullptr) {
      // nothing recorded has run; run() falls back to enqueuing eagerly
      (void)hipGetLastError();
      if (g) (void)hipGraphDestroy(g);
      capture_failed_ = true;
      if (verbosity(inst_) >= 2) fprintf(stderr, "MPLA: graph capture failed, running eagerly\n");
      return;
    }
    graph_ = g;
    hip_ck(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (slot_ != nullptr) {
      slot_->reset();
      slot_->graph = graph_;
      slot_->exec = exec_;
      slot_->key = key_;
    }
  }

  mpla_instance* inst_;
  bool capture_;
  int L_;
  std::function<void(int)> iteration_;
  std::function<void()> normalize_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool capture_failed_ = false;
  GraphSlot* slot_ = nullptr;
  std::vector<unsigned char> key_;
};

}  // namespace mpla_rt
