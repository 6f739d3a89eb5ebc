// Generic system adapter cores (src/generic_system_adapter.cu of the
// reference): a matrix-free product whose entries come from the virtual
// system_assembler::get_matrix_entry, evaluated on the fly in tiles.
//
// The reference materialises row panels through one virtual call per entry
// and then calls cublasDgemv; it also reads max_row_count_per_dgemv with a
// <<<1,1>>> kernel and a device-to-host copy on every product
// (src/generic_system_adapter.cu:83-137).  Here one thread owns one row, x is
// staged through LDS in chunks of 256, the columns are split into slabs so the
// launch fills the GPU, and slab partial sums are added in slab order
// (deterministic).  Compiled with -fgpu-rdc so the virtual calls reach
// assemblers defined in other translation units.

#include <hip/hip_runtime.h>

#include "../../../include/generic_system_adapter.h"
#include "../../../include/system_assembler.h"
#include "runtime.hpp"

using namespace mpla_rt;

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 256;
constexpr int kTargetBlocks = 2048;

struct Plan {
  int row_tiles, slabs, slab_cols;
};

Plan plan(int rows, int cols) {
  Plan p;
  p.row_tiles = (rows + kThreads - 1) / kThreads;
  int ks = (kTargetBlocks + p.row_tiles - 1) / p.row_tiles;
  const int chunks = (cols + kChunk - 1) / kChunk;
  ks = std::max(1, std::min(ks, std::min(chunks, 65535)));
  const int per = (chunks + ks - 1) / ks;
  p.slab_cols = per * kChunk;
  p.slabs = (cols + p.slab_cols - 1) / p.slab_cols;
  return p;
}

__global__ __launch_bounds__(kThreads) void generic_partial_kernel(system_assembler* assem, int row_off, int rows,
                                                                   int col_off, int cols, int slab_cols,
                                                                   const double* __restrict__ x,
                                                                   double* __restrict__ out,
                                                                   const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  __shared__ double lx[kChunk];
  const int i = blockIdx.x * kThreads + threadIdx.x;
  const bool active = i < rows;
  const int c0 = blockIdx.y * slab_cols;
  const int c1 = min(cols, c0 + slab_cols);
  double acc = 0.0;
  for (int cb = c0; cb < c1; cb += kChunk) {
    __syncthreads();
    const int j = cb + threadIdx.x;
    lx[threadIdx.x] = j < c1 ? x[j] : 0.0;
    __syncthreads();
    const int len = min(kChunk, c1 - cb);
    if (active)
      for (int jj = 0; jj < len; ++jj)
        acc = fma(assem->get_matrix_entry(row_off + i, col_off + cb + jj), lx[jj], acc);
  }
  if (active) out[(size_t)blockIdx.y * rows + i] = acc;
}

__global__ __launch_bounds__(kThreads) void slab_finish_kernel(int rows, int slabs, const double* __restrict__ part,
                                                               double* __restrict__ y,
                                                               const int32_t* __restrict__ skip) {
  if (skip != nullptr && *skip != 0) return;
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= rows) return;
  double s = part[i];
  for (int k = 1; k < slabs; ++k) s += part[(size_t)k * rows + i];
  y[i] = s;
}

__global__ __launch_bounds__(kThreads) void generic_rhs_kernel(system_assembler* assem, int row_off, int rows,
                                                               double* __restrict__ b) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i < rows) b[i] = assem->get_rhs_entry(row_off + i);
}

void generic_core(mpla_vector* b, mpla_generic_matrix* A, mpla_vector* x_redist, mpla_instance* inst,
                  const char* what) {
  const int rows = A->cur_proc_row_count, cols = A->cur_proc_col_count;
  if (rows <= 0) return;
  system_assembler* assem = static_cast<system_assembler*>(A->data);
  if (assem == nullptr) fail(what, "A->data holds no assembler");
  hipStream_t s = runtime(inst)->stream;
  const Plan p = plan(rows, std::max(cols, 1));
  double* part = static_cast<double*>(arena(inst, kCore, sizeof(double) * (size_t)p.slabs * (size_t)rows));
  hipLaunchKernelGGL(generic_partial_kernel, dim3(p.row_tiles, p.slabs), dim3(kThreads), 0, s, assem,
                     A->cur_proc_row_offset, rows, A->cur_proc_col_offset, cols, p.slab_cols, x_redist->data,
                     part, inst->skip_flag);
  hip_ck(hipGetLastError(), what);
  hipLaunchKernelGGL(slab_finish_kernel, dim3(p.row_tiles), dim3(kThreads), 0, s, rows, p.slabs, part, b->data,
                     inst->skip_flag);
  hip_ck(hipGetLastError(), what);
}

}  // namespace

extern "C" {

void mpla_set_generic_system_rhs(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_instance* instance) {
  const int rows = A->cur_proc_row_count;
  if (rows <= 0) return;
  hipLaunchKernelGGL(generic_rhs_kernel, dim3((rows + kThreads - 1) / kThreads), dim3(kThreads), 0,
                     runtime(instance)->stream, static_cast<system_assembler*>(A->data), A->cur_proc_row_offset, rows,
                     b->data);
  hip_ck(hipGetLastError(), "mpla_set_generic_system_rhs");
}

void mpla_dgemv_core_generic_system_matrix_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A,
                                                  struct mpla_vector* x_redist, struct mpla_instance* instance) {
  generic_core(b, A, x_redist, instance, "mpla_dgemv_core_generic_system_matrix_cublas");
}

void mpla_dgemv_core_generic_system_matrix_streamed_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A,
                                                           struct mpla_vector* x_redist,
                                                           struct mpla_instance* instance) {
  generic_core(b, A, x_redist, instance, "mpla_dgemv_core_generic_system_matrix_streamed_cublas");
}

}  // extern "C"
