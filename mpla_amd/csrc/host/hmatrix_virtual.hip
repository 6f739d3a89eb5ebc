// H-matrix kernels whose entries come from a device-side virtual
// system_assembler (src/system_assembler.h:6-12): entry (i, j) of the sorted
// operator is assem->get_matrix_entry(id of row i, id of column j).  Built as
// relocatable device code with the assembler's translation units.

#include "../../../include/system_assembler.h"
#include "../hmatrix_kernels.hpp"
#include "hmatrix.hpp"

namespace {

struct VirtualEntries {
  system_assembler* a;
  const int64_t* rid;  // sorted row -> point id
  const int64_t* cid;  // sorted column -> point id
  static constexpr bool kTable = false;
  __device__ double operator()(int64_t i, int64_t j, const double*) const {
    return a->get_matrix_entry((int)rid[i], (int)cid[j]);
  }
};

__global__ void rhs_kernel(system_assembler* a, const int64_t* __restrict__ ids, int64_t n, double* __restrict__ b) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    b[i] = a->get_rhs_entry((int)ids[i]);
}

int status() { return hipGetLastError() == hipSuccess ? MPLAK_OK : MPLAK_ERR_HIP; }

}  // namespace

namespace mpla_hm {

using mplak::hm::kHmThreads;

int virtual_aca(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                const mplak_hm_block* blocks, int64_t nblocks, int kmax, double eps, double* U, double* V,
                int32_t* rank, void* stream) {
  if (nblocks == 0) return MPLAK_OK;
  VirtualEntries e{assem, row_ids, col_ids};
  hipLaunchKernelGGL(mplak::hm::aca_kernel<VirtualEntries>, dim3((unsigned)nblocks), dim3(kHmThreads), 0,
                     static_cast<hipStream_t>(stream), e, blocks, kmax, eps, U, V, rank);
  return status();
}

int virtual_apply(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                  const mplak_hm_chunk* chunks, int64_t nchunks, const int64_t* near_ptr, const int64_t* near_cols,
                  const int64_t* far_ptr, const int32_t* far_blk, const mplak_hm_block* blocks, const int32_t* rank,
                  int kmax, const double* U, const double* t, const double* xs, double* ys, void* stream) {
  if (nchunks == 0) return MPLAK_OK;
  VirtualEntries e{assem, row_ids, col_ids};
  hipLaunchKernelGGL(mplak::hm::apply_generic_kernel<VirtualEntries>, dim3((unsigned)nchunks), dim3(mplak::hm::kHmApplyThreads), 0,
                     static_cast<hipStream_t>(stream), e, chunks, near_ptr, near_cols, far_ptr, far_blk, blocks, rank,
                     kmax, U, t, xs, ys);
  return status();
}

int virtual_near_fill(system_assembler* assem, const int64_t* row_ids, const int64_t* col_ids,
                      const int64_t* range_leaf_lo, const int64_t* near_cols, const int64_t* near_off,
                      int64_t nranges, double* stored, void* stream) {
  if (nranges == 0) return MPLAK_OK;
  VirtualEntries e{assem, row_ids, col_ids};
  hipLaunchKernelGGL(mplak::hm::near_fill_kernel<VirtualEntries>, dim3((unsigned)nranges), dim3(kHmThreads), 0,
                     static_cast<hipStream_t>(stream), e, range_leaf_lo, near_cols, near_off, stored);
  return status();
}

int virtual_rhs(system_assembler* assem, const int64_t* ids, int64_t n, double* b, void* stream) {
  if (n == 0) return MPLAK_OK;
  int64_t g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(rhs_kernel, dim3((unsigned)g), dim3(256), 0, static_cast<hipStream_t>(stream), assem, ids, n, b);
  return status();
}

}  // namespace mpla_hm
