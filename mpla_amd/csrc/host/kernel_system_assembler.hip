// Factory of the Gaussian kernel system assembler
// (src/kernel_system_assembler.cu:10-36 of the reference).
//
// The object must be constructed in device code so that its vtable pointer
// refers to the device vtable.  The reference uses device-side `new`; here
// the storage comes from hipMalloc and a one-thread kernel constructs the
// object in place, so the device heap is never involved.  Compiled with
// -fgpu-rdc together with the generic system adapter.

#include <hip/hip_runtime.h>

#include <new>

#include "../../../include/kernel_system_assembler.h"
#include "runtime.hpp"

using namespace mpla_rt;

namespace {

__global__ void construct_kernel(gaussian_kernel_system_assembler** assem, void* storage, double** points,
                                 int max_row_count_per_dgemv, int dim) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  gaussian_kernel_system_assembler* a = new (storage) gaussian_kernel_system_assembler();
  a->points = points;
  a->max_row_count_per_dgemv = max_row_count_per_dgemv;
  a->dim = dim;
  a->regularization = 0.0;
  *assem = a;
}

__global__ void destruct_kernel(gaussian_kernel_system_assembler* a) {
  if (threadIdx.x == 0 && blockIdx.x == 0) a->gaussian_kernel_system_assembler::~gaussian_kernel_system_assembler();
}

}  // namespace

extern "C" {

void create_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem, double** points,
                                                    int max_row_count_per_dgemv, int dim) {
  void* storage = nullptr;
  hip_ck(hipMalloc(&storage, sizeof(gaussian_kernel_system_assembler)), "create_gaussian_kernel_system_assembler_object");
  hipLaunchKernelGGL(construct_kernel, dim3(1), dim3(1), 0, 0, assem, storage, points, max_row_count_per_dgemv, dim);
  hip_ck(hipGetLastError(), "create_gaussian_kernel_system_assembler_object");
  hip_ck(hipDeviceSynchronize(), "create_gaussian_kernel_system_assembler_object");
}

void destroy_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem) {
  gaussian_kernel_system_assembler* a = nullptr;
  hip_ck(hipMemcpy(&a, assem, sizeof(a), hipMemcpyDeviceToHost), "destroy_gaussian_kernel_system_assembler_object");
  if (a == nullptr) return;
  hipLaunchKernelGGL(destruct_kernel, dim3(1), dim3(1), 0, 0, a);
  hip_ck(hipGetLastError(), "destroy_gaussian_kernel_system_assembler_object");
  hip_ck(hipDeviceSynchronize(), "destroy_gaussian_kernel_system_assembler_object");
  hip_ck(hipFree(a), "destroy_gaussian_kernel_system_assembler_object");
}

}  // extern "C"
