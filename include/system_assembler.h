/*
 * system_assembler.h -- drop-in for the reference's abstract device-side
 * system assembler (src/system_assembler.h:6-12).  HIP C++ only.
 *
 * A concrete assembler is constructed in device memory (so its vtable is a
 * device vtable) and handed to the generic system adapter cores through
 * mpla_generic_matrix::data.  Code that defines its own assembler must be
 * compiled with -fgpu-rdc and linked with generic_system_adapter.hip so that
 * the adapter's kernels can reach the virtual functions.
 */
#ifndef system_assembler_h__
#define system_assembler_h__

#include <hip/hip_runtime.h>

#include "mpla.h"

class system_assembler
{
	public:
		int max_row_count_per_dgemv;
		virtual	__device__ double get_matrix_entry(int i, int j) =0;
		virtual __device__ double get_rhs_entry(int i) =0;
};

#endif
