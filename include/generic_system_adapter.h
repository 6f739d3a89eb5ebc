/*
 * generic_system_adapter.h -- drop-in for the reference's generic system
 * adapter (src/generic_system_adapter.h, src/generic_system_adapter.cu).
 *
 * A->data is a device pointer to a system_assembler.  The cores evaluate
 * get_matrix_entry on the fly inside a tiled matrix-free product (nothing is
 * materialised, so the assembler's max_row_count_per_dgemv is not needed).
 */
#ifndef generic_system_adapter_h__
#define generic_system_adapter_h__

#include <stdio.h>

#include "mpla.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/generic_system_adapter.cu:39-50: b_i = get_rhs_entry(global i) */
void mpla_set_generic_system_rhs(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_instance* instance);

/* src/generic_system_adapter.cu:53-74 and :83-137 */
void mpla_dgemv_core_generic_system_matrix_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);
void mpla_dgemv_core_generic_system_matrix_streamed_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

#ifdef __cplusplus
}
#endif

#endif
