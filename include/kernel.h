/*
 * kernel.h -- drop-in for the reference's matrix-free kernel-matrix plugin
 * (src/kernel.h, src/kernel.cu).
 *
 * Same payload struct and core names.  Both cores evaluate the reference's
 * entries, exp(-r*r) of the point distance r plus 1e-8 where r < 1e-14
 * (src/kernel.cu:12-15, 39-50), in one fused gfx950 kernel: nothing of the
 * matrix is stored, so max_row_count_per_dgemv is accepted and ignored.
 */
#ifndef mpla_kernel_plugin_h__
#define mpla_kernel_plugin_h__

#include <stdio.h>

#include "mpla.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/kernel.h:3-8.  points: device array of `dim` device pointers (SoA). */
struct kernel_matrix_data
{
	double** points;
	int dim;
	int max_row_count_per_dgemv;
};

/* src/kernel.cu:70-82: b_i = prod_d (sqrt(pi)/2)(erf(1 - p_d) - erf(-p_d)) */
void mpla_set_gaussian_kernel_rhs(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_instance* instance);

/* src/kernel.cu:87-114 and :118-171 (A->data is a host struct kernel_matrix_data*) */
void mpla_dgemv_core_kernel_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);
void mpla_dgemv_core_kernel_streamed_cublas(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

#ifdef __cplusplus
}
#endif

#endif
