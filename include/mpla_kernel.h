/*
 * mpla_kernel.h -- kernel-interpolation systems beyond the reference plugin:
 * length scale, Matern-3/2, nugget, FP32 entry evaluation and the
 * mixed-precision solver.  No reference counterpart (the reference's
 * Gaussian has no length scale and Matern is unused, src/kernel.cu:7-15).
 */
#ifndef mpla_kernel_ext_h__
#define mpla_kernel_ext_h__

#include "mpla.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MPLA_KERNEL_GAUSSIAN 0 /* exp(-r^2 / (2 l^2)) */
#define MPLA_KERNEL_MATERN32 1 /* (1 + s) exp(-s), s = sqrt(3) r / l */

#define MPLA_PREC_F64 0
#define MPLA_PREC_F32 1
#define MPLA_PREC_BF16 2
#define MPLA_PREC_BF16X3 3

/* Payload of mpla_dgemv_core_kernel_fused (A->data, host struct).  The first
 * three fields match struct kernel_matrix_data. */
struct mpla_kernel_matrix_data
{
	double** points;             /* device array of dim device pointers (SoA) */
	int dim;
	int max_row_count_per_dgemv; /* ignored */
	int kernel;                  /* MPLA_KERNEL_* */
	int precision;               /* entry arithmetic of the core: MPLA_PREC_F64 or _F32 */
	double length_scale;
	double nugget;               /* added to K_ii (global row == global column) */
	double coincident_shift;     /* added where the distance is < 1e-14 (0: never) */
};

/* b = K x with entries on the fly (K + nugget I, in the payload's precision). */
void mpla_dgemv_core_kernel_fused(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x_redist, struct mpla_instance* instance);

/* Solve (K + nugget I) x = b to ||b - A x||/||b|| < epsilon, with the
 * residual, norms, dot products and updates in FP64 and the inner CG's
 * operator in `inner_precision` (MPLA_PREC_F32).  MPLA_PREC_F64 runs plain
 * FP64 CG (the recurrences and relative rule of
 * mpla_generic_conjugate_gradient_with_checkpoint_restart): on a square block
 * of up to 3 dimensions without a coincidence shift, on the points in
 * Hilbert-curve order with the FP64 product screened at 2^-(53 + log2 n),
 * whose dropped entries change no row by more than an ulp of its largest
 * term; otherwise with the fused core on every entry.
 * Iterative refinement: r = b - A64 x (FP64 operator), inner CG on A_low e = r
 * to a relative tolerance, x += e, until the true FP64 residual meets
 * epsilon.  iter_max bounds the total inner iterations.  On a square block
 * the solve runs on the points in Hilbert-curve order: the inner operator
 * drops entries below 2^-(bit below the inner tolerance), looser as the
 * inner residual falls (MPLA_IR_DROP_LOG2, MPLA_IR_RELAX), in single-reduction
 * CG; the FP64 residual products drop only entries below
 * 2^-(log2 n + log2(1/nugget) + log2(1/epsilon) + 12), whose sum moves the
 * residual by at most 2^-12 epsilon ||b|| (K is positive semi-definite; with
 * nugget <= 0 or a negative coincidence shift they drop only entries below an
 * FP64 ulp of a row's largest term).  On exit instance->last_rel_res is the
 * FP64 relative residual of x to that accuracy, instance->last_iterations the total inner
 * iterations and instance->last_matvecs the operator applications. */
void mpla_kernel_conjugate_gradient_mixed(struct mpla_vector* b, struct mpla_generic_matrix* A, struct mpla_vector* x, int iter_max, double epsilon, int inner_precision, struct mpla_instance* instance);

/* Solve (K + nugget I) X = B for nrhs right-hand sides: B and X are device
 * arrays, column-major N x nrhs with leading dimension A->cur_proc_row_count.
 * Batched CG (per-column scalars and stopping tests, one shared multi-RHS
 * product per iteration) inside FP64 iterative refinement; inner_precision is
 * MPLA_PREC_BF16, MPLA_PREC_BF16X3 or MPLA_PREC_F32 for the shared product
 * (raised automatically if refinement stagnates), or MPLA_PREC_F64 for plain
 * batched FP64 CG.  iter_max bounds the shared products.  rel_res (host array
 * of nrhs, may be NULL) receives every column's true FP64 relative residual;
 * instance->last_rel_res is the largest of them. */
void mpla_kernel_block_conjugate_gradient_mixed(double* B, double* X, int nrhs, struct mpla_generic_matrix* A, int iter_max, double epsilon, int inner_precision, double* rel_res, struct mpla_instance* instance);

#ifdef __cplusplus
}
#endif

#endif
