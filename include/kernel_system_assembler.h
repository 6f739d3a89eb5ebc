/*
 * kernel_system_assembler.h -- drop-in for the reference's Gaussian kernel
 * system assembler (src/kernel_system_assembler.h:10-50,
 * src/kernel_system_assembler.cu:10-36).  HIP C++ only.
 *
 * Entries: exp(-r*r) of the point distance r, plus `regularization` on the
 * diagonal i == j.  The factory value-initialises the object, so
 * regularization is 0 unless the caller sets it on the device.
 */
#ifndef kernel_system_assembler_h__
#define kernel_system_assembler_h__

#include <hip/hip_runtime.h>

#include "mpla.h"
#include "system_assembler.h"

class gaussian_kernel_system_assembler : public system_assembler
{
	public:
		double** points;       /* device array of dim device pointers (SoA) */
		int dim;
		double regularization;

		__device__ double gaussian_kernel(double r)
		{
			return exp(-r * r);
		}

		__device__ double get_matrix_entry(int i, int j)
		{
			double s = 0.0;
			for (int d = 0; d < dim; d++)
			{
				const double df = points[d][i] - points[d][j];
				s += df * df;
			}
			const double k = gaussian_kernel(sqrt(s));
			return (i == j) ? k + regularization : k;
		}

		__device__ double get_rhs_entry(int i)
		{
			/* integral of exp(-|p_i - y|^2) over y in [0,1]^dim */
			double v = 1.0;
			for (int d = 0; d < dim; d++)
				v = v * 0.88622692545275801365 * (erf(1.0 - points[d][i]) - erf(0.0 - points[d][i]));
			return v;
		}
};

#ifdef __cplusplus
extern "C" {
#endif

/* *assem (a device pointer-to-pointer, as in the reference) receives a device
 * pointer to a new gaussian_kernel_system_assembler. */
void create_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem, double** points, int max_row_count_per_dgemv, int dim);

void destroy_gaussian_kernel_system_assembler_object(struct gaussian_kernel_system_assembler** assem);

#ifdef __cplusplus
}
#endif

#endif
