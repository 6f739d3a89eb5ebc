/*
 * quadrature.h -- kernel-quadrature helpers of the drop-in library (libmpla.so).
 *
 * The reference's quadrature driver (src/generate_quadrature_weights.cpp)
 * draws its nodes from GSL's Halton generator (gsl_qrng_alloc(
 * gsl_qrng_halton, dim) + gsl_qrng_get, :107-117), a dependency this library
 * replaces with its own generator of the same sequence.
 */
#ifndef MPLA_QUADRATURE_H
#define MPLA_QUADRATURE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Halton points into HOST memory: out[d * ld + i] is coordinate d of point
 * first + i + 1 of the Halton sequence, i.e. the radical inverse of
 * k = first + i + 1 in the d-th prime base (2, 3, 5, ...), with the digit
 * weights formed by repeated multiplication with 1/base, as GSL's generator
 * does.  first = 0 gives the points gsl_qrng_get returns from a fresh
 * generator, in order (src/generate_quadrature_weights.cpp:107-117).
 * 1 <= dim <= 64, n >= 0, ld >= n.  Returns 0, or -1 on invalid arguments. */
int mpla_halton_points(double* out, int32_t dim, int64_t n, int64_t first, int64_t ld);

#ifdef __cplusplus
}
#endif

#endif /* MPLA_QUADRATURE_H */
