// Halton quadrature nodes (replaces the GSL generator the reference's
// quadrature driver uses, src/generate_quadrature_weights.cpp:107-117).

#include "../../../include/quadrature.h"

#include <vector>

// Unfused multiply-add, so that the digit sums round as in a plain C build of
// the generator (and identically on every host compiler).
#pragma STDC FP_CONTRACT OFF

namespace {

// The first `count` primes, by trial division against the primes found so far.
std::vector<int64_t> first_primes(int count) {
  std::vector<int64_t> primes;
  for (int64_t c = 2; (int)primes.size() < count; ++c) {
    bool prime = true;
    for (int64_t p : primes) {
      if (p * p > c) break;
      if (c % p == 0) {
        prime = false;
        break;
      }
    }
    if (prime) primes.push_back(c);
  }
  return primes;
}

// Radical inverse of k in base b; the weight of digit m is (1/b)^(m+1) built
// by repeated multiplication, the evaluation order of GSL's generator.
double radical_inverse(int64_t k, int64_t b, double inv_b) {
  double r = 0.0, weight = 1.0;
  for (; k > 0; k /= b) {
    weight *= inv_b;
    r += weight * (double)(k % b);
  }
  return r;
}

}  // namespace

extern "C" int mpla_halton_points(double* out, int32_t dim, int64_t n, int64_t first, int64_t ld) {
  if (dim < 1 || dim > 64 || n < 0 || first < 0 || ld < n || (n > 0 && out == nullptr)) return -1;
  const std::vector<int64_t> bases = first_primes(dim);
  for (int d = 0; d < dim; ++d) {
    const double inv_b = 1.0 / (double)bases[d];
    double* row = out + (int64_t)d * ld;
    for (int64_t i = 0; i < n; ++i) row[i] = radical_inverse(first + i + 1, bases[d], inv_b);
  }
  return 0;
}
