// Division of doubles by a divisor fixed for a loop (a window period, its
// weight total, a smoothing length): with rd = RN(1 / d) precomputed,
// Markstein's correction of x * rd -- q = RN(x * rd), r = x - d q (exact
// with an fma), RN(q + r rd) -- is the correctly rounded quotient x / d for
// every finite nonzero x whose quotient is a normal double and every d whose
// reciprocal is normal, so the results are the IEEE division's bit for bit
// at a few instructions instead of a full division.  Zeros, infinities and
// NaNs pass through as in x / d (q itself).  Shared by the kernels and their
// host check (tests/div_check.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace tca {

struct DivBy {
  double d, rd;
};

__host__ __device__ __forceinline__ DivBy div_by_of(double d) { return DivBy{d, 1.0 / d}; }

__host__ __device__ __forceinline__ double div_by(double x, const DivBy& k) {
  const double q = x * k.rd;
  const double c = fma(fma(-q, k.d, x), k.rd, q);
  return (__builtin_isfinite(q) && q != 0.0) ? c : q;
}

}  // namespace tca
