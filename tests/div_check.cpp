// Host check of tacuda_amd/csrc/divide.hpp: div_by equals IEEE division bit
// for bit for every divisor the kernels use (integer periods 1..100000, the
// WMA weight totals p(p+1)/2, 3 and 0.015-scaled values) over random and
// boundary numerators, and passes infinities and NaNs through.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <random>

#include "../tacuda_amd/csrc/divide.hpp"

static bool same(double a, double b) {
  if (std::isnan(a) && std::isnan(b)) return true;
  uint64_t x, y;
  std::memcpy(&x, &a, 8);
  std::memcpy(&y, &b, 8);
  return x == y;
}

int main() {
  std::mt19937_64 g(12345);
  std::uniform_real_distribution<double> mant(1.0, 2.0);
  std::uniform_int_distribution<int> ex(-60, 60);
  long long bad = 0, n = 0;
  auto check = [&](double x, double d) {
    const tca::DivBy k = tca::div_by_of(d);
    const double a = tca::div_by(x, k), b = x / d;
    ++n;
    if (!same(a, b)) {
      if (bad < 5) std::printf("mismatch x=%.17g d=%.17g got %.17g want %.17g\n", x, d, a, b);
      ++bad;
    }
  };
  for (int p = 1; p <= 100000; ++p) {
    const double ds[2] = {(double)p, (double)p * (p + 1) / 2.0};
    for (double d : ds) {
      for (int i = 0; i < 200; ++i) {
        const double x = std::ldexp(mant(g), ex(g)) * ((i & 1) ? -1.0 : 1.0);
        check(x, d);
      }
      // sums of float32 prices: multiples of small powers of two
      for (int i = 0; i < 50; ++i) {
        const float f = (float)(std::ldexp(mant(g), ex(g) % 20));
        check((double)f * (double)(1 + (i % 37)), d);
      }
      check(0.0, d);
      check(-0.0, d);
      check(std::numeric_limits<double>::infinity(), d);
      check(-std::numeric_limits<double>::infinity(), d);
      check(std::numeric_limits<double>::quiet_NaN(), d);
    }
  }
  for (int i = 0; i < 2000000; ++i) check(std::ldexp(mant(g), ex(g)), 3.0);
  if (bad) {
    std::printf("div_by: %lld of %lld differ\n", bad, n);
    return 1;
  }
  std::printf("div_by ok (%lld cases)\n", n);
  return 0;
}
