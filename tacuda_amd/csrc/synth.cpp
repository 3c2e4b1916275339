// Seeded synthetic OHLCV generator (host).  Documented PRNG so any machine
// regenerates the exact same float32 series: splitmix64 seeds a
// xoshiro256** stream per (seed, series); normals by Box-Muller in double.
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/tacuda_amd.h"
#include "host.hpp"

namespace {

inline uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed, uint64_t stream) {
    uint64_t k = seed * 0x100000001b3ull ^ (stream + 0x632be59bd9b4e019ull);
    for (auto& v : s) v = splitmix64(k);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in (0, 1]
  double u01() { return ((next() >> 11) + 1) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = u01(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

void gen_series(uint64_t seed, int series, int n, float* o, float* h, float* l, float* c,
                float* v) {
  Xoshiro rng(seed, (uint64_t)series);
  double close = 100.0;
  float prevc = 100.0f;
  for (int t = 0; t < n; ++t) {
    const double zr = rng.normal(), zh = rng.normal(), zl = rng.normal(), zv = rng.normal();
    if (t > 0) close = close * std::exp(0.01 * zr);
    const float cf = (float)close;
    const float of = (t == 0) ? 100.0f : prevc;
    const double hi = std::fmax((double)of, (double)cf) * (1.0 + std::fabs(0.005 * zh));
    const double lo = std::fmin((double)of, (double)cf) * (1.0 - std::fabs(0.005 * zl));
    if (o) o[t] = of;
    if (h) h[t] = (float)hi;
    if (l) l[t] = (float)lo;
    if (c) c[t] = cf;
    if (v) v[t] = (float)std::exp(10.0 + zv);
    prevc = cf;
  }
}

}  // namespace

extern "C" CTAPI_EXPORT ctStatus_t ctx_synth_ohlcv(unsigned long long seed, int first, int rows,
                                                   int n, float* o, float* h, float* l, float* c,
                                                   float* v, int threads) {
  return tca::guarded([&]() -> ctStatus_t {
    if (rows <= 0 || n <= 0 || first < 0)
      return tca::fail(CT_STATUS_INVALID_ARGUMENT, "ctx_synth_ohlcv: bad shape");
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > rows) nt = rows;
    auto work = [&](int t0) {
      for (int r = t0; r < rows; r += nt) {
        const size_t off = (size_t)r * n;
        gen_series(seed, first + r, n, o ? o + off : nullptr, h ? h + off : nullptr,
                   l ? l + off : nullptr, c ? c + off : nullptr, v ? v + off : nullptr);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    return CT_STATUS_SUCCESS;
  });
}

extern "C" CTAPI_EXPORT ctStatus_t ctx_workspace_reserve(hipStream_t stream, size_t bytes) {
  return tca::guarded([&]() -> ctStatus_t {
    tca::Ws ws;
    for (int part = 0; part < tca::kWsParts; ++part) {
      ctStatus_t st = tca::workspace_for(stream, bytes, &ws, part);
      if (st != CT_STATUS_SUCCESS) return st;
    }
    return CT_STATUS_SUCCESS;
  });
}
