// CPU baseline of the 7-indicator pass with the reference's own host build of
// TA-Lib 0.8.3 (src/compat/Talib08Formulas.h), linked into
// oracle/_ref/libtaref.so beside ref_driver.cpp.
//
// TEST AND BENCHMARK INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times
// it; no product code links it.  Inputs are widened to double and outputs
// allocated once per thread before the clock starts, so the timed loop is
// the TA_* calls alone: SMA(20), EMA(20), RSI(14), MACD(12, 26, 9),
// BBANDS(20, 2, 2), ATR(14), STOCH(14, 3, 3) over the whole series.
#include "Talib08Formulas.h"

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

using namespace tacuda::talib08::ref;

namespace {

struct Worker {
  std::vector<double> h, l, c, o[12];
  std::vector<unsigned char> arena;
  long long passes = 0;
  int rc = 0;
};

int one_pass(Worker& w, int n) {
  Arena a{w.arena.data(), w.arena.size()};
  int b = 0, nb = 0, rc = 0;
  rc |= TA_SMA(a, 0, n - 1, w.c.data(), 20, &b, &nb, w.o[0].data());
  rc |= TA_EMA(a, 0, n - 1, w.c.data(), 20, &b, &nb, w.o[1].data());
  rc |= TA_RSI(a, 0, n - 1, w.c.data(), 14, &b, &nb, w.o[2].data());
  rc |= TA_MACD(a, 0, n - 1, w.c.data(), 12, 26, 9, &b, &nb, w.o[3].data(), w.o[4].data(),
                w.o[5].data());
  rc |= TA_BBANDS(a, 0, n - 1, w.c.data(), 20, 2.0, 2.0, 0, &b, &nb, w.o[6].data(),
                  w.o[7].data(), w.o[8].data());
  rc |= TA_ATR(a, 0, n - 1, w.h.data(), w.l.data(), w.c.data(), 14, &b, &nb, w.o[9].data());
  rc |= TA_STOCH(a, 0, n - 1, w.h.data(), w.l.data(), w.c.data(), 14, 3, 0, 3, 0, &b, &nb,
                 w.o[10].data(), w.o[11].data());
  return rc;
}

}  // namespace

#define API extern "C" __attribute__((visibility("default")))

API int ref_hardware_threads() { return (int)std::thread::hardware_concurrency(); }

// Runs the pass over the series on `threads` threads (each its own copy: one
// series per thread) until `seconds` have elapsed after the first pass of
// every thread; returns bars per second over all threads and the passes run.
API int ref_bench7(const float* hi, const float* lo, const float* cl, int n, int threads,
                   double seconds, double* barsPerSec, long long* passes) {
  if (!hi || !lo || !cl || n <= 0 || threads <= 0 || !barsPerSec || !passes) return 4;
  std::vector<Worker> ws(threads);
  for (Worker& w : ws) {
    w.h.assign(hi, hi + n);
    w.l.assign(lo, lo + n);
    w.c.assign(cl, cl + n);
    for (auto& v : w.o) v.assign(n, 0.0);
    w.arena.resize((size_t)n * 64 + (1u << 20));
    w.rc = one_pass(w, n);  // warm-up: touches every buffer once
    if (w.rc) return w.rc;
  }
  std::atomic<int> ready{0};
  std::atomic<bool> go{false};
  using clk = std::chrono::steady_clock;
  clk::time_point t0;
  const auto budget = std::chrono::duration<double>(seconds);
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t) {
    th.emplace_back([&, t]() {
      Worker& w = ws[t];
      ready.fetch_add(1);
      while (!go.load()) std::this_thread::yield();
      const clk::time_point start = t0;
      do {
        w.rc |= one_pass(w, n);
        ++w.passes;
      } while (clk::now() - start < budget);
    });
  }
  while (ready.load() < threads) std::this_thread::yield();
  t0 = clk::now();
  go.store(true);
  for (auto& x : th) x.join();
  const double el = std::chrono::duration<double>(clk::now() - t0).count();
  long long total = 0;
  for (Worker& w : ws) {
    if (w.rc) return w.rc;
    total += w.passes;
  }
  *passes = total;
  *barsPerSec = (double)total * n / el;
  return 0;
}
