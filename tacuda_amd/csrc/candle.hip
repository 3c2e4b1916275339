// The 61 TA-Lib candlestick patterns on the GPU.  One workgroup per 256-bar
// tile of a row: the tile and a 15-bar halo (5 candles back plus 10-bar
// setting windows) are staged in LDS once, the four window totals every
// pattern averages over (real body and high-low over 10 bars, shadows over
// 10, high-low over 5) are built once per bar in LDS, and each work-item
// then evaluates any subset of the 61 rules for its bar, writing one int32
// (or legacy float) signal plane per pattern -- coalesced across the tile.
// Reference surface: ct_cdl_*_d (tacuda_device.h:438+), ct_cdl_*_batch_d
// (tacuda.h:949+); signals {0, +-80, +-100, +-200}, 0 during warm-up.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../include/tacuda_amd.h"
#include "host.hpp"
#include "hostio.hpp"

namespace tca {
namespace cdl {

enum { BodyLong, BodyVeryLong, BodyShort, BodyDoji, ShadowLong, ShadowVeryLong, ShadowShort,
       ShadowVeryShort, Near, Far, Equal };

constexpr int kTile = 256;
constexpr int kHalo = 15;  // bars read before a tile: 5 candles back + 10-bar totals
constexpr int kSumHalo = 5;  // averages are taken up to 5 candles back (HIKKAKEMOD)

// LDS tile accessor: bars from `base`, window totals from `sbase`.
// rb10, hl10, sh10, hl5 hold the window means (total / avgPeriod), divided
// once per bar when the tile is built rather than at every use.
struct CdlCtx {
  const float *so, *sh, *sl, *sc;
  const double *rb10, *hl10, *sh10, *hl5;
  int base, sbase;
  __device__ double o(int j) const { return so[j - base]; }
  __device__ double h(int j) const { return sh[j - base]; }
  __device__ double l(int j) const { return sl[j - base]; }
  __device__ double c(int j) const { return sc[j - base]; }
  // TA_CANDLEAVERAGE: factor * (window total / avgPeriod) / (2 for shadows)
  __device__ double avg(int s, int j) const {
    const int k = j - sbase;
    switch (s) {
      case BodyLong: return 1.0 * rb10[k] / 1.0;
      case BodyVeryLong: return 3.0 * rb10[k] / 1.0;
      case BodyShort: return 1.0 * rb10[k] / 1.0;
      case BodyDoji: return 0.1 * hl10[k] / 1.0;
      case ShadowLong: return 1.0 * fabs(c(j) - o(j)) / 1.0;
      case ShadowVeryLong: return 2.0 * fabs(c(j) - o(j)) / 1.0;
      case ShadowShort: return 1.0 * sh10[k] / 2.0;
      case ShadowVeryShort: return 0.1 * hl10[k] / 1.0;
      case Near: return 0.2 * hl5[k] / 1.0;
      case Far: return 0.6 * hl5[k] / 1.0;
      default: return 0.05 * hl5[k] / 1.0;  // Equal
    }
  }
};

}  // namespace cdl
}  // namespace tca

#include "candle_rules.hpp"

namespace tca {
namespace cdl {

// Pattern order, names (the reference's ct_cdl_<name>_d) and look-backs
// (reference src/indicators/<Pattern>.cu warm-ups).
#define TCA_CDL_LIST(X)                                                                        \
  X(two_crows, 12, twocrows) X(three_black_crows, 13, threeblackcrows)                          \
  X(three_inside, 12, threeinside) X(three_line_strike, 8, threelinestrike)                     \
  X(three_outside, 3, threeoutside) X(three_stars_in_south, 12, threestarsinsouth)              \
  X(three_white_soldiers, 12, threewhitesoldiers) X(abandoned_baby, 12, abandonedbaby)          \
  X(advance_block, 12, advanceblock) X(belt_hold, 10, belthold) X(breakaway, 14, breakaway)     \
  X(closing_marubozu, 10, closingmarubozu) X(conceal_baby_swallow, 13, concealbabyswall)        \
  X(counterattack, 11, counterattack) X(dark_cloud_cover, 11, darkcloudcover)                   \
  X(doji, 10, doji) X(doji_star, 11, dojistar) X(dragonfly_doji, 10, dragonflydoji)             \
  X(engulfing, 2, engulfing) X(evening_doji_star, 12, eveningdojistar)                          \
  X(evening_star, 12, eveningstar) X(gap_side_side_white, 7, gapsidesidewhite)                  \
  X(gravestone_doji, 10, gravestonedoji) X(hammer, 11, hammer) X(hanging_man, 11, hangingman)   \
  X(harami, 11, harami) X(harami_cross, 11, haramicross) X(high_wave, 10, highwave)             \
  X(hikkake, 5, hikkake) X(hikkake_mod, 10, hikkakemod) X(homing_pigeon, 11, homingpigeon)      \
  X(identical_three_crows, 12, identical3crows) X(in_neck, 11, inneck)                          \
  X(inverted_hammer, 11, invertedhammer) X(kicking, 11, kicking)                                \
  X(kicking_by_length, 11, kickingbylength) X(ladder_bottom, 14, ladderbottom)                  \
  X(long_legged_doji, 10, longleggeddoji) X(long_line, 10, longline) X(marubozu, 10, marubozu)  \
  X(matching_low, 6, matchinglow) X(mat_hold, 14, mathold)                                      \
  X(morning_doji_star, 12, morningdojistar) X(morning_star, 12, morningstar)                    \
  X(on_neck, 11, onneck) X(piercing, 11, piercing) X(rickshaw_man, 10, rickshawman)             \
  X(rise_fall3_methods, 14, risefall3methods) X(separating_lines, 11, separatinglines)          \
  X(shooting_star, 11, shootingstar) X(short_line, 10, shortline)                               \
  X(spinning_top, 10, spinningtop) X(stalled_pattern, 12, stalledpattern)                       \
  X(stick_sandwich, 7, sticksandwich) X(takuri, 10, takuri) X(tasuki_gap, 7, tasukigap)         \
  X(thrusting, 11, thrusting) X(tristar, 12, tristar) X(unique_3_river, 12, unique3river)       \
  X(upside_gap_2_crows, 12, upsidegap2crows) X(xside_gap_3_methods, 2, xsidegap3methods)

enum Id {
#define TCA_ID(name, lb, fn) k_##name,
  TCA_CDL_LIST(TCA_ID)
#undef TCA_ID
  kCount
};
static_assert(kCount == 61, "61 TA-Lib patterns");

__constant__ int c_lookback[kCount] = {
#define TCA_LB(name, lb, fn) lb,
    TCA_CDL_LIST(TCA_LB)
#undef TCA_LB
};
const char* const kNames[kCount] = {
#define TCA_NM(name, lb, fn) #name,
    TCA_CDL_LIST(TCA_NM)
#undef TCA_NM
};

__device__ __forceinline__ int eval(int k, const CdlCtx& X, int i) {
  switch (k) {
#define TCA_EV(name, lb, fn) \
  case k_##name:             \
    return fn(X, i);
    TCA_CDL_LIST(TCA_EV)
#undef TCA_EV
    default:
      return 0;
  }
}

// out: plane k of pattern k at out + k*planeStride (only set bits of mask
// are written; a single pattern uses planeStride 0).
// ALL: every pattern, as straight-line code (shared subexpressions of the
// rules are computed once); otherwise the patterns of `mask`, one at a time.
template <bool I32, bool ALL>
__global__ __launch_bounds__(kTile) void k_cdl(const float* o, const float* h, const float* l,
                                               const float* c, void* out, size_t planeStride,
                                               unsigned long long mask, int n, int rows,
                                               int tilesPerRow) {
  __shared__ float so[kTile + kHalo], sh[kTile + kHalo], sl[kTile + kHalo], sc[kTile + kHalo];
  __shared__ double rb10[kTile + kSumHalo], hl10[kTile + kSumHalo], sh10[kTile + kSumHalo],
      hl5[kTile + kSumHalo];
  const int tile = blockIdx.x;
  const int row = tile / tilesPerRow;
  const int j0 = (tile - row * tilesPerRow) * kTile;
  if (row >= rows) return;
  const size_t ro = (size_t)row * n;
  const int base = j0 - kHalo, sbase = j0 - kSumHalo;
  const float nan = __builtin_nanf("");
  for (int e = threadIdx.x; e < kTile + kHalo; e += kTile) {
    const int j = base + e;
    const bool in = j >= 0 && j < n;
    so[e] = in ? o[ro + j] : nan;
    sh[e] = in ? h[ro + j] : nan;
    sl[e] = in ? l[ro + j] : nan;
    sc[e] = in ? c[ro + j] : nan;
  }
  __syncthreads();
  // window totals over the candles strictly before bar j, oldest first
  for (int e = threadIdx.x; e < kTile + kSumHalo; e += kTile) {
    const int j = sbase + e;
    double r = 0.0, hl = 0.0, s = 0.0, h5 = 0.0;
    for (int u = j - 10; u < j; ++u) {
      const double ou = so[u - base], hu = sh[u - base], lu = sl[u - base], cu = sc[u - base];
      r += fabs(cu - ou);
      hl += hu - lu;
      s += (hu - fmax(ou, cu)) + (fmin(ou, cu) - lu);
      if (u >= j - 5) h5 += hu - lu;
    }
    rb10[e] = r / 10;
    hl10[e] = hl / 10;
    sh10[e] = s / 10;
    hl5[e] = h5 / 5;
  }
  __syncthreads();
  const int i = j0 + threadIdx.x;
  if (i >= n) return;
  const CdlCtx X{so, sh, sl, sc, rb10, hl10, sh10, hl5, base, sbase};
  auto put = [&](int k, int v) {
    const size_t idx = (size_t)k * planeStride + ro + i;
    if (I32)
      __builtin_nontemporal_store(v, static_cast<int32_t*>(out) + idx);
    else
      __builtin_nontemporal_store((float)v, static_cast<float*>(out) + idx);
  };
  if (ALL) {
#define TCA_ALL(name, lb, fn) put(k_##name, i < lb ? 0 : fn(X, i));
    TCA_CDL_LIST(TCA_ALL)
#undef TCA_ALL
    return;
  }
  unsigned long long m = mask;
  while (m) {
    const int k = __builtin_ctzll(m);
    m &= m - 1;
    put(k, i < c_lookback[k] ? 0 : eval(k, X, i));
  }
}

ctStatus_t launch(const float* o, const float* h, const float* l, const float* c, void* out,
                  size_t planeStride, unsigned long long mask, int n, int rows, bool i32,
                  hipStream_t s, const char* fn) {
  if (!o || !h || !l || !c || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (n <= 0 || rows <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen and numSeries must be positive", fn);
  const int tiles = (n + kTile - 1) / kTile;
  const long long blocks = (long long)tiles * rows;
  if (blocks > 0x7fffffffLL) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: matrix too large", fn);
  const bool all = mask == (~0ull >> (64 - kCount)) && planeStride != 0;
  KTime kt(s, i32 && all ? "k_cdl<I32,ALL>" : i32 ? "k_cdl<I32>" : "k_cdl");
  if (i32 && all)
    hipLaunchKernelGGL((k_cdl<true, true>), dim3((unsigned)blocks), dim3(kTile), 0, s, o, h, l, c,
                       out, planeStride, mask, n, rows, tiles);
  else if (i32)
    hipLaunchKernelGGL((k_cdl<true, false>), dim3((unsigned)blocks), dim3(kTile), 0, s, o, h, l,
                       c, out, planeStride, mask, n, rows, tiles);
  else
    hipLaunchKernelGGL((k_cdl<false, false>), dim3((unsigned)blocks), dim3(kTile), 0, s, o, h, l,
                       c, out, planeStride, mask, n, rows, tiles);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// Legacy engulfing signal (reference CandleUtils.h:233-260 as used by
// BullishEngulfing.cu:10-23 and BearishEngulfing.cu): TA-Lib's engulfing
// rule on bars t-1, t, written from bar 1 of every row (0 at bar 0).
__global__ void k_engulf_legacy(const float* __restrict__ o, const float* __restrict__ c,
                                float* __restrict__ out, int n, long long total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  if (i % n == 0) {
    out[i] = 0.0f;
    return;
  }
  const float co = o[i], cc = c[i], po = o[i - 1], pc = c[i - 1];
  const int cur = cc >= co ? 1 : -1, prev = pc >= po ? 1 : -1;
  const bool bull = cur == 1 && prev == -1 && ((cc >= po && co < pc) || (cc > po && co <= pc));
  const bool bear = cur == -1 && prev == 1 && ((co >= pc && cc < po) || (co > pc && cc <= po));
  out[i] = (bull || bear) ? (float)(cur * ((co != pc && cc != po) ? 100 : 80)) : 0.0f;
}

ctStatus_t engulf_legacy(const float* o, const float* h, const float* l, const float* c,
                         float* out, int n, int rows, hipStream_t s, const char* fn) {
  if (!o || !h || !l || !c || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (n <= 0 || rows <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen and numSeries must be positive", fn);
  const long long total = (long long)n * rows;
  const long long blocks = (total + 255) / 256;
  if (blocks > 0x7fffffffLL) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: matrix too large", fn);
  hipLaunchKernelGGL(k_engulf_legacy, dim3((unsigned)blocks), dim3(256), 0, s, o, c, out, n, total);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// Host form of a float-signal call: mask selects one pattern of the list, 0
// the legacy engulfing alias.
ctStatus_t host_run(const float* ho, const float* hh, const float* hl, const float* hc, float* hy,
                    int n, int rows, unsigned long long mask, hipStream_t s, const char* fn) {
  if (!ho || !hh || !hl || !hc || !hy) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (n <= 0 || rows <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen and numSeries must be positive", fn);
  const size_t e = (size_t)n * rows;
  HostRun r(s);
  const float *o, *h, *l, *c;
  float* y;
  ctStatus_t st = r.in(ho, e, &o);
  if (st == CT_STATUS_SUCCESS) st = r.in(hh, e, &h);
  if (st == CT_STATUS_SUCCESS) st = r.in(hl, e, &l);
  if (st == CT_STATUS_SUCCESS) st = r.in(hc, e, &c);
  if (st == CT_STATUS_SUCCESS) st = r.out(hy, e, &y);
  if (st == CT_STATUS_SUCCESS)
    st = mask ? launch(o, h, l, c, y, 0, mask, n, rows, false, s, fn)
              : engulf_legacy(o, h, l, c, y, n, rows, s, fn);
  if (st == CT_STATUS_SUCCESS) st = r.finish();
  return st;
}

}  // namespace cdl
}  // namespace tca

using namespace tca;

extern "C" {

CTAPI_EXPORT int ctx_cdl_count(void) { return cdl::kCount; }

CTAPI_EXPORT const char* ctx_cdl_name(int pattern) {
  return (pattern >= 0 && pattern < cdl::kCount) ? cdl::kNames[pattern] : nullptr;
}

CTAPI_EXPORT ctStatus_t ctx_cdl_batch_d(int pattern, const float* o, const float* h,
                                        const float* l, const float* c, int32_t* out,
                                        int seriesLen, int numSeries, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (pattern < 0 || pattern >= cdl::kCount)
      return fail(CT_STATUS_INVALID_ARGUMENT, "cdl: pattern index out of range");
    ctStatus_t r = cdl::launch(o, h, l, c, out, 0, 1ull << pattern, seriesLen, numSeries, true, s,
                               cdl::kNames[pattern]);
    if (r != CT_STATUS_SUCCESS) return r;
    return finish_batch(s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_cdl_all_batch_d(const float* o, const float* h, const float* l,
                                            const float* c, int32_t* out, int seriesLen,
                                            int numSeries, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    const unsigned long long all = (1ull << cdl::kCount) - 1;
    ctStatus_t r = cdl::launch(o, h, l, c, out, (size_t)seriesLen * numSeries, all, seriesLen,
                               numSeries, true, s, "cdl_all");
    if (r != CT_STATUS_SUCCESS) return r;
    return finish_batch(s);
  });
}

#define TCA_ABI(name, lb, fn)                                                                   \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_d(const float* o, const float* h, const float* l,    \
                                            const float* c, float* out, int size,              \
                                            hipStream_t s) {                                   \
    return guarded([&] {                                                                       \
      return cdl::launch(o, h, l, c, out, 0, 1ull << cdl::k_##name, size, 1, false, s,          \
                         "cdl_" #name);                                                        \
    });                                                                                        \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch_d(const float* o, const float* h,              \
                                                  const float* l, const float* c, float* out,  \
                                                  int seriesLen, int numSeries, int,           \
                                                  hipStream_t s) {                             \
    return guarded([&]() -> ctStatus_t {                                                       \
      ctStatus_t r = cdl::launch(o, h, l, c, out, 0, 1ull << cdl::k_##name, seriesLen,          \
                                 numSeries, false, s, "cdl_" #name);                           \
      if (r != CT_STATUS_SUCCESS) return r;                                                    \
      return finish_batch(s);                                                                  \
    });                                                                                        \
  }
TCA_CDL_LIST(TCA_ABI)
#undef TCA_ABI

// Host forms ct_cdl_<name> and ct_cdl_<name>_batch (tacuda.h:432+,
// :1870+): staged through the host-call buffers, the float signals of the
// device call.
#define TCA_HOST(name, lb, fn)                                                                  \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name(const float* ho, const float* hh, const float* hl,     \
                                        const float* hc, float* hy, int size, hipStream_t s) { \
    return guarded([&]() -> ctStatus_t {                                                       \
      return cdl::host_run(ho, hh, hl, hc, hy, size, 1, 1ull << cdl::k_##name, s,              \
                           "cdl_" #name);                                                      \
    });                                                                                        \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch(const float* ho, const float* hh,              \
                                                const float* hl, const float* hc, float* hy,   \
                                                int seriesLen, int numSeries, int,             \
                                                hipStream_t s) {                               \
    return guarded([&]() -> ctStatus_t {                                                       \
      return cdl::host_run(ho, hh, hl, hc, hy, seriesLen, numSeries, 1ull << cdl::k_##name, s, \
                           "cdl_" #name);                                                      \
    });                                                                                        \
  }
TCA_CDL_LIST(TCA_HOST)
#undef TCA_HOST

// The reference's legacy engulfing names (tacuda.h:447-452, :1917-1952;
// tacuda_device.h:571-582): bullish and bearish both return the signed
// engulfing signal from bar 1 on.
#define TCA_ENGULF(name)                                                                        \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_d(const float* o, const float* h, const float* l,    \
                                            const float* c, float* out, int size,              \
                                            hipStream_t s) {                                   \
    return guarded([&] { return cdl::engulf_legacy(o, h, l, c, out, size, 1, s, #name); });    \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch_d(const float* o, const float* h,              \
                                                  const float* l, const float* c, float* out,  \
                                                  int seriesLen, int numSeries, int,           \
                                                  hipStream_t s) {                             \
    return guarded([&]() -> ctStatus_t {                                                       \
      ctStatus_t r = cdl::engulf_legacy(o, h, l, c, out, seriesLen, numSeries, s, #name);      \
      if (r != CT_STATUS_SUCCESS) return r;                                                    \
      return finish_batch(s);                                                                  \
    });                                                                                        \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name(const float* ho, const float* hh, const float* hl,     \
                                        const float* hc, float* hy, int size, hipStream_t s) { \
    return guarded([&]() -> ctStatus_t {                                                       \
      return cdl::host_run(ho, hh, hl, hc, hy, size, 1, 0, s, #name);                          \
    });                                                                                        \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch(const float* ho, const float* hh,              \
                                                const float* hl, const float* hc, float* hy,   \
                                                int seriesLen, int numSeries, int,             \
                                                hipStream_t s) {                               \
    return guarded([&]() -> ctStatus_t {                                                       \
      return cdl::host_run(ho, hh, hl, hc, hy, seriesLen, numSeries, 0, s, #name);             \
    });                                                                                        \
  }
TCA_ENGULF(bullish_engulfing)
TCA_ENGULF(bearish_engulfing)
#undef TCA_ENGULF

}  // extern "C"
