// Host entry points (ct_<name>, ct_<name>_batch), the storage/compute-typed
// _ex calls, the ct_<name>_device aliases and the moving-average selector
// ct_ma / ct_ma_device.  Host calls stage through the pooled device buffers
// and pinned chunks of hostio.hpp, run the device path and wait for the
// stream (reference api.cpp:230-287).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "../../include/tacuda_amd.h"
#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"
#include "series.hpp"

namespace tca {

namespace {

ctStatus_t bad(const char* fn, const char* why) {
  return fail(CT_STATUS_INVALID_ARGUMENT, "%s: %s", fn, why);
}

ctStatus_t check_shape(const char* fn, int len, int rows) {
  if (len <= 0) return bad(fn, "size/seriesLen must be positive");
  if (rows <= 0) return bad(fn, "numSeries must be positive");
  if ((long long)len * rows > (long long)INT32_MAX * 16) return bad(fn, "matrix too large");
  return CT_STATUS_SUCCESS;
}

size_t elems(int len, int rows) { return (size_t)len * (size_t)rows; }

// osc[t] = up[t] - down[t] from bar `start` of every row, NaN before it
// (reference Aroon.cu:53-55: the oscillator of the host AROON call).
__global__ void k_aroon_osc(const float* __restrict__ up, const float* __restrict__ dn,
                            float* __restrict__ osc, int len, long long total, int start) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int t = (int)(i % len);
  osc[i] = t >= start ? up[i] - dn[i] : __builtin_nanf("");
}

ctStatus_t aroon_osc(const float* up, const float* dn, float* osc, int len, int rows, int start,
                     hipStream_t s) {
  const long long total = (long long)len * rows;
  const int threads = 256;
  const long long blocks = (total + threads - 1) / threads;
  hipLaunchKernelGGL(k_aroon_osc, dim3((unsigned)blocks), dim3(threads), 0, s, up, dn, osc, len,
                     total, start);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// Element size of a storage type, 0 when the value is not a ctDataType_t.
size_t dtype_size(ctDataType_t t) {
  return t == CT_DTYPE_F16 ? 2 : t == CT_DTYPE_F32 ? 4 : t == CT_DTYPE_F64 ? 8 : 0;
}

// The _ex calls: float32 storage with float32 or float64 compute runs the
// float32 path, whose recurrences, window sums and ratios are already
// evaluated in float64 (TA-Lib's arithmetic) and rounded once per output.
// Half and double storage are not provided.  Pairs the reference rejects
// (include/utils/Precision.h:34-42: half storage needs float32 compute)
// are rejected the same way.
ctStatus_t ex_types(const char* fn, ctDataType_t storage, ctDataType_t compute) {
  const bool valid = dtype_size(storage) && dtype_size(compute) &&
                     (storage == CT_DTYPE_F16 ? compute == CT_DTYPE_F32
                                              : (compute == CT_DTYPE_F32 || compute == CT_DTYPE_F64));
  if (!valid) return fail(CT_STATUS_UNSUPPORTED, "%s: unsupported storage/compute pair", fn);
  if (storage != CT_DTYPE_F32)
    return fail(CT_STATUS_UNSUPPORTED, "%s: only float32 storage is provided", fn);
  return CT_STATUS_SUCCESS;
}

}  // namespace

// ct_ma / ct_ma_device (reference MA.cu:23-113): the selected moving average
// (SMA, EMA, WMA, DEMA, TEMA, TRIMA, KAMA with TA-Lib's 2/30, T3 with 0.7).
ctStatus_t dev_ma(const float* x, float* y, int len, int rows, int p, int type, hipStream_t s) {
  switch (type) {
    case CT_MA_SMA: return dev_sma(x, y, len, rows, p, s);
    case CT_MA_EMA: return dev_ema(x, y, len, rows, p, s);
    case CT_MA_WMA: return dev_wma(x, y, len, rows, p, s);
    case CT_MA_DEMA: return dev_dema(x, y, len, rows, p, s);
    case CT_MA_TEMA: return dev_tema(x, y, len, rows, p, s);
    case CT_MA_TRIMA: return dev_trima(x, y, len, rows, p, s);
    case CT_MA_KAMA: return dev_kama(x, y, len, rows, p, 2, 30, s);
    case CT_MA_T3: return dev_t3(x, y, len, rows, p, 0.7f, s);
    default: break;
  }
  if (!x || !y) return bad("ma", "null pointer");
  return fail(CT_STATUS_UNSUPPORTED, "ma: moving-average type %d is not provided", type);
}

// Host AROON with separate up/down periods (reference api.cpp:2511-2570 and
// Aroon.cu:12-66): up over upPeriod, down over downPeriod, oscillator
// up - down from max(upPeriod, downPeriod); all three lines NaN when either
// period exceeds the series.
ctStatus_t dev_aroon2(const float* h, const float* l, float* up, float* dn, float* osc, int len,
                      int rows, int upP, int dnP, float* scratch, hipStream_t s) {
  if (upP <= 0 || dnP <= 0) return bad("aroon", "invalid period");
  // the unused line of each pass lands in scratch
  ctStatus_t r = dev_aroon(h, l, scratch, up, len, rows, upP, s);
  if (r != CT_STATUS_SUCCESS) return r;
  r = dev_aroon(h, l, dn, scratch, len, rows, dnP, s);
  if (r != CT_STATUS_SUCCESS) return r;
  return aroon_osc(up, dn, osc, len, rows, std::max(upP, dnP), s);
}

}  // namespace tca

using namespace tca;

#define R_(x)                                   \
  do {                                          \
    ctStatus_t st_ = (x);                       \
    if (st_ != CT_STATUS_SUCCESS) return st_;   \
  } while (0)

namespace {

void fill_nan(float* p, size_t n) {
  const float q = std::nanf("");
  for (size_t i = 0; i < n; ++i) p[i] = q;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ host single --
#define H1(NAME, DEVCALL)                                                                          \
  CTAPI_EXPORT ctStatus_t ct_##NAME(const float* hx, float* hy, int n, int p, hipStream_t s) {     \
    return guarded([&]() -> ctStatus_t {                                                           \
      R_(check_shape(#NAME, n, 1));                                                                \
      HostRun r(s);                                                                                \
      const float* x;                                                                              \
      float* y;                                                                                    \
      R_(r.in(hx, n, &x));                                                                         \
      R_(r.out(hy, n, &y));                                                                        \
      R_(DEVCALL);                                                                                 \
      return r.finish();                                                                           \
    });                                                                                            \
  }
H1(sma, dev_sma(x, y, n, 1, p, s))
H1(wma, dev_wma(x, y, n, 1, p, s))
H1(ema, dev_ema(x, y, n, 1, p, s))
H1(dema, dev_dema(x, y, n, 1, p, s))
H1(tema, dev_tema(x, y, n, 1, p, s))
H1(max, dev_max(x, y, n, 1, p, s))
H1(min, dev_min(x, y, n, 1, p, s))
H1(stddev, dev_stddev(x, y, n, 1, p, s))
H1(var, dev_var(x, y, n, 1, p, s))
H1(sum, dev_sum(x, y, n, 1, p, s))
H1(rsi, dev_rsi(x, y, n, 1, p, s))
#undef H1

CTAPI_EXPORT ctStatus_t ct_t3(const float* hx, float* hy, int n, int p, float v, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("t3", n, 1));
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_t3(x, y, n, 1, p, v, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_kama(const float* hx, float* hy, int n, int p, int fast, int slow, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("kama", n, 1));
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_kama(x, y, n, 1, p, fast, slow, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_ma(const float* hx, float* hy, int n, int p, ctMaType_t type, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ma", n, 1));
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_ma(x, y, n, 1, p, (int)type, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_macd(const float* hx, float* hm, float* hs, float* hh, int n, int fast, int slow, int sig, ctMaType_t t, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("macd", n, 1));
    HostRun r(s); const float* x; float *m, *g, *h;
    R_(r.in(hx, n, &x)); R_(r.out(hm, n, &m)); R_(r.out(hs, n, &g)); R_(r.out(hh, n, &h));
    R_(dev_macd(x, m, g, h, n, 1, fast, slow, sig, (int)t, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands(const float* hx, float* hu, float* hm, float* hl, int n, int p, float up, float dn, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("bbands", n, 1));
    HostRun r(s); const float* x; float *u, *m, *l;
    R_(r.in(hx, n, &x)); R_(r.out(hu, n, &u)); R_(r.out(hm, n, &m)); R_(r.out(hl, n, &l));
    R_(dev_bbands(x, u, m, l, n, 1, p, up, dn, s));
    return r.finish();
  });
}

#define HHLC(NAME, DEVCALL)                                                                        \
  CTAPI_EXPORT ctStatus_t ct_##NAME(const float* hh, const float* hl, const float* hc, float* hy,  \
                                    int n, int p, hipStream_t s) {                                 \
    return guarded([&]() -> ctStatus_t {                                                           \
      R_(check_shape(#NAME, n, 1));                                                                \
      HostRun r(s);                                                                                \
      const float *h, *l, *c;                                                                      \
      float* y;                                                                                    \
      R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));         \
      R_(DEVCALL);                                                                                 \
      return r.finish();                                                                           \
    });                                                                                            \
  }
HHLC(natr, dev_natr(h, l, c, y, n, 1, p, s))
HHLC(cci, dev_cci(h, l, c, y, n, 1, p, s))
HHLC(adx, dev_adx(h, l, c, y, n, 1, p, s))
HHLC(dx, dev_dx(h, l, c, y, n, 1, p, s))
HHLC(plus_di, dev_plus_di(h, l, c, y, n, 1, p, s))
HHLC(minus_di, dev_minus_di(h, l, c, y, n, 1, p, s))
HHLC(willr, dev_willr(h, l, c, y, n, 1, p, s))
#undef HHLC

CTAPI_EXPORT ctStatus_t ct_atr(const float* hh, const float* hl, const float* hc, float* hy, int n, int p, float, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("atr", n, 1));
    HostRun r(s); const float *h, *l, *c; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));
    R_(dev_atr(h, l, c, y, n, 1, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_trange(const float* hh, const float* hl, const float* hc, float* hy, int n, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("trange", n, 1));
    HostRun r(s); const float *h, *l, *c; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));
    R_(dev_trange(h, l, c, y, n, 1, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_stochastic(const float* hh, const float* hl, const float* hc, float* hk, float* hd, int n, int kp, int dp, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("stochastic", n, 1));
    HostRun r(s); const float *h, *l, *c; float *k, *d;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hk, n, &k)); R_(r.out(hd, n, &d));
    R_(dev_stoch(h, l, c, k, d, n, 1, kp, dp, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_plus_dm(const float* hh, const float* hl, float* hy, int n, int p, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("plus_dm", n, 1));
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_plus_dm(h, l, nullptr, y, n, 1, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_minus_dm(const float* hh, const float* hl, float* hy, int n, int p, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("minus_dm", n, 1));
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_minus_dm(h, l, nullptr, y, n, 1, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_aroonosc(const float* hh, const float* hl, float* hy, int n, int p, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("aroonosc", n, 1));
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_aroonosc(h, l, y, n, 1, p, s));
    return r.finish();
  });
}
// Host AROON keeps the reference's host order (up, down, osc) and its two
// periods (tacuda.h:385-388).
CTAPI_EXPORT ctStatus_t ct_aroon(const float* hh, const float* hl, float* hup, float* hdn, float* hosc, int n, int upP, int dnP, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!hh || !hl || !hup || !hdn || !hosc) return bad("aroon", "null pointer");
    R_(check_shape("aroon", n, 1));
    if (upP <= 0 || dnP <= 0) return bad("aroon", "invalid period");
    if (upP > n || dnP > n) {
      // TA-Lib's zero-count convention: every line is warm-up
      fill_nan(hup, n);
      fill_nan(hdn, n);
      fill_nan(hosc, n);
      return CT_STATUS_SUCCESS;
    }
    HostRun r(s); const float *h, *l; float *u, *d, *o; float* scratch;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hup, n, &u)); R_(r.out(hdn, n, &d)); R_(r.out(hosc, n, &o));
    void* sp = nullptr;
    R_(pool_acquire((size_t)n * 4, &sp));
    scratch = static_cast<float*>(sp);
    ctStatus_t st = dev_aroon2(h, l, u, d, o, n, 1, upP, dnP, scratch, s);
    if (st == CT_STATUS_SUCCESS) st = r.finish();
    else (void)hipStreamSynchronize(s);
    pool_release(sp);
    return st;
  });
}

// ------------------------------------------------------------- host batch --
#define HB1(NAME)                                                                                   \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch(const float* hx, float* hy, int len, int rows, int p,   \
                                            int, hipStream_t s) {                                  \
    return guarded([&]() -> ctStatus_t {                                                            \
      R_(check_shape(#NAME, len, rows));                                                            \
      const size_t n = elems(len, rows);                                                            \
      HostRun r(s);                                                                                 \
      const float* x;                                                                               \
      float* y;                                                                                     \
      R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y));                                                    \
      R_(dev_##NAME(x, y, len, rows, p, s));                                                        \
      return r.finish();                                                                            \
    });                                                                                             \
  }
HB1(sma) HB1(ema) HB1(rsi) HB1(wma) HB1(dema) HB1(tema) HB1(stddev) HB1(var) HB1(sum) HB1(max) HB1(min)
#undef HB1

#define HBHLC(NAME)                                                                                 \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch(const float* hh, const float* hl, const float* hc,      \
                                            float* hy, int len, int rows, int p, int,              \
                                            hipStream_t s) {                                       \
    return guarded([&]() -> ctStatus_t {                                                            \
      R_(check_shape(#NAME, len, rows));                                                            \
      const size_t n = elems(len, rows);                                                            \
      HostRun r(s);                                                                                 \
      const float *h, *l, *c;                                                                       \
      float* y;                                                                                     \
      R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));          \
      R_(dev_##NAME(h, l, c, y, len, rows, p, s));                                                  \
      return r.finish();                                                                            \
    });                                                                                             \
  }
HBHLC(natr) HBHLC(adx) HBHLC(dx) HBHLC(plus_di) HBHLC(minus_di) HBHLC(willr) HBHLC(cci)
#undef HBHLC

CTAPI_EXPORT ctStatus_t ct_atr_batch(const float* hh, const float* hl, const float* hc, float* hy, int len, int rows, int p, float, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("atr", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l, *c; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));
    R_(dev_atr(h, l, c, y, len, rows, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_macd_batch(const float* hx, float* hm, float* hs, float* hh, int len, int rows, int fast, int slow, int sig, int maType, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("macd", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float* x; float *m, *g, *h;
    R_(r.in(hx, n, &x)); R_(r.out(hm, n, &m)); R_(r.out(hs, n, &g)); R_(r.out(hh, n, &h));
    R_(dev_macd(x, m, g, h, len, rows, fast, slow, sig, maType, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands_batch(const float* hx, float* hu, float* hm, float* hl, int len, int rows, int p, float up, float dn, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("bbands", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float* x; float *u, *m, *l;
    R_(r.in(hx, n, &x)); R_(r.out(hu, n, &u)); R_(r.out(hm, n, &m)); R_(r.out(hl, n, &l));
    R_(dev_bbands(x, u, m, l, len, rows, p, up, dn, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_stochastic_batch(const float* hh, const float* hl, const float* hc, float* hk, float* hd, int len, int rows, int kp, int dp, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("stochastic", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l, *c; float *k, *d;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hk, n, &k)); R_(r.out(hd, n, &d));
    R_(dev_stoch(h, l, c, k, d, len, rows, kp, dp, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_kama_batch(const float* hx, float* hy, int len, int rows, int p, int fast, int slow, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("kama", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_kama(x, y, len, rows, p, fast, slow, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_t3_batch(const float* hx, float* hy, int len, int rows, int p, float v, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("t3", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_t3(x, y, len, rows, p, v, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_obv(const float* hc, const float* hv, float* hy, int n, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("obv", n, 1));
    HostRun r(s); const float *c, *v; float* y;
    R_(r.in(hc, n, &c)); R_(r.in(hv, n, &v)); R_(r.out(hy, n, &y));
    R_(dev_obv(c, v, y, n, 1, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_obv_batch(const float* hc, const float* hv, float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("obv", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *c, *v; float* y;
    R_(r.in(hc, n, &c)); R_(r.in(hv, n, &v)); R_(r.out(hy, n, &y));
    R_(dev_obv(c, v, y, len, rows, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_ad(const float* hh, const float* hl, const float* hc, const float* hv, float* hy, int n, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ad", n, 1));
    HostRun r(s); const float *h, *l, *c, *v; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.in(hv, n, &v)); R_(r.out(hy, n, &y));
    R_(dev_ad(h, l, c, v, y, n, 1, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_ad_batch(const float* hh, const float* hl, const float* hc, const float* hv, float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ad", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l, *c, *v; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.in(hv, n, &v)); R_(r.out(hy, n, &y));
    R_(dev_ad(h, l, c, v, y, len, rows, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_trange_batch(const float* hh, const float* hl, const float* hc, float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("trange", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l, *c; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.in(hc, n, &c)); R_(r.out(hy, n, &y));
    R_(dev_trange(h, l, c, y, len, rows, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_aroonosc_batch(const float* hh, const float* hl, float* hy, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("aroonosc", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_aroonosc(h, l, y, len, rows, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_aroon_batch(const float* hh, const float* hl, float* hd, float* hu, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("aroon", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l; float *d, *u;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hd, n, &d)); R_(r.out(hu, n, &u));
    R_(dev_aroon(h, l, d, u, len, rows, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_plus_dm_batch(const float* hh, const float* hl, float* hy, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("plus_dm", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_plus_dm(h, l, nullptr, y, len, rows, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_minus_dm_batch(const float* hh, const float* hl, float* hy, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("minus_dm", len, rows)); const size_t n = elems(len, rows);
    HostRun r(s); const float *h, *l; float* y;
    R_(r.in(hh, n, &h)); R_(r.in(hl, n, &l)); R_(r.out(hy, n, &y));
    R_(dev_minus_dm(h, l, nullptr, y, len, rows, p, s));
    return r.finish();
  });
}

// ----------------------------------------------------------- device aliases --
// ct_<name>_device (reference tacuda.h:202-290): the same device-pointer
// calls as ct_<name>_d, enqueued on the stream without waiting.
#define DEV1(NAME)                                                                                \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_device(const float* x, float* y, int n, int p,            \
                                             hipStream_t s) {                                   \
    return guarded([&] { return dev_##NAME(x, y, n, 1, p, s); });                                \
  }
DEV1(sma) DEV1(wma) DEV1(ema) DEV1(dema) DEV1(tema) DEV1(max) DEV1(min) DEV1(stddev) DEV1(var)
DEV1(sum) DEV1(rsi)
#undef DEV1
CTAPI_EXPORT ctStatus_t ct_t3_device(const float* x, float* y, int n, int p, float v, hipStream_t s) {
  return guarded([&] { return dev_t3(x, y, n, 1, p, v, s); });
}
CTAPI_EXPORT ctStatus_t ct_kama_device(const float* x, float* y, int n, int p, int fast, int slow, hipStream_t s) {
  return guarded([&] { return dev_kama(x, y, n, 1, p, fast, slow, s); });
}
CTAPI_EXPORT ctStatus_t ct_ma_d(const float* x, float* y, int n, int p, int type, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ma", n, 1));
    return dev_ma(x, y, n, 1, p, type, s);
  });
}
CTAPI_EXPORT ctStatus_t ct_ma_batch_d(const float* x, float* y, int len, int rows, int p, int type, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ma", len, rows));
    R_(dev_ma(x, y, len, rows, p, type, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_ma_batch(const float* hx, float* hy, int len, int rows, int p, int type, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ma", len, rows));
    const size_t n = elems(len, rows);
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_ma(x, y, len, rows, p, type, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_ma_device(const float* x, float* y, int n, int p, ctMaType_t type, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("ma", n, 1));
    return dev_ma(x, y, n, 1, p, (int)type, s);
  });
}

// ------------------------------------------------------------ typed (_ex) --
CTAPI_EXPORT ctStatus_t ct_device_upload_ex(void* d, const void* h, int count, ctDataType_t t, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    const size_t es = dtype_size(t);
    if (!d || !h || count <= 0 || !es) return bad("ct_device_upload_ex", "bad arguments");
    hipError_t e = hipMemcpyAsync(d, h, (size_t)count * es, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      hip_fail(e, "hipMemcpyAsync");
      return CT_STATUS_COPY_FAILED;
    }
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_device_download_ex(void* h, const void* d, int count, ctDataType_t t, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    const size_t es = dtype_size(t);
    if (!d || !h || count <= 0 || !es) return bad("ct_device_download_ex", "bad arguments");
    hipError_t e = hipMemcpyAsync(h, d, (size_t)count * es, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      hip_fail(e, "hipMemcpyAsync");
      return CT_STATUS_COPY_FAILED;
    }
    return CT_STATUS_SUCCESS;
  });
}

#define EX1(NAME)                                                                                  \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_ex_d(const void* x, void* y, int n, int p, ctDataType_t st,  \
                                           ctDataType_t ct, hipStream_t s) {                      \
    return guarded([&]() -> ctStatus_t {                                                           \
      if (!x || !y) return bad(#NAME "_ex", "null pointer");                                       \
      R_(ex_types(#NAME "_ex", st, ct));                                                           \
      return dev_##NAME(static_cast<const float*>(x), static_cast<float*>(y), n, 1, p, s);         \
    });                                                                                            \
  }                                                                                                \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_ex(const void* hx, void* hy, int n, int p, ctDataType_t st,  \
                                         ctDataType_t ct, hipStream_t s) {                        \
    return guarded([&]() -> ctStatus_t {                                                           \
      if (!hx || !hy) return bad(#NAME "_ex", "null pointer");                                     \
      R_(check_shape(#NAME "_ex", n, 1));                                                          \
      R_(ex_types(#NAME "_ex", st, ct));                                                           \
      HostRun r(s);                                                                                \
      const float* x;                                                                              \
      float* y;                                                                                    \
      R_(r.in(static_cast<const float*>(hx), n, &x));                                              \
      R_(r.out(static_cast<float*>(hy), n, &y));                                                   \
      R_(dev_##NAME(x, y, n, 1, p, s));                                                            \
      return r.finish();                                                                           \
    });                                                                                            \
  }
EX1(sma) EX1(ema) EX1(rsi)
#undef EX1

CTAPI_EXPORT ctStatus_t ct_atr_ex_d(const void* h, const void* l, const void* c, void* y, int n, int p, float, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!h || !l || !c || !y) return bad("atr_ex", "null pointer");
    R_(ex_types("atr_ex", st, ct));
    return dev_atr(static_cast<const float*>(h), static_cast<const float*>(l),
                   static_cast<const float*>(c), static_cast<float*>(y), n, 1, p, s);
  });
}
CTAPI_EXPORT ctStatus_t ct_atr_ex(const void* hh, const void* hl, const void* hc, void* hy, int n, int p, float, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!hh || !hl || !hc || !hy) return bad("atr_ex", "null pointer");
    R_(check_shape("atr_ex", n, 1));
    R_(ex_types("atr_ex", st, ct));
    HostRun r(s); const float *h, *l, *c; float* y;
    R_(r.in(static_cast<const float*>(hh), n, &h)); R_(r.in(static_cast<const float*>(hl), n, &l));
    R_(r.in(static_cast<const float*>(hc), n, &c)); R_(r.out(static_cast<float*>(hy), n, &y));
    R_(dev_atr(h, l, c, y, n, 1, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_macd_ex_d(const void* x, void* m, void* g, void* h, int n, int fast, int slow, int sig, int maType, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!x || !m || !g || !h) return bad("macd_ex", "null pointer");
    R_(ex_types("macd_ex", st, ct));
    return dev_macd(static_cast<const float*>(x), static_cast<float*>(m), static_cast<float*>(g),
                    static_cast<float*>(h), n, 1, fast, slow, sig, maType, s);
  });
}
CTAPI_EXPORT ctStatus_t ct_macd_ex(const void* hx, void* hm, void* hs, void* hh, int n, int fast, int slow, int sig, ctMaType_t t, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!hx || !hm || !hs || !hh) return bad("macd_ex", "null pointer");
    R_(check_shape("macd_ex", n, 1));
    R_(ex_types("macd_ex", st, ct));
    HostRun r(s); const float* x; float *m, *g, *h;
    R_(r.in(static_cast<const float*>(hx), n, &x)); R_(r.out(static_cast<float*>(hm), n, &m));
    R_(r.out(static_cast<float*>(hs), n, &g)); R_(r.out(static_cast<float*>(hh), n, &h));
    R_(dev_macd(x, m, g, h, n, 1, fast, slow, sig, (int)t, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands_ex_d(const void* x, void* u, void* m, void* l, int n, int p, float up, float dn, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!x || !u || !m || !l) return bad("bbands_ex", "null pointer");
    R_(ex_types("bbands_ex", st, ct));
    return dev_bbands(static_cast<const float*>(x), static_cast<float*>(u), static_cast<float*>(m),
                      static_cast<float*>(l), n, 1, p, up, dn, s);
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands_ex(const void* hx, void* hu, void* hm, void* hl, int n, int p, float up, float dn, ctDataType_t st, ctDataType_t ct, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!hx || !hu || !hm || !hl) return bad("bbands_ex", "null pointer");
    R_(check_shape("bbands_ex", n, 1));
    R_(ex_types("bbands_ex", st, ct));
    HostRun r(s); const float* x; float *u, *m, *l;
    R_(r.in(static_cast<const float*>(hx), n, &x)); R_(r.out(static_cast<float*>(hu), n, &u));
    R_(r.out(static_cast<float*>(hm), n, &m)); R_(r.out(static_cast<float*>(hl), n, &l));
    R_(dev_bbands(x, u, m, l, n, 1, p, up, dn, s));
    return r.finish();
  });
}

// --------------------------------------------------------- staging control --
CTAPI_EXPORT ctStatus_t ctx_host_staging(int mode) {
  return guarded([&]() -> ctStatus_t {
    if (mode < 0 || mode > 2) return bad("ctx_host_staging", "mode must be 0, 1 or 2");
    set_host_staging_mode(mode);
    return CT_STATUS_SUCCESS;
  });
}

}  // extern "C"
