// TA-Lib's MACDEXT (MACD whose fast, slow and signal lines may each be any
// of SMA, EMA, WMA, DEMA, TEMA, TRIMA, KAMA or T3) and TRIMA.
//
// MACDEXT is the reference's 0.8.3 formula backend path
// (src/compat/Talib08.cu:23-142: one thread per row running the TA-Lib
// routine in double, a leading NaN prefix skipped).  Here too each row is
// one work-item, walking its bars in double; the macd line is the
// difference of two moving averages of nearly equal magnitude, so it is
// formed in double (rounding the two averages to float32 first would cost
// up to an ulp of the price).  Two double lines per row of row scratch.
// TRIMA stands alone as a parallel window kernel (one work-item per bar).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/tacuda_amd.h"
#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"
#include "series.hpp"

namespace tca {
namespace {

// TA_MAType numbering (ctMaType_t agrees for these): 0 SMA, 1 EMA, 2 WMA,
// 3 DEMA, 4 TEMA, 5 TRIMA, 6 KAMA, 8 T3.
__host__ __device__ inline int ma_lookback(int p, int type) {
  if (p == 1) return 0;
  switch (type) {
    case 0: case 1: case 2: case 5: return p - 1;
    case 3: return 2 * (p - 1);
    case 4: return 3 * (p - 1);
    case 6: return p;
    case 8: return 6 * (p - 1);
  }
  return -1;
}

// TRIMA weights min(j+1, p-j) over the window and their reciprocal total
// (ta_TRIMA.c: 1/((p/2+1)^2) for odd p, 1/((p/2)(p/2+1)) for even p).
__host__ __device__ inline double trima_factor(int p) {
  const int i = p >> 1;
  return (p % 2 == 1) ? 1.0 / ((double)(i + 1) * (i + 1)) : 1.0 / ((double)i * (i + 1));
}

// Row cursor: in[t] as double, for float rows (the input) or double lines.
template <class T>
struct Line {
  const T* p;
  __device__ double operator[](int t) const { return (double)p[t]; }
};

// The moving average of in[s..n) written to out[s + lookback .. n) in
// double (out elsewhere untouched).  TA-Lib's arithmetic per type: SMA-seeded
// EMA cascades with fma(beta, prev, k*x), KAMA with its 2/30 constants, T3's
// six-level cascade and c1..c4 combination with volume factor 0.7, direct
// window sums for SMA, WMA and TRIMA.
template <class In>
__device__ void ma_line(const In in, double* out, int s, int n, int p, int type) {
  const int lb = ma_lookback(p, type);
  if (s + lb >= n) return;
  if (p == 1) {
    for (int t = s; t < n; ++t) out[t] = in[t];
    return;
  }
  double beta = (double)(p - 1) / (double)(p + 1);
  const double k = 1.0 - beta;
  if (beta < 0.5) beta = 1.0 - k;
  if (type == 0 || type == 2 || type == 5) {
    const double div = type == 0 ? (double)p : type == 2 ? (double)p * (p + 1) / 2.0 : 0.0;
    const double fac = type == 5 ? trima_factor(p) : 0.0;
    for (int t = s + p - 1; t < n; ++t) {
      double a = 0.0;
      for (int j = 0; j < p; ++j) {
        const double v = in[t - p + 1 + j];
        const int w = type == 0 ? 1 : type == 2 ? j + 1 : (j + 1 < p - j ? j + 1 : p - j);
        a += type == 0 ? v : v * (double)w;
      }
      out[t] = type == 5 ? a * fac : a / div;
    }
    return;
  }
  if (type == 6) {
    const double cmax = 2.0 / 31.0, cdiff = 2.0 / 3.0 - cmax;
    double prev = in[s + p - 1];
    for (int t = s + p; t < n; ++t) {
      double sum = 0.0;
      for (int j = t - p + 1; j <= t; ++j) sum += fabs(in[j] - in[j - 1]);
      const double x = in[t];
      const double roc = x - in[t - p];
      double er = 1.0;
      if (!(sum <= 0.0 || sum <= roc)) er = fmin(fabs(roc / sum), 1.0);
      double sc = fma(er, cdiff, cmax);
      sc *= sc;
      prev = fma(x - prev, sc, prev);
      out[t] = prev;
    }
    return;
  }
  // EMA-family cascades: `levels` SMA-seeded EMAs, level q seeded by the
  // mean of the first p values of level q-1.
  const bool t3 = type == 8;
  const int levels = type == 1 ? 1 : type == 3 ? 2 : type == 4 ? 3 : 6;
  const double kk = t3 ? 2.0 / (p + 1.0) : k, bb = t3 ? 1.0 - kk : beta;
  double e[6] = {0, 0, 0, 0, 0, 0};
  int today = s;
  double acc = 0.0;
  for (int i = 0; i < p; ++i) acc += in[today++];
  e[0] = acc / p;
  for (int lev = 1; lev < levels; ++lev) {
    acc = e[lev - 1];
    for (int i = 0; i < p - 1; ++i) {
      e[0] = fma(bb, e[0], kk * in[today++]);
      for (int q = 1; q < lev; ++q) e[q] = fma(bb, e[q], kk * e[q - 1]);
      acc += e[lev - 1];
    }
    e[lev] = acc / p;
  }
  const double v = (double)0.7f, v2 = v * v;
  const double c1 = -(v2 * v), c2 = 3.0 * (v2 - c1), c3 = -6.0 * v2 - 3.0 * (v - c1),
               c4 = fma(3.0, v2, fma(3.0, v, 1.0) - c1);
  auto value = [&]() -> double {
    switch (levels) {
      case 1: return e[0];
      case 2: return 2.0 * e[0] - e[1];
      case 3: return e[2] + (3.0 * e[0] - 3.0 * e[1]);
    }
    return fma(c4, e[2], fma(c3, e[3], fma(c1, e[5], c2 * e[4])));
  };
  out[today - 1] = value();
  for (; today < n; ++today) {
    e[0] = fma(bb, e[0], kk * in[today]);
    for (int q = 1; q < levels; ++q) e[q] = fma(bb, e[q], kk * e[q - 1]);
    out[today] = value();
  }
}

struct MacdExt {
  int fast, ft, slow, st, sig, sgt;
};

// One work-item per row: fast and slow averages as TA_MA computes them when
// started at L = max of their look-backs (the bars before L - lookback are
// not read, which places the recursive seeds where TA-Lib's are), the macd
// line from L in double, its signal average, float outputs from L +
// lookback(signal).
__global__ void k_macdext_rows(const float* __restrict__ x, float* __restrict__ m,
                               float* __restrict__ sg, float* __restrict__ h,
                               double* __restrict__ scratchA, double* __restrict__ scratchB,
                               int len, int rows, MacdExt q) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  const size_t off = (size_t)row * len;
  const float* xr = x + off;
  float *mr = m + off, *sr = sg + off, *hr = h + off;
  double *a = scratchA + off, *b = scratchB + off;
  const float qn = __builtin_nanf("");
  int f = 0;
  while (f < len && isnan(xr[f])) ++f;
  const int lbF = ma_lookback(q.fast, q.ft), lbS = ma_lookback(q.slow, q.st),
            lbG = ma_lookback(q.sig, q.sgt);
  const int L = lbF > lbS ? lbF : lbS;
  const int start = f + L + lbG;
  for (int t = 0; t < len && t < start; ++t) mr[t] = sr[t] = hr[t] = qn;
  if (start >= len) return;
  ma_line(Line<float>{xr}, a, f + L - lbF, len, q.fast, q.ft);
  ma_line(Line<float>{xr}, b, f + L - lbS, len, q.slow, q.st);
  for (int t = f + L; t < len; ++t) a[t] -= b[t];
  ma_line(Line<double>{a}, b, f + L, len, q.sig, q.sgt);
  for (int t = start; t < len; ++t) {
    mr[t] = (float)a[t];
    sr[t] = (float)b[t];
    hr[t] = (float)(a[t] - b[t]);
  }
}

// TRIMA, one work-item per bar: the triangular weighted window total in
// double (exact for float inputs, as TA-Lib's running sums are), NaN during
// the p-1 warm-up bars; a NaN or Inf affects only the windows holding it.
__global__ void k_trima(const float* __restrict__ x, float* __restrict__ y, int len,
                        long long total, int p, double fac) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const int t = (int)(i % len);
  if (t < p - 1) {
    y[i] = __builtin_nanf("");
    return;
  }
  const float* w = x + (i - (p - 1));
  double a = 0.0;
  for (int j = 0; j < p; ++j) {
    const int wt = j + 1 < p - j ? j + 1 : p - j;
    a += (double)wt * (double)w[j];
  }
  y[i] = (float)(a * fac);
}

ctStatus_t bad(const char* fn, const char* why) {
  return fail(CT_STATUS_INVALID_ARGUMENT, "%s: %s", fn, why);
}

ctStatus_t check_shape(const char* fn, int len, int rows) {
  if (len <= 0) return bad(fn, "size/seriesLen must be positive");
  if (rows <= 0) return bad(fn, "numSeries must be positive");
  if ((long long)len * rows > (long long)INT32_MAX * 16) return bad(fn, "matrix too large");
  return CT_STATUS_SUCCESS;
}

}  // namespace

ctStatus_t dev_trima(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  if (!x || !y) return bad("trima", "null pointer");
  ctStatus_t r = check_shape("trima", len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (p <= 0) return bad("trima", "invalid period");
  const long long total = (long long)len * rows;
  const long long blocks = (total + 255) / 256;
  hipLaunchKernelGGL(k_trima, dim3((unsigned)blocks), dim3(256), 0, s, x, y, len, total, p,
                     trima_factor(p));
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t dev_macdext(const float* x, float* m, float* sg, float* h, int len, int rows, int fast,
                       int ft, int slow, int st, int sig, int sgt, hipStream_t s) {
  if (!x || !m || !sg || !h) return bad("macdext", "null pointer");
  ctStatus_t r = check_shape("macdext", len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  for (int t : {ft, st, sgt}) {
    if (t == CT_MA_MAMA || (t > CT_MA_T3 && t <= CT_MA_ALMA))
      return fail(CT_STATUS_UNSUPPORTED, "macdext: moving-average type %d is not provided", t);
    if (t < 0 || t > CT_MA_T3) return bad("macdext", "unknown moving-average type");
  }
  // TA_MACDEXT's ranges: fast and slow from 2, signal from 1
  if (fast < 2 || slow < 2 || sig < 1) return bad("macdext", "invalid periods");
  if (slow < fast) {
    std::swap(fast, slow);
    std::swap(ft, st);
  }
  const size_t bytes = (size_t)len * rows * sizeof(double);
  void *a = nullptr, *b = nullptr;
  r = temp_for(s, bytes, 0, &a);
  if (r == CT_STATUS_SUCCESS) r = temp_for(s, bytes, 1, &b);
  if (r != CT_STATUS_SUCCESS) return r;
  const MacdExt q{fast, ft, slow, st, sig, sgt};
  hipLaunchKernelGGL(k_macdext_rows, dim3((rows + 63) / 64), dim3(64), 0, s, x, m, sg, h,
                     static_cast<double*>(a), static_cast<double*>(b), len, rows, q);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

}  // namespace tca

using namespace tca;

#define R_(x)                                   \
  do {                                          \
    ctStatus_t st_ = (x);                       \
    if (st_ != CT_STATUS_SUCCESS) return st_;   \
  } while (0)

extern "C" {

// tacuda_device.h:151, tacuda.h:134, :243, :982-991
CTAPI_EXPORT ctStatus_t ct_trima_d(const float* x, float* y, int n, int p, hipStream_t s) {
  return guarded([&] { return dev_trima(x, y, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_trima_device(const float* x, float* y, int n, int p, hipStream_t s) {
  return guarded([&] { return dev_trima(x, y, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_trima_batch_d(const float* x, float* y, int len, int rows, int p, int,
                                         hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(dev_trima(x, y, len, rows, p, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_trima(const float* hx, float* hy, int n, int p, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("trima", n, 1));
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_trima(x, y, n, 1, p, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_trima_batch(const float* hx, float* hy, int len, int rows, int p, int,
                                       hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("trima", len, rows));
    const size_t n = (size_t)len * rows;
    HostRun r(s); const float* x; float* y;
    R_(r.in(hx, n, &x)); R_(r.out(hy, n, &y)); R_(dev_trima(x, y, len, rows, p, s));
    return r.finish();
  });
}

// tacuda_device.h:388, tacuda.h:770, :1387-1404
CTAPI_EXPORT ctStatus_t ct_macdext_d(const float* x, float* m, float* sg, float* h, int n, int fast,
                                     int ft, int slow, int st, int sig, int sgt, hipStream_t s) {
  return guarded([&] { return dev_macdext(x, m, sg, h, n, 1, fast, ft, slow, st, sig, sgt, s); });
}
CTAPI_EXPORT ctStatus_t ct_macdext_batch_d(const float* x, float* m, float* sg, float* h, int len,
                                           int rows, int fast, int ft, int slow, int st, int sig,
                                           int sgt, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(dev_macdext(x, m, sg, h, len, rows, fast, ft, slow, st, sig, sgt, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_macdext(const float* hx, float* hm, float* hs, float* hh, int n,
                                   int fast, ctMaType_t ft, int slow, ctMaType_t st, int sig,
                                   ctMaType_t sgt, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("macdext", n, 1));
    HostRun r(s); const float* x; float *m, *g, *h;
    R_(r.in(hx, n, &x)); R_(r.out(hm, n, &m)); R_(r.out(hs, n, &g)); R_(r.out(hh, n, &h));
    R_(dev_macdext(x, m, g, h, n, 1, fast, (int)ft, slow, (int)st, sig, (int)sgt, s));
    return r.finish();
  });
}
CTAPI_EXPORT ctStatus_t ct_macdext_batch(const float* hx, float* hm, float* hs, float* hh, int len,
                                         int rows, int fast, int ft, int slow, int st, int sig,
                                         int sgt, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    R_(check_shape("macdext", len, rows));
    const size_t n = (size_t)len * rows;
    HostRun r(s); const float* x; float *m, *g, *h;
    R_(r.in(hx, n, &x)); R_(r.out(hm, n, &m)); R_(r.out(hs, n, &g)); R_(r.out(hh, n, &h));
    R_(dev_macdext(x, m, g, h, len, rows, fast, ft, slow, st, sig, sgt, s));
    return r.finish();
  });
}

}  // extern "C"
