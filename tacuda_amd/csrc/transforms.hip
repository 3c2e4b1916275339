// Batch-surface families beyond the fused pass, on one elementwise / window
// kernel each: the price transforms (AVGPRICE, MEDPRICE, TYPPRICE,
// WCLPRICE), the momentum family (MOMENTUM, ROC, ROCP, ROCR, ROCR100), the
// window midpoints (MIDPRICE, MIDPOINT) and the linear-regression family
// (LINEARREG, LINEARREG_SLOPE, LINEARREG_INTERCEPT, LINEARREG_ANGLE, TSF).
// Reference entry points: include/tacuda.h:111-121, 212-227, 403-424,
// 742-753, 992-1148, 1294-1312, 1334-1339, 1421-1428, 1511-1516, 1721-1727;
// include/tacuda_device.h:171-183, 338-369, 447, 511.
//
// One thread per output over the series-major matrix (rows of seriesLen
// bars), grid-stride.  Arithmetic is TA-Lib 0.8.3's in double where the
// reference vendors the function (ta_ROC.c, ta_ROCP.c, ta_MEDPRICE.c,
// ta_TYPPRICE.c, ta_MIDPRICE.c, ta_LINEARREG.c) and TA-Lib's definitions of
// the others in the same style; window extrema follow the reference's
// MIDPRICE.cu / MIDPOINT.cu (NaN samples skipped; a MIDPOINT window that
// starts inside the row's leading NaN run is NaN).  Window sums are
// evaluated directly per output in TA-Lib's priming order: for float
// inputs of ordinary range every partial sum is exact, so they equal
// TA-Lib's running sums bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <utility>

#include "../../include/tacuda.h"
#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"

namespace tca {
namespace {

constexpr int kTB = 256;  // threads per block

__device__ __forceinline__ float qnan_f() { return __int_as_float(0x7fc00000); }

dim3 grid_for(long long total) {
  const long long b = (total + kTB - 1) / kTB;
  return dim3((unsigned)(b < 65536 ? b : 65536));
}

// ------------------------------------------------------ price transforms --
enum PriceOp { kAvgPrice, kMedPrice, kTypPrice, kWclPrice };

template <int OP>
__global__ __launch_bounds__(kTB) void k_price(const float* __restrict__ o,
                                               const float* __restrict__ h,
                                               const float* __restrict__ l,
                                               const float* __restrict__ c,
                                               float* __restrict__ y, long long total) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    double v;
    if constexpr (OP == kAvgPrice) {
      v = ((double)o[i] + (double)h[i] + (double)l[i] + (double)c[i]) / 4.0;
    } else if constexpr (OP == kMedPrice) {
      v = ((double)h[i] + (double)l[i]) / 2.0;  // ta_MEDPRICE.c:70
    } else if constexpr (OP == kTypPrice) {
      v = ((double)h[i] + (double)l[i] + (double)c[i]) / 3.0;  // ta_TYPPRICE.c:68
    } else {
      v = ((double)h[i] + (double)l[i] + (double)c[i] * 2.0) / 4.0;
    }
    y[i] = (float)v;
  }
}

// -------------------------------------------------------------- momentum --
enum MomOp { kMom, kRoc, kRocp, kRocr, kRocr100, kChange };

// y[t] = f(x[t], x[t-p]) from bar p of every row (NaN before); a zero
// earlier price gives 0 for the ratios (ta_ROC.c:117-123, ta_ROCP.c:116-122).
template <int OP>
__global__ __launch_bounds__(kTB) void k_mom(const float* __restrict__ x, float* __restrict__ y,
                                             int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p) {
      y[i] = qnan_f();
      continue;
    }
    if constexpr (OP == kChange) {
      y[i] = x[i] - x[i - p];  // float difference (reference Change.cu:24-29)
      continue;
    }
    const double cur = (double)x[i], prev = (double)x[i - p];
    double v;
    if constexpr (OP == kMom) {
      v = cur - prev;
    } else if constexpr (OP == kRoc) {
      v = prev != 0.0 ? (cur / prev - 1.0) * 100.0 : 0.0;
    } else if constexpr (OP == kRocp) {
      v = prev != 0.0 ? (cur - prev) / prev : 0.0;
    } else if constexpr (OP == kRocr) {
      v = prev != 0.0 ? cur / prev : 0.0;
    } else {
      v = prev != 0.0 ? (cur / prev) * 100.0 : 0.0;
    }
    y[i] = (float)v;
  }
}

// ------------------------------------------------------ window midpoints --
// (max + min) / 2 over the p bars ending at t: MIDPRICE of high and low,
// MIDPOINT of one input (hi == lo).
template <bool ONE>
__global__ __launch_bounds__(kTB) void k_mid(const float* __restrict__ hi,
                                             const float* __restrict__ lo, float* __restrict__ y,
                                             int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p - 1) {
      y[i] = qnan_f();
      continue;
    }
    const long long s = i - (p - 1);
    float mx = hi[s], mn = lo[s];
    for (int k = 1; k < p; ++k) {
      mx = fmaxf(mx, hi[s + k]);
      mn = fminf(mn, lo[s + k]);
    }
    if constexpr (ONE) {
      // a window starting inside the row's leading NaN run is NaN
      // (reference MIDPOINT.cu:28-33)
      if (!__builtin_isnan(mx) && __builtin_isnan(hi[s])) {
        const long long r0 = i - t;
        bool lead = true;
        for (long long q = s - 1; q >= r0 && lead; --q) lead = __builtin_isnan(hi[q]);
        if (lead) {
          y[i] = qnan_f();
          continue;
        }
      }
    }
    y[i] = (float)(((double)mx + (double)mn) / 2.0);
  }
}

// ------------------------------------------------------- bar transforms --
// BOP (close - open) / (high - low) in float, 0 on a zero range (reference
// BOP.cu:14-24); IBS (close - low) / (high - low) in double, 0.5 when the
// range is not positive (TA-Lib 0.8.3 ta_IBS).  No look-back.
template <bool IBS>
__global__ __launch_bounds__(kTB) void k_bar(const float* __restrict__ o,
                                             const float* __restrict__ h,
                                             const float* __restrict__ l,
                                             const float* __restrict__ c, float* __restrict__ y,
                                             long long total) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    if constexpr (IBS) {
      const double range = (double)h[i] - (double)l[i];
      double v = ((double)c[i] - (double)l[i]) / range;
      if (range <= 0.0) v = 0.5;
      y[i] = (float)v;
    } else {
      const float d = h[i] - l[i];
      y[i] = d == 0.0f ? 0.0f : (c[i] - o[i]) / d;
    }
  }
}

// ---------------------------------------------------- window statistics --
// AVGDEV: the mean absolute deviation from the window mean over the p bars
// ending at t, in float with the reference's order (AVGDEV.cu:18-33: the
// sum oldest first, mean = sum / p, deviations summed oldest first, / p).
__global__ __launch_bounds__(kTB) void k_avgdev(const float* __restrict__ x,
                                                float* __restrict__ y, int len, long long total,
                                                int p) {
  const float fp = (float)p;
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p - 1) {
      y[i] = qnan_f();
      continue;
    }
    const float* w = x + (i - (p - 1));
    float sum = 0.0f;
    for (int k = 0; k < p; ++k) sum += w[k];
    const float mean = sum / fp;
    float dev = 0.0f;
    for (int k = 0; k < p; ++k) dev += fabsf(w[k] - mean);
    y[i] = dev / fp;
  }
}

// MINMAX: fminf / fmaxf folded from the oldest bar of the p-bar window (a
// NaN sample is skipped; an all-NaN window gives NaN; reference
// MINMAX.cu:17-31).
__global__ __launch_bounds__(kTB) void k_minmax(const float* __restrict__ x,
                                                float* __restrict__ mn, float* __restrict__ mx,
                                                int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p - 1) {
      mn[i] = qnan_f();
      mx[i] = qnan_f();
      continue;
    }
    const float* w = x + (i - (p - 1));
    float a = w[0], b = w[0];
    for (int k = 1; k < p; ++k) {
      a = fminf(a, w[k]);
      b = fmaxf(b, w[k]);
    }
    mn[i] = a;
    mx[i] = b;
  }
}

// ------------------------------------------------------------ pivots --
// PIVOT_HIGH / PIVOT_LOW (reference ClubPack.cu:357-374): bar b = t - s is
// a pivot when it is strictly above (below) each of the s bars before it
// and at or above (at or below) each of the s bars after it; the pivot's
// value is reported at t, when the s bars after it are known; NaN otherwise.
template <bool HIGH>
__global__ __launch_bounds__(kTB) void k_pivot(const float* __restrict__ x,
                                               float* __restrict__ y, int len, long long total,
                                               int st) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    float o = qnan_f();
    if (t - st >= st) {
      const long long b = i - st;
      const float v = x[b];
      bool piv = true;
      for (int k = 1; k <= st && piv; ++k)
        piv = HIGH ? (v > x[b - k] && v >= x[b + k]) : (v < x[b - k] && v <= x[b + k]);
      if (piv) o = v;
    }
    y[i] = o;
  }
}

// ------------------------------------------- TRENDSCORE, MARKET_STRENGTH --
// TRENDSCORE (reference ClubPack.cu:294-306): the sum over k = 1..10 of +1
// when x[t] >= x[t - lookback - k], else -1, from bar lookback + 10.
// MARKET_STRENGTH (ClubPack.cu:274-292): over the p one-bar changes ending
// at t (double), total / ups (total >= 0) or total / downs, x100, 0 when the
// divisor is 0; from bar p.
template <bool TREND>
__global__ __launch_bounds__(kTB) void k_club(const float* __restrict__ x, float* __restrict__ y,
                                              int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    float o = qnan_f();
    if constexpr (TREND) {
      if (t >= p + 10) {
        int sc = 0;
        for (int k = 1; k <= 10; ++k) sc += x[i] >= x[i - p - k] ? 1 : -1;
        o = (float)sc;
      }
    } else if (t >= p) {
      double tot = 0.0, ups = 0.0, dns = 0.0;
      for (int k = 0; k < p; ++k) {
        const double d = (double)x[i - k] - x[i - k - 1];
        tot += d;
        if (d > 0.0) ups += d;
        else dns += -d;
      }
      if (tot >= 0.0) o = ups > 0.0 ? (float)(100.0 * tot / ups) : 0.0f;
      else o = dns > 0.0 ? (float)(100.0 * tot / dns) : 0.0f;
    }
    y[i] = o;
  }
}

// ------------------------------------------------------------- SKEW --
// SKEW (reference ClubPack.cu:327-355, RollingSkew): over the p bars ending
// at t, newest first, the double mean, then the central moments m2 and m3
// (/ p); 0 when m2 <= 0, else m3 / m2^1.5.
__global__ __launch_bounds__(kTB) void k_skew(const float* __restrict__ x, float* __restrict__ y,
                                              int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    float o = qnan_f();
    if (t >= p - 1) {
      double mean = 0.0;
      for (int k = 0; k < p; ++k) mean += x[i - k];
      mean /= p;
      double m2 = 0.0, m3 = 0.0;
      for (int k = 0; k < p; ++k) {
        const double d = (double)x[i - k] - mean, d2 = d * d;
        m2 += d2;
        m3 += d2 * d;
      }
      m2 /= p;
      m3 /= p;
      o = m2 <= 0.0 ? 0.0f : (float)(m3 / pow(m2, 1.5));
    }
    y[i] = o;
  }
}

// ------------------------------------------------------ strict rolling --
// ROLLING_MEAN / _MAX / _MIN / _STD (reference ClubPack.cu:377-407): the p
// bars ending at t, newest first; any non-finite bar leaves the output NaN;
// sums and squares in double, mean = sum / p, std = sqrt(max(0, sq / p -
// mean^2)); extremes folded with fmaxf / fminf from x[t].
enum RollOp { kRollMean, kRollMax, kRollMin, kRollStd };

template <int OP>
__global__ __launch_bounds__(kTB) void k_rolling(const float* __restrict__ x,
                                                 float* __restrict__ y, int len, long long total,
                                                 int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    float o = qnan_f();
    if (t >= p - 1) {
      bool ok = true;
      double sum = 0.0, sq = 0.0;
      float mx = x[i], mn = x[i];
      for (int k = 0; k < p; ++k) {
        const float v = x[i - k];
        if (!__builtin_isfinite(v)) {
          ok = false;
          break;
        }
        sum += v;
        sq += (double)v * v;
        mx = fmaxf(mx, v);
        mn = fminf(mn, v);
      }
      if (ok) {
        if constexpr (OP == kRollMean) {
          o = (float)(sum / p);
        } else if constexpr (OP == kRollMax) {
          o = mx;
        } else if constexpr (OP == kRollMin) {
          o = mn;
        } else {
          const double mean = sum / p;
          o = (float)sqrt(fmax(0.0, sq / p - mean * mean));
        }
      }
    }
    y[i] = o;
  }
}

// ------------------------------------------------------ ULTOSC and MFI --
// ULTOSC (reference ULTOSC.cu:15-39): over the p3 bars ending at t, newest
// first, float buying pressure close - min(low, prev close) and true range
// max(high, prev close) - min(low, prev close) (the first bar of a row is
// its own previous close), summed into the p1-, p2- and p3-bar windows;
// 100 (4 a1 + 2 a2 + a3) / 7 with a = bp / tr (0 on a zero range).  The
// periods arrive sorted.
__global__ __launch_bounds__(kTB) void k_ultosc(const float* __restrict__ h,
                                                const float* __restrict__ l,
                                                const float* __restrict__ c,
                                                float* __restrict__ y, int len, long long total,
                                                int p1, int p2, int p3) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p3) {
      y[i] = qnan_f();
      continue;
    }
    float bp1 = 0.0f, tr1 = 0.0f, bp2 = 0.0f, tr2 = 0.0f, bp3 = 0.0f, tr3 = 0.0f;
    for (int j = 0; j < p3; ++j) {
      const long long k = i - j;
      const float pc = (t - j > 0) ? c[k - 1] : c[k];
      const float lo = fminf(l[k], pc);
      const float bp = c[k] - lo;
      const float tr = fmaxf(h[k], pc) - lo;
      if (j < p1) {
        bp1 += bp;
        tr1 += tr;
      }
      if (j < p2) {
        bp2 += bp;
        tr2 += tr;
      }
      bp3 += bp;
      tr3 += tr;
    }
    const float a1 = tr1 == 0.0f ? 0.0f : bp1 / tr1;
    const float a2 = tr2 == 0.0f ? 0.0f : bp2 / tr2;
    const float a3 = tr3 == 0.0f ? 0.0f : bp3 / tr3;
    y[i] = 100.0f * (4.0f * a1 + 2.0f * a2 + a3) / 7.0f;
  }
}

// MFI (reference MFI.cu:15-95): typical price tp = (h + l + c) / 3 in
// double, flow tp x volume counted positive when tp rose from the bar
// before and negative when it fell (non-finite prices count nothing), summed
// over the p bars ending at t; 0 when the total flow is below 1, else
// 100 pos / total.  The windows are summed directly per bar (the reference
// differences two prefix scans, which rounds differently and re-sums a
// window whose difference went negative).
__global__ __launch_bounds__(kTB) void k_mfi(const float* __restrict__ h,
                                             const float* __restrict__ l,
                                             const float* __restrict__ c,
                                             const float* __restrict__ v, float* __restrict__ y,
                                             int len, long long total, int p) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p) {
      y[i] = qnan_f();
      continue;
    }
    auto tp = [&](long long k) {
      return ((double)h[k] + (double)l[k] + (double)c[k]) / 3.0;
    };
    double pos = 0.0, neg = 0.0, prev = tp(i - p);
    for (long long k = i - p + 1; k <= i; ++k) {
      const double cur = tp(k);
      if (__builtin_isfinite(cur) && __builtin_isfinite(prev)) {
        const double f = cur * (double)v[k];
        if (cur > prev) pos += f;
        else if (cur < prev) neg += f;
      }
      prev = cur;
    }
    const double tot = pos + neg;
    y[i] = tot < 1.0 ? 0.0f : (float)(100.0 * (pos / tot));
  }
}

// ------------------------------------------------------- linear regression --
enum RegOut { kLinReg, kSlope, kIntercept, kAngle, kTsf };

// The least-squares line over the p bars ending at t with TA-Lib's sums
// (ta_LINEARREG.c: x the distance back from t, SumX and SumXSqr of 0..p-1,
// the window primed oldest first), m = (p SumXY - SumX SumY) / Divisor,
// b = (SumY - m SumX) / p; a one-bar window has slope 0 (the reference's
// denominator-zero rule, include/indicators/detail/LinearRegCommon.cuh:35-40).
template <int OUT>
__global__ __launch_bounds__(kTB) void k_linreg(const float* __restrict__ x,
                                                float* __restrict__ y, int len, long long total,
                                                int p, double sumX, double divisor) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB) {
    const int t = (int)(i % len);
    if (t < p - 1) {
      y[i] = qnan_f();
      continue;
    }
    double sumY = 0.0, sumXY = 0.0;
    for (int k = p - 1; k >= 0; --k) {
      const double v = (double)x[i - k];
      sumY += v;
      sumXY += (double)k * v;
    }
    double m, b;
    if (divisor == 0.0) {
      m = 0.0;
      b = sumY / (double)p;
    } else {
      m = ((double)p * sumXY - sumX * sumY) / divisor;
      b = (sumY - m * sumX) / (double)p;
    }
    double v;
    if constexpr (OUT == kLinReg) v = fma(m, (double)(p - 1), b);
    else if constexpr (OUT == kSlope) v = m;
    else if constexpr (OUT == kIntercept) v = b;
    else if constexpr (OUT == kAngle) v = atan(m) * (180.0 / 3.14159265358979323846);
    else v = b + m * (double)p;
    y[i] = (float)v;
  }
}

// -------------------------------------------------------------- dispatch --
ctStatus_t shape_ok(const char* fn, int len, int rows) {
  if (len <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen must be positive", fn);
  if (rows <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: numSeries must be positive", fn);
  return CT_STATUS_SUCCESS;
}

// period > (or >= for the momentum family) the row length: success with an
// all-NaN line (reference ROC.cu / MIDPRICE.cu / LINEARREG.cu, TA-Lib's
// zero-count convention)
__global__ void k_fill_nan(float* y, long long total) {
  for (long long i = (long long)blockIdx.x * kTB + threadIdx.x; i < total;
       i += (long long)gridDim.x * kTB)
    y[i] = __int_as_float(0x7fc00000);
}

ctStatus_t all_nan(float* y, long long total, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_nan, grid_for(total), dim3(kTB), 0, s, y, total);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <int OP>
ctStatus_t run_price(const char* fn, const float* o, const float* h, const float* l,
                     const float* c, float* y, int len, int rows, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  const bool needO = OP == kAvgPrice, needC = OP != kMedPrice;
  if (!h || !l || !y || (needO && !o) || (needC && !c))
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  const long long total = (long long)len * rows;
  hipLaunchKernelGGL(k_price<OP>, grid_for(total), dim3(kTB), 0, s, o, h, l, c, y, total);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <int OP>
ctStatus_t run_mom(const char* fn, const float* x, float* y, int len, int rows, int p,
                   hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p >= len) return all_nan(y, total, s);
  hipLaunchKernelGGL(k_mom<OP>, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <bool ONE>
ctStatus_t run_mid(const char* fn, const float* hi, const float* lo, float* y, int len,
                   int rows, int p, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!hi || !lo || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p > len) return all_nan(y, total, s);
  hipLaunchKernelGGL(k_mid<ONE>, grid_for(total), dim3(kTB), 0, s, hi, lo, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <int OUT>
ctStatus_t run_linreg(const char* fn, const float* x, float* y, int len, int rows, int p,
                      hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p > len) return all_nan(y, total, s);
  // ta_LINEARREG.c: SumX, SumXSqr and Divisor of x = 0..p-1
  const double sumX = (double)p * (p - 1) * 0.5;
  const double sumXSqr = (double)p * (p - 1) * (2 * p - 1) / 6.0;
  const double divisor = sumX * sumX - p * sumXSqr;
  hipLaunchKernelGGL(k_linreg<OUT>, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p, sumX,
                     divisor);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <bool IBS>
ctStatus_t run_bar(const char* fn, const float* o, const float* h, const float* l, const float* c,
                   float* y, int len, int rows, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!h || !l || !c || !y || (!IBS && !o))
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  const long long total = (long long)len * rows;
  hipLaunchKernelGGL(k_bar<IBS>, grid_for(total), dim3(kTB), 0, s, o, h, l, c, y, total);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t run_avgdev(const char* fn, const float* x, float* y, int len, int rows, int p,
                      hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p > len) return all_nan(y, total, s);
  hipLaunchKernelGGL(k_avgdev, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t run_minmax(const char* fn, const float* x, float* mn, float* mx, int len, int rows,
                      int p, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !mn || !mx) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p > len) {
    r = all_nan(mn, total, s);
    return r != CT_STATUS_SUCCESS ? r : all_nan(mx, total, s);
  }
  hipLaunchKernelGGL(k_minmax, grid_for(total), dim3(kTB), 0, s, x, mn, mx, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <bool HIGH>
ctStatus_t run_pivot(const char* fn, const float* x, float* y, int len, int rows, int st,
                     hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  // ClubPack.cu:637-639: strength > 0 and both sides inside the series
  if (st <= 0 || 2ll * st >= len) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid strength", fn);
  const long long total = (long long)len * rows;
  hipLaunchKernelGGL(k_pivot<HIGH>, grid_for(total), dim3(kTB), 0, s, x, y, len, total, st);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <bool TREND>
ctStatus_t run_club(const char* fn, const float* x, float* y, int len, int rows, int p,
                    hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  const long long total = (long long)len * rows;
  if (TREND) {
    // ClubPack.cu:574-582: lookback >= 0 and the series longer than lookback + 10
    if (p < 0 || len <= p + 10) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid lookback", fn);
  } else {
    if (p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
    if (p >= len) return all_nan(y, total, s);
  }
  hipLaunchKernelGGL(k_club<TREND>, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t run_skew(const char* fn, const float* x, float* y, int len, int rows, int p,
                    hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p <= 0 || p > len) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  hipLaunchKernelGGL(k_skew, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <int OP>
ctStatus_t run_rolling(const char* fn, const float* x, float* y, int len, int rows, int p,
                       hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!x || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  // period beyond the series is an invalid argument here (ClubPack.cu:409-413)
  if (p <= 0 || p > len) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  hipLaunchKernelGGL(k_rolling<OP>, grid_for(total), dim3(kTB), 0, s, x, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t run_ultosc(const char* fn, const float* h, const float* l, const float* c, float* y,
                      int len, int rows, int p1, int p2, int p3, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!h || !l || !c || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p1 <= 0 || p2 <= 0 || p3 <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  // sorted as TA-Lib sorts them (reference ULTOSC.cu:50-54)
  if (p1 > p2) std::swap(p1, p2);
  if (p2 > p3) std::swap(p2, p3);
  if (p1 > p2) std::swap(p1, p2);
  const long long total = (long long)len * rows;
  if (p3 >= len) return all_nan(y, total, s);
  hipLaunchKernelGGL(k_ultosc, grid_for(total), dim3(kTB), 0, s, h, l, c, y, len, total, p1, p2,
                     p3);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t run_mfi(const char* fn, const float* h, const float* l, const float* c, const float* v,
                   float* y, int len, int rows, int p, hipStream_t s) {
  ctStatus_t r = shape_ok(fn, len, rows);
  if (r != CT_STATUS_SUCCESS) return r;
  if (!h || !l || !c || !v || !y) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (p < 2) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid period", fn);
  const long long total = (long long)len * rows;
  if (p > len) return all_nan(y, total, s);
  hipLaunchKernelGGL(k_mfi, grid_for(total), dim3(kTB), 0, s, h, l, c, v, y, len, total, p);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

size_t elems(int len, int rows) {
  return (len > 0 && rows > 0) ? (size_t)len * (size_t)rows : 0;
}

}  // namespace
}  // namespace tca

using namespace tca;

#define T_RET(x)                                \
  do {                                          \
    const ctStatus_t r_ = (x);                  \
    if (r_ != CT_STATUS_SUCCESS) return r_;     \
  } while (0)

// One-input, one-period families: device, device alias, batch device, host
// and host batch forms.
#define TCA_X1P(NAME, RUN)                                                                        \
  extern "C" {                                                                                    \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_d(const float* x, float* y, int n, int p, hipStream_t s) {  \
    return guarded([&] { return RUN("ct_" #NAME, x, y, n, 1, p, s); });                           \
  }                                                                                               \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch_d(const float* x, float* y, int len, int rows, int p, \
                                              int, hipStream_t s) {                               \
    return guarded([&]() -> ctStatus_t {                                                          \
      T_RET(RUN("ct_" #NAME "_batch_d", x, y, len, rows, p, s));                                  \
      return finish_batch(s);                                                                     \
    });                                                                                           \
  }                                                                                               \
  CTAPI_EXPORT ctStatus_t ct_##NAME(const float* hx, float* hy, int n, int p, hipStream_t s) {    \
    return guarded([&]() -> ctStatus_t {                                                          \
      T_RET(shape_ok("ct_" #NAME, n, 1));                                                         \
      if (!hx || !hy) return fail(CT_STATUS_INVALID_ARGUMENT, "ct_" #NAME ": null pointer");      \
      HostRun hr(s);                                                                              \
      const float* x;                                                                             \
      float* y;                                                                                   \
      T_RET(hr.in(hx, elems(n, 1), &x));                                                          \
      T_RET(hr.out(hy, elems(n, 1), &y));                                                         \
      T_RET(RUN("ct_" #NAME, x, y, n, 1, p, s));                                                  \
      return hr.finish();                                                                         \
    });                                                                                           \
  }                                                                                               \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch(const float* hx, float* hy, int len, int rows, int p, \
                                            int, hipStream_t s) {                                 \
    return guarded([&]() -> ctStatus_t {                                                          \
      T_RET(shape_ok("ct_" #NAME "_batch", len, rows));                                           \
      if (!hx || !hy) return fail(CT_STATUS_INVALID_ARGUMENT, "ct_" #NAME "_batch: null pointer"); \
      HostRun hr(s);                                                                              \
      const float* x;                                                                             \
      float* y;                                                                                   \
      T_RET(hr.in(hx, elems(len, rows), &x));                                                     \
      T_RET(hr.out(hy, elems(len, rows), &y));                                                    \
      T_RET(RUN("ct_" #NAME "_batch", x, y, len, rows, p, s));                                    \
      return hr.finish();                                                                         \
    });                                                                                           \
  }                                                                                               \
  }

#define TCA_DEVICE_ALIAS(NAME)                                                                  \
  extern "C" CTAPI_EXPORT ctStatus_t ct_##NAME##_device(const float* x, float* y, int n, int p, \
                                                        hipStream_t s) {                        \
    return ct_##NAME##_d(x, y, n, p, s);                                                        \
  }

namespace {
template <int OP>
ctStatus_t mom(const char* fn, const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  return run_mom<OP>(fn, x, y, len, rows, p, s);
}
template <int OUT>
ctStatus_t reg(const char* fn, const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  return run_linreg<OUT>(fn, x, y, len, rows, p, s);
}
ctStatus_t midpoint(const char* fn, const float* x, float* y, int len, int rows, int p,
                    hipStream_t s) {
  return run_mid<true>(fn, x, x, y, len, rows, p, s);
}
template <int OP>
ctStatus_t roll(const char* fn, const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  return run_rolling<OP>(fn, x, y, len, rows, p, s);
}
template <bool HIGH>
ctStatus_t pivot(const char* fn, const float* x, float* y, int len, int rows, int st,
                 hipStream_t s) {
  return run_pivot<HIGH>(fn, x, y, len, rows, st, s);
}
}  // namespace

TCA_X1P(skew, run_skew)
TCA_X1P(trendscore, run_club<true>)
TCA_X1P(market_strength, run_club<false>)
TCA_X1P(pivot_high, pivot<true>)
TCA_X1P(pivot_low, pivot<false>)
TCA_X1P(rolling_mean, roll<kRollMean>)
TCA_X1P(rolling_max, roll<kRollMax>)
TCA_X1P(rolling_min, roll<kRollMin>)
TCA_X1P(rolling_std, roll<kRollStd>)
TCA_X1P(change, mom<kChange>)
TCA_DEVICE_ALIAS(change)
TCA_X1P(avgdev, run_avgdev)

TCA_X1P(momentum, mom<kMom>)
TCA_X1P(roc, mom<kRoc>)
TCA_X1P(rocp, mom<kRocp>)
TCA_X1P(rocr, mom<kRocr>)
TCA_X1P(rocr100, mom<kRocr100>)
TCA_DEVICE_ALIAS(momentum)
TCA_DEVICE_ALIAS(roc)
TCA_DEVICE_ALIAS(rocp)
TCA_DEVICE_ALIAS(rocr)
TCA_DEVICE_ALIAS(rocr100)
TCA_X1P(midpoint, midpoint)
TCA_X1P(linearreg, reg<kLinReg>)
TCA_X1P(linearreg_slope, reg<kSlope>)
TCA_X1P(linearreg_intercept, reg<kIntercept>)
TCA_X1P(linearreg_angle, reg<kAngle>)
TCA_X1P(tsf, reg<kTsf>)

// ------------------------------------------------------ price transforms --
extern "C" {

CTAPI_EXPORT ctStatus_t ct_avgprice_d(const float* o, const float* h, const float* l,
                                      const float* c, float* y, int n, hipStream_t s) {
  return guarded([&] { return run_price<kAvgPrice>("ct_avgprice", o, h, l, c, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_avgprice_batch_d(const float* o, const float* h, const float* l,
                                            const float* c, float* y, int len, int rows, int,
                                            hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_price<kAvgPrice>("ct_avgprice_batch_d", o, h, l, c, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_medprice_d(const float* h, const float* l, float* y, int n,
                                      hipStream_t s) {
  return guarded(
      [&] { return run_price<kMedPrice>("ct_medprice", nullptr, h, l, nullptr, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_medprice_batch_d(const float* h, const float* l, float* y, int len,
                                            int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_price<kMedPrice>("ct_medprice_batch_d", nullptr, h, l, nullptr, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_typprice_d(const float* h, const float* l, const float* c, float* y,
                                      int n, hipStream_t s) {
  return guarded([&] { return run_price<kTypPrice>("ct_typprice", nullptr, h, l, c, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_typprice_batch_d(const float* h, const float* l, const float* c,
                                            float* y, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_price<kTypPrice>("ct_typprice_batch_d", nullptr, h, l, c, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_wclprice_d(const float* h, const float* l, const float* c, float* y,
                                      int n, hipStream_t s) {
  return guarded([&] { return run_price<kWclPrice>("ct_wclprice", nullptr, h, l, c, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_wclprice_batch_d(const float* h, const float* l, const float* c,
                                            float* y, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_price<kWclPrice>("ct_wclprice_batch_d", nullptr, h, l, c, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_midprice_d(const float* h, const float* l, float* y, int n, int p,
                                      hipStream_t s) {
  return guarded([&] { return run_mid<false>("ct_midprice", h, l, y, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_midprice_batch_d(const float* h, const float* l, float* y, int len,
                                            int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_mid<false>("ct_midprice_batch_d", h, l, y, len, rows, p, s));
    return finish_batch(s);
  });
}

}  // extern "C"

// Host forms of the multi-input families: the inputs through the pooled
// staging of hostio, then the device form.
namespace {
template <class F>
ctStatus_t host_multi(const char* fn, int len, int rows, int nin, const float* const* hin,
                      float* hy, hipStream_t s, F&& body) {
  T_RET(shape_ok(fn, len, rows));
  for (int k = 0; k < nin; ++k)
    if (!hin[k]) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (!hy) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  HostRun hr(s);
  const float* d[4] = {nullptr, nullptr, nullptr, nullptr};
  float* y;
  for (int k = 0; k < nin; ++k) T_RET(hr.in(hin[k], elems(len, rows), &d[k]));
  T_RET(hr.out(hy, elems(len, rows), &y));
  T_RET(body(d, y));
  return hr.finish();
}
}  // namespace

extern "C" {

CTAPI_EXPORT ctStatus_t ct_avgprice(const float* ho, const float* hh, const float* hl,
                                    const float* hc, float* hy, int n, hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {ho, hh, hl, hc};
    return host_multi("ct_avgprice", n, 1, 4, in, hy, s, [&](const float* const* d, float* y) {
      return run_price<kAvgPrice>("ct_avgprice", d[0], d[1], d[2], d[3], y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_avgprice_batch(const float* ho, const float* hh, const float* hl,
                                          const float* hc, float* hy, int len, int rows, int,
                                          hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {ho, hh, hl, hc};
    return host_multi("ct_avgprice_batch", len, rows, 4, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_price<kAvgPrice>("ct_avgprice_batch", d[0], d[1], d[2], d[3],
                                                    y, len, rows, s);
                      });
  });
}
CTAPI_EXPORT ctStatus_t ct_medprice(const float* hh, const float* hl, float* hy, int n,
                                    hipStream_t s) {
  return guarded([&] {
    const float* in[2] = {hh, hl};
    return host_multi("ct_medprice", n, 1, 2, in, hy, s, [&](const float* const* d, float* y) {
      return run_price<kMedPrice>("ct_medprice", nullptr, d[0], d[1], nullptr, y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_medprice_batch(const float* hh, const float* hl, float* hy, int len,
                                          int rows, int, hipStream_t s) {
  return guarded([&] {
    const float* in[2] = {hh, hl};
    return host_multi("ct_medprice_batch", len, rows, 2, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_price<kMedPrice>("ct_medprice_batch", nullptr, d[0], d[1],
                                                    nullptr, y, len, rows, s);
                      });
  });
}
CTAPI_EXPORT ctStatus_t ct_typprice(const float* hh, const float* hl, const float* hc, float* hy,
                                    int n, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_typprice", n, 1, 3, in, hy, s, [&](const float* const* d, float* y) {
      return run_price<kTypPrice>("ct_typprice", nullptr, d[0], d[1], d[2], y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_typprice_batch(const float* hh, const float* hl, const float* hc,
                                          float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_typprice_batch", len, rows, 3, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_price<kTypPrice>("ct_typprice_batch", nullptr, d[0], d[1],
                                                    d[2], y, len, rows, s);
                      });
  });
}
CTAPI_EXPORT ctStatus_t ct_wclprice(const float* hh, const float* hl, const float* hc, float* hy,
                                    int n, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_wclprice", n, 1, 3, in, hy, s, [&](const float* const* d, float* y) {
      return run_price<kWclPrice>("ct_wclprice", nullptr, d[0], d[1], d[2], y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_wclprice_batch(const float* hh, const float* hl, const float* hc,
                                          float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_wclprice_batch", len, rows, 3, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_price<kWclPrice>("ct_wclprice_batch", nullptr, d[0], d[1],
                                                    d[2], y, len, rows, s);
                      });
  });
}
CTAPI_EXPORT ctStatus_t ct_midprice(const float* hh, const float* hl, float* hy, int n, int p,
                                    hipStream_t s) {
  return guarded([&] {
    const float* in[2] = {hh, hl};
    return host_multi("ct_midprice", n, 1, 2, in, hy, s, [&](const float* const* d, float* y) {
      return run_mid<false>("ct_midprice", d[0], d[1], y, n, 1, p, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_midprice_batch(const float* hh, const float* hl, float* hy, int len,
                                          int rows, int p, int, hipStream_t s) {
  return guarded([&] {
    const float* in[2] = {hh, hl};
    return host_multi("ct_midprice_batch", len, rows, 2, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_mid<false>("ct_midprice_batch", d[0], d[1], y, len, rows, p, s);
                      });
  });
}

}  // extern "C"

// ------------------------------------------- BOP, IBS and MINMAX forms --
extern "C" {

CTAPI_EXPORT ctStatus_t ct_bop_d(const float* o, const float* h, const float* l, const float* c,
                                 float* y, int n, hipStream_t s) {
  return guarded([&] { return run_bar<false>("ct_bop_d", o, h, l, c, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_bop_batch_d(const float* o, const float* h, const float* l,
                                       const float* c, float* y, int len, int rows, int,
                                       hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_bar<false>("ct_bop_batch_d", o, h, l, c, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_bop(const float* ho, const float* hh, const float* hl, const float* hc,
                               float* hy, int n, hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {ho, hh, hl, hc};
    return host_multi("ct_bop", n, 1, 4, in, hy, s, [&](const float* const* d, float* y) {
      return run_bar<false>("ct_bop", d[0], d[1], d[2], d[3], y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_bop_batch(const float* ho, const float* hh, const float* hl,
                                     const float* hc, float* hy, int len, int rows, int,
                                     hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {ho, hh, hl, hc};
    return host_multi("ct_bop_batch", len, rows, 4, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_bar<false>("ct_bop_batch", d[0], d[1], d[2], d[3], y, len,
                                              rows, s);
                      });
  });
}

CTAPI_EXPORT ctStatus_t ct_ibs_d(const float* h, const float* l, const float* c, float* y, int n,
                                 hipStream_t s) {
  return guarded([&] { return run_bar<true>("ct_ibs_d", nullptr, h, l, c, y, n, 1, s); });
}
CTAPI_EXPORT ctStatus_t ct_ibs_batch_d(const float* h, const float* l, const float* c, float* y,
                                       int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_bar<true>("ct_ibs_batch_d", nullptr, h, l, c, y, len, rows, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_ibs(const float* hh, const float* hl, const float* hc, float* hy,
                               int n, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_ibs", n, 1, 3, in, hy, s, [&](const float* const* d, float* y) {
      return run_bar<true>("ct_ibs", nullptr, d[0], d[1], d[2], y, n, 1, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_ibs_batch(const float* hh, const float* hl, const float* hc,
                                     float* hy, int len, int rows, int, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_ibs_batch", len, rows, 3, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_bar<true>("ct_ibs_batch", nullptr, d[0], d[1], d[2], y, len,
                                             rows, s);
                      });
  });
}

CTAPI_EXPORT ctStatus_t ct_minmax_d(const float* x, float* mn, float* mx, int n, int p,
                                    hipStream_t s) {
  return guarded([&] { return run_minmax("ct_minmax_d", x, mn, mx, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_minmax_batch_d(const float* x, float* mn, float* mx, int len, int rows,
                                          int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_minmax("ct_minmax_batch_d", x, mn, mx, len, rows, p, s));
    return finish_batch(s);
  });
}
}  // extern "C"

namespace {
ctStatus_t host_minmax(const char* fn, const float* hx, float* hmn, float* hmx, int len, int rows,
                       int p, hipStream_t s) {
  T_RET(shape_ok(fn, len, rows));
  if (!hx || !hmn || !hmx) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  HostRun hr(s);
  const float* x;
  float *mn, *mx;
  T_RET(hr.in(hx, elems(len, rows), &x));
  T_RET(hr.out(hmn, elems(len, rows), &mn));
  T_RET(hr.out(hmx, elems(len, rows), &mx));
  T_RET(run_minmax(fn, x, mn, mx, len, rows, p, s));
  return hr.finish();
}
}  // namespace

extern "C" {
CTAPI_EXPORT ctStatus_t ct_minmax(const float* hx, float* hmn, float* hmx, int n, int p,
                                  hipStream_t s) {
  return guarded([&] { return host_minmax("ct_minmax", hx, hmn, hmx, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_minmax_batch(const float* hx, float* hmn, float* hmx, int len, int rows,
                                        int p, int, hipStream_t s) {
  return guarded([&] { return host_minmax("ct_minmax_batch", hx, hmn, hmx, len, rows, p, s); });
}
}  // extern "C"

// ------------------------------------------------------ ULTOSC and MFI --
extern "C" {
CTAPI_EXPORT ctStatus_t ct_ultosc_d(const float* h, const float* l, const float* c, float* y,
                                    int n, int p1, int p2, int p3, hipStream_t s) {
  return guarded([&] { return run_ultosc("ct_ultosc_d", h, l, c, y, n, 1, p1, p2, p3, s); });
}
CTAPI_EXPORT ctStatus_t ct_ultosc_batch_d(const float* h, const float* l, const float* c, float* y,
                                          int len, int rows, int p1, int p2, int p3, int,
                                          hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_ultosc("ct_ultosc_batch_d", h, l, c, y, len, rows, p1, p2, p3, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_ultosc(const float* hh, const float* hl, const float* hc, float* hy,
                                  int n, int p1, int p2, int p3, hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_ultosc", n, 1, 3, in, hy, s, [&](const float* const* d, float* y) {
      return run_ultosc("ct_ultosc", d[0], d[1], d[2], y, n, 1, p1, p2, p3, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_ultosc_batch(const float* hh, const float* hl, const float* hc,
                                        float* hy, int len, int rows, int p1, int p2, int p3, int,
                                        hipStream_t s) {
  return guarded([&] {
    const float* in[3] = {hh, hl, hc};
    return host_multi("ct_ultosc_batch", len, rows, 3, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_ultosc("ct_ultosc_batch", d[0], d[1], d[2], y, len, rows, p1,
                                          p2, p3, s);
                      });
  });
}
CTAPI_EXPORT ctStatus_t ct_mfi_d(const float* h, const float* l, const float* c, const float* v,
                                 float* y, int n, int p, hipStream_t s) {
  return guarded([&] { return run_mfi("ct_mfi_d", h, l, c, v, y, n, 1, p, s); });
}
CTAPI_EXPORT ctStatus_t ct_mfi_batch_d(const float* h, const float* l, const float* c,
                                       const float* v, float* y, int len, int rows, int p, int,
                                       hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    T_RET(run_mfi("ct_mfi_batch_d", h, l, c, v, y, len, rows, p, s));
    return finish_batch(s);
  });
}
CTAPI_EXPORT ctStatus_t ct_mfi(const float* hh, const float* hl, const float* hc, const float* hv,
                               float* hy, int n, int p, hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {hh, hl, hc, hv};
    return host_multi("ct_mfi", n, 1, 4, in, hy, s, [&](const float* const* d, float* y) {
      return run_mfi("ct_mfi", d[0], d[1], d[2], d[3], y, n, 1, p, s);
    });
  });
}
CTAPI_EXPORT ctStatus_t ct_mfi_batch(const float* hh, const float* hl, const float* hc,
                                     const float* hv, float* hy, int len, int rows, int p, int,
                                     hipStream_t s) {
  return guarded([&] {
    const float* in[4] = {hh, hl, hc, hv};
    return host_multi("ct_mfi_batch", len, rows, 4, in, hy, s,
                      [&](const float* const* d, float* y) {
                        return run_mfi("ct_mfi_batch", d[0], d[1], d[2], d[3], y, len, rows, p, s);
                      });
  });
}
}  // extern "C"
