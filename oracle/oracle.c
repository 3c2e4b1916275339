/*
 * CPU oracle for the indicator core -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a plain-C restatement of the TA-Lib 0.8.3 algorithms that the
 * reference vendors (third_party/ta_lib_0_8_3/ta_*.c) and of the reference's
 * own semantics for the families TA-Lib C is not vendored for.  It is the
 * checker for the HIP kernels: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  No product code links it.
 *
 * Conventions shared by every entry point (the drop-in contract of the
 * reference C ABI, include/tacuda.h:32-40 and docs/IMPLEMENTATION_NOTES.md):
 *   - inputs are float32, arithmetic is double, outputs are float32 arrays of
 *     the full input length with NaN warm-up;
 *   - return 0 on success, 4 (CT_STATUS_INVALID_ARGUMENT) on a bad argument;
 *   - a period longer than the data is success with an all-NaN line;
 *   - EMA, DEMA, TEMA, T3, MACD, RSI and KAMA skip a leading NaN prefix
 *     (reference BatchFused.cu:34-40, EMA.cu:65, MACDEXT.cu:199, KAMA.cu:186)
 *     and then run the TA-Lib recurrence on the remaining samples;
 *   - window families (SMA, SUM, WMA, VAR, STDDEV, BBANDS) evaluate every
 *     window directly, so a NaN/Inf affects exactly the windows containing it
 *     (docs/IMPLEMENTATION_NOTES.md:60-84).  For float32 inputs of ordinary
 *     magnitude the double window sums are exact, so this equals TA-Lib's
 *     running totals bit for bit on finite data.
 *
 * Parity status per function is listed in the header comment of each one.
 * "pinned" names the independent check in tests/ that fixes it.
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_BAD 4
#define API __attribute__((visibility("default")))

static void fill_nan(float* o, int n) {
  for (int i = 0; i < n; ++i) o[i] = NAN;
}
static int lead_nan(const float* x, int n) {
  int f = 0;
  while (f < n && isnan(x[f])) ++f;
  return f;
}
/* TA_IS_ZERO / TA_IS_ZERO_SCALED with TA_EPSILON = 1e-14
 * (reference src/compat/Talib08Formulas.h:82-84). */
static int is_zero(double v) { return (-1e-14 < v) && (v < 1e-14); }
static int is_zero_scaled(double v, double s) { return fabs(v) <= 1e-14 * s; }

/* EMA coefficients, beta first (ta_EMA.c:88-93). */
static void ema_coef(int p, double* beta, double* k) {
  double b = (double)(p - 1) / (double)(p + 1);
  double kk = 1.0 - b;
  if (b < 0.5) b = 1.0 - kk;
  *beta = b;
  *k = kk;
}

/* ------------------------------------------------------------------ SMA ---
 * ta_SMA.c:36-117 evaluated per window (window semantics).  pinned: _ref
 * TA_SMA bit-equality on finite data; brute-force rational sums. */
API int orc_sma(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = t - p + 1; j <= t; ++j) s += (double)x[j];
    out[t] = (float)(s / (double)p);
  }
  return ORC_OK;
}

/* SUM: window total (reference SUM.cu; TA-Lib ta_SUM).  pinned: brute force. */
API int orc_sum(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = t - p + 1; j <= t; ++j) s += (double)x[j];
    out[t] = (float)s;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ WMA ---
 * ta_WMA.c:36-244: weights 1..p oldest to newest, divider p(p+1)/2.
 * pinned: _ref TA_WMA. */
API int orc_wma(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  if (p == 1) {
    for (int t = 0; t < n; ++t) out[t] = x[t];
    return ORC_OK;
  }
  const double divider = (double)p * (p + 1) / 2.0;
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    int w = 1;
    for (int j = t - p + 1; j <= t; ++j, ++w) s += (double)x[j] * w;
    out[t] = (float)(s / divider);
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ EMA ---
 * ta_EMA.c:37-147 on x[f..n) (f = leading NaNs): SMA seed of the first p
 * samples, then prev = fma(beta, prev, k*x).  pinned: _ref TA_EMA bit
 * equality; closed form on a constant and on a step input. */
API int orc_ema(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  if (p == 1) {
    for (int t = 0; t < n; ++t) out[t] = x[t];
    return ORC_OK;
  }
  const int f = lead_nan(x, n);
  const int m = n - f;
  if (p > m) return ORC_OK;
  double beta, k;
  ema_coef(p, &beta, &k);
  const float* in = x + f;
  double t = 0.0;
  for (int i = 0; i < p; ++i) t += (double)in[i];
  double prev = t / p;
  out[f + p - 1] = (float)prev;
  for (int i = p; i < m; ++i) {
    prev = fma(beta, prev, k * (double)in[i]);
    out[f + i] = (float)prev;
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------- DEMA ---
 * ta_DEMA.c:37-206: EMA2 is seeded with the mean of the first p EMA1
 * values (kept in double); out = 2*e1 - e2 from index 2(p-1).
 * pinned: _ref TA_DEMA. */
API int orc_dema(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  if (p == 1) {
    for (int t = 0; t < n; ++t) out[t] = x[t];
    return ORC_OK;
  }
  const int f = lead_nan(x, n);
  const int m = n - f;
  const int lb = 2 * (p - 1);
  if (lb >= m) return ORC_OK;
  double beta, k;
  ema_coef(p, &beta, &k);
  const float* in = x + f;
  int today = 0;
  double tmp = 0.0;
  for (int i = 0; i < p; ++i) tmp += (double)in[today++];
  double e1 = tmp / p;
  tmp = 0.0;
  tmp += e1;
  for (int i = 0; i < p - 1; ++i) {
    e1 = fma(beta, e1, k * (double)in[today++]);
    tmp += e1;
  }
  double e2 = tmp / p;
  out[f + lb] = (float)(2.0 * e1 - e2);
  for (; today < m; ++today) {
    e1 = fma(beta, e1, k * (double)in[today]);
    e2 = fma(beta, e2, k * e1);
    out[f + today] = (float)(2.0 * e1 - e2);
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------- TEMA ---
 * ta_TEMA.c:37-227: three cascaded SMA-seeded EMAs, out = e3 + (3e1 - 3e2)
 * from index 3(p-1).  pinned: _ref TA_TEMA. */
API int orc_tema(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  if (p == 1) {
    for (int t = 0; t < n; ++t) out[t] = x[t];
    return ORC_OK;
  }
  const int f = lead_nan(x, n);
  const int m = n - f;
  const int lb = 3 * (p - 1);
  if (lb >= m) return ORC_OK;
  double beta, k;
  ema_coef(p, &beta, &k);
  const float* in = x + f;
  int today = 0;
  double tmp = 0.0;
  for (int i = 0; i < p; ++i) tmp += (double)in[today++];
  double e1 = tmp / p;
  tmp = 0.0;
  tmp += e1;
  for (int i = 0; i < p - 1; ++i) {
    e1 = fma(beta, e1, k * (double)in[today++]);
    tmp += e1;
  }
  double e2 = tmp / p;
  tmp = 0.0;
  tmp += e2;
  for (int i = 0; i < p - 1; ++i) {
    e1 = fma(beta, e1, k * (double)in[today++]);
    e2 = fma(beta, e2, k * e1);
    tmp += e2;
  }
  double e3 = tmp / p;
  out[f + lb] = (float)(e3 + (3.0 * e1 - 3.0 * e2));
  for (; today < m; ++today) {
    e1 = fma(beta, e1, k * (double)in[today]);
    e2 = fma(beta, e2, k * e1);
    e3 = fma(beta, e3, k * e2);
    out[f + today] = (float)(e3 + (3.0 * e1 - 3.0 * e2));
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------- T3 ---
 * ta_T3.c:37-226: six cascaded EMAs with k = 2/(p+1), each seeded by the
 * mean of the first p values of the level below; combined with c1..c4 from
 * the volume factor.  Lookback 6(p-1).  pinned: _ref TA_T3. */
API int orc_t3(const float* x, float* out, int n, int p, double v) {
  if (!x || !out || n <= 0 || p < 2) return ORC_BAD;
  fill_nan(out, n);
  const int f = lead_nan(x, n);
  const int m = n - f;
  const int lb = 6 * (p - 1);
  if (lb >= m) return ORC_OK;
  const float* in = x + f;
  const double k = 2.0 / (p + 1.0);
  const double omk = 1.0 - k;
  int today = 0;
  double tmp = (double)in[today++];
  for (int i = p - 1; i > 0; --i) tmp += (double)in[today++];
  double e1 = tmp / p;
  tmp = e1;
  for (int i = p - 1; i > 0; --i) {
    e1 = fma(omk, e1, k * (double)in[today++]);
    tmp += e1;
  }
  double e2 = tmp / p;
  tmp = e2;
  for (int i = p - 1; i > 0; --i) {
    e1 = fma(omk, e1, k * (double)in[today++]);
    e2 = fma(omk, e2, k * e1);
    tmp += e2;
  }
  double e3 = tmp / p;
  tmp = e3;
  for (int i = p - 1; i > 0; --i) {
    e1 = fma(omk, e1, k * (double)in[today++]);
    e2 = fma(omk, e2, k * e1);
    e3 = fma(omk, e3, k * e2);
    tmp += e3;
  }
  double e4 = tmp / p;
  tmp = e4;
  for (int i = p - 1; i > 0; --i) {
    e1 = fma(omk, e1, k * (double)in[today++]);
    e2 = fma(omk, e2, k * e1);
    e3 = fma(omk, e3, k * e2);
    e4 = fma(omk, e4, k * e3);
    tmp += e4;
  }
  double e5 = tmp / p;
  tmp = e5;
  for (int i = p - 1; i > 0; --i) {
    e1 = fma(omk, e1, k * (double)in[today++]);
    e2 = fma(omk, e2, k * e1);
    e3 = fma(omk, e3, k * e2);
    e4 = fma(omk, e4, k * e3);
    e5 = fma(omk, e5, k * e4);
    tmp += e5;
  }
  double e6 = tmp / p;
  /* ta_T3.c:200-205 coefficient algebra */
  tmp = v * v;
  const double c1 = -(tmp * v);
  const double c2 = 3.0 * (tmp - c1);
  const double c3 = -6.0 * tmp - 3.0 * (v - c1);
  const double c4 = fma(3.0, tmp, fma(3.0, v, 1.0) - c1);
  /* today == lb + 1 here; the first output uses the states at lb. */
  out[f + lb] = (float)fma(c4, e3, fma(c3, e4, fma(c1, e6, c2 * e5)));
  for (; today < m; ++today) {
    e1 = fma(omk, e1, k * (double)in[today]);
    e2 = fma(omk, e2, k * e1);
    e3 = fma(omk, e3, k * e2);
    e4 = fma(omk, e4, k * e3);
    e5 = fma(omk, e5, k * e4);
    e6 = fma(omk, e6, k * e5);
    out[f + today] = (float)fma(c4, e3, fma(c3, e4, fma(c1, e6, c2 * e5)));
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------- KAMA ---
 * ta_KAMA.c:37-297 with fast/slow smoothing periods (TA-Lib fixes 2/30):
 * constMax = 2/(slow+1), constDiff = 2/(fast+1) - constMax.  The |delta| sum
 * is a window total, evaluated directly (TA-Lib runs it add-then-subtract
 * with a flat-window reset; both give 0 on a flat window).
 * pinned: _ref TA_KAMA (fast 2, slow 30). */
API int orc_kama(const float* x, float* out, int n, int p, int fast, int slow) {
  if (!x || !out || n <= 0 || p <= 0 || fast <= 0 || slow <= 0) return ORC_BAD;
  fill_nan(out, n);
  const int f = lead_nan(x, n);
  const int m = n - f;
  if (p == 1) {
    for (int t = 0; t < n; ++t) out[t] = x[t];
    return ORC_OK;
  }
  if (p >= m) return ORC_OK;
  const float* in = x + f;
  const double cmax = 2.0 / (slow + 1.0);
  const double cdiff = 2.0 / (fast + 1.0) - cmax;
  /* first KAMA at index p, previous KAMA = in[p-1] (ta_KAMA.c:148-178) */
  double prev = (double)in[p - 1];
  for (int t = p; t < m; ++t) {
    double sum = 0.0;
    for (int j = t - p + 1; j <= t; ++j) sum += fabs((double)in[j] - (double)in[j - 1]);
    const double roc = (double)in[t] - (double)in[t - p];
    double er;
    if (sum <= 0.0 || sum <= roc) {
      er = 1.0;
    } else {
      er = fabs(roc / sum);
      if (er > 1.0) er = 1.0;
    }
    double sc = fma(er, cdiff, cmax);
    sc *= sc;
    prev = fma((double)in[t] - prev, sc, prev);
    out[f + t] = (float)prev;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ RMA ---
 * ta_RMA.c:37-135: Wilder smoothing, wBeta = (p-1)/p, wAlpha = 1 - wBeta,
 * SMA seed at p-1.  pinned: reference golden fixture
 * tests/fixtures/talib08_typed.json "rma" (TA-Lib 0.8.3 output) and _ref. */
API int orc_rma(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  if (p > n) return ORC_OK;
  const double wb = (double)(p - 1) / (double)p;
  const double wa = 1.0 - wb;
  double t = 0.0;
  for (int i = 0; i < p; ++i) t += (double)x[i];
  double prev = t / p;
  out[p - 1] = (float)prev;
  for (int i = p; i < n; ++i) {
    prev = fma(wb, prev, wa * (double)x[i]);
    out[i] = (float)prev;
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------- MACD ---
 * ta_MACD.c:37-288 on x[f..]: reversed periods swapped; slow EMA seeded
 * over the first `slow` samples, fast over the fast window ENDING at slow-1;
 * signal EMA seeded by the mean of the first `signal` MACD values; all
 * three outputs start at slow+signal-2.  pinned: _ref TA_MACD. */
API int orc_macd(const float* x, float* macd, float* sig, float* hist, int n,
                 int fast, int slow, int signal) {
  if (!x || !macd || !sig || !hist || n <= 0 || fast <= 0 || slow <= 0 || signal <= 0)
    return ORC_BAD;
  fill_nan(macd, n);
  fill_nan(sig, n);
  fill_nan(hist, n);
  if (slow < fast) {
    int t = slow;
    slow = fast;
    fast = t;
  }
  const int f = lead_nan(x, n);
  const int m = n - f;
  const int lbSlow = slow - 1, lbSig = signal - 1, lbFast = fast - 1;
  const int lb = lbSlow + lbSig;
  if (lb >= m) return ORC_OK;
  const float* in = x + f;
  double sb, sk, fb, fk, gb, gk;
  ema_coef(slow, &sb, &sk);
  ema_coef(fast, &fb, &fk);
  ema_coef(signal, &gb, &gk);
  int today = 0;
  double tmp = 0.0;
  for (int i = 0; i < slow; ++i) tmp += (double)in[today++];
  double ps = tmp / slow;
  int ft = lbSlow - lbFast;
  double pf = 0.0;
  for (int i = 0; i < fast; ++i) pf += (double)in[ft++];
  pf = pf / fast;
  while (today < ft) ps = fma(sb, ps, sk * (double)in[today++]);
  while (today <= lb - lbSig) {
    const double v = (double)in[today++];
    pf = fma(fb, pf, fk * v);
    ps = fma(sb, ps, sk * v);
  }
  double mv = pf - ps;
  double psig = 0.0;
  psig += mv;
  for (int i = 0; i < signal - 1; ++i) {
    const double v = (double)in[today++];
    pf = fma(fb, pf, fk * v);
    ps = fma(sb, ps, sk * v);
    mv = pf - ps;
    psig += mv;
  }
  psig = psig / signal;
  /* today == lb + 1 */
  macd[f + lb] = (float)mv;
  sig[f + lb] = (float)psig;
  hist[f + lb] = (float)(mv - psig);
  for (; today < m; ++today) {
    const double v = (double)in[today];
    pf = fma(fb, pf, fk * v);
    ps = fma(sb, ps, sk * v);
    mv = pf - ps;
    if (signal == 1)
      psig = mv;
    else
      psig = fma(gb, psig, gk * mv);
    macd[f + today] = (float)mv;
    sig[f + today] = (float)psig;
    hist[f + today] = (float)(mv - psig);
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ RSI ---
 * ta_RSI.c:36-200 on x[f..]: Wilder averages seeded by the mean of the
 * first p gains/losses (scaled by invPeriod), updated as
 * (avg*(p-1)+delta)*invPeriod; RSI = 100*G/(G+L), 50 when G+L == 0.
 * pinned: _ref TA_RSI; flat pivot 200/3 and 100/3 (reference
 * tests/cpp/test_rsi.cpp:26-53); flat series = 50. */
API int orc_rsi(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  const int f = lead_nan(x, n);
  const int m = n - f;
  if (p >= m) return ORC_OK;
  const float* in = x + f;
  const double inv = 1.0 / (double)p;
  double prevValue = (double)in[0];
  double g = 0.0, l = 0.0;
  int today = 1;
  for (int i = p; i > 0; --i) {
    const double v = (double)in[today++];
    const double d = v - prevValue;
    prevValue = v;
    const double gd = (d > 0.0) ? d : 0.0;
    g += gd;
    l += gd - d;
  }
  l *= inv;
  g *= inv;
  {
    const double s = g + l;
    out[f + p] = (float)(s > 0.0 ? 100.0 * (g / s) : 50.0);
  }
  for (; today < m; ++today) {
    const double v = (double)in[today];
    const double d = v - prevValue;
    prevValue = v;
    l *= (double)(p - 1);
    g *= (double)(p - 1);
    const double gd = (d > 0.0) ? d : 0.0;
    g += gd;
    l += gd - d;
    l *= inv;
    g *= inv;
    const double s = g + l;
    out[f + today] = (float)(s > 0.0 ? 100.0 * (g / s) : 50.0);
  }
  return ORC_OK;
}

/* True range, ta_ATR.c:135-150 order: start from high-low, replace by the
 * previous-close distances only when strictly greater. */
static double true_range(double h, double l, double pc) {
  double g = h - l;
  double v2 = fabs(pc - h);
  if (v2 > g) g = v2;
  double v3 = fabs(pc - l);
  if (v3 > g) g = v3;
  return g;
}

/* ---------------------------------------------------------------- TRANGE ---
 * ta_TRANGE.c:36-123: NaN at bar 0, TR afterwards.  pinned: _ref. */
API int orc_trange(const float* h, const float* l, const float* c, float* out, int n) {
  if (!h || !l || !c || !out || n <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = 1; t < n; ++t) out[t] = (float)true_range(h[t], l[t], c[t - 1]);
  return ORC_OK;
}

/* ------------------------------------------------------------------ ATR ---
 * ta_ATR.c:37-222: seed = mean TR[1..p] at index p; then
 * fma(wBeta, prev, wAlpha*TR).  No leading-NaN skip (reference
 * StreamStates.h:675-706).  pinned: _ref TA_ATR; ATR == RMA(TRANGE). */
static int atr_core(const float* h, const float* l, const float* c, int n, int p,
                    double* dst /* n doubles, NaN-filled */) {
  for (int i = 0; i < n; ++i) dst[i] = NAN;
  if (p >= n) return 0;
  const double wb = (double)(p - 1) / (double)p;
  const double wa = 1.0 - wb;
  double tot = 0.0;
  int today = 1;
  for (int i = 0; i < p; ++i, ++today) tot += true_range(h[today], l[today], c[today - 1]);
  double prev = tot / p;
  dst[p] = prev;
  for (; today < n; ++today) {
    prev = fma(wb, prev, wa * true_range(h[today], l[today], c[today - 1]));
    dst[today] = prev;
  }
  return 1;
}

/* Period 1 of ATR and NATR is TRANGE: unnormalized, NaN at bar 0, and a NaN
 * bar does not poison later bars (reference docs/BATCH_API.md:155-159,
 * NATR.cu:171-174, ATR.cu:163; classic TA-Lib's period-1 TRANGE arm). */
static void trange_p1(const float* h, const float* l, const float* c, float* out, int n) {
  out[0] = NAN;
  for (int t = 1; t < n; ++t) out[t] = (float)true_range(h[t], l[t], c[t - 1]);
}

API int orc_atr(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  if (p == 1) {
    trange_p1(h, l, c, out, n);
    return ORC_OK;
  }
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
  atr_core(h, l, c, n, p, tmp);
  for (int i = 0; i < n; ++i) out[i] = (float)tmp[i];
  free(tmp);
  return ORC_OK;
}

/* NATR = ATR / close * 100, 0 where close is zero (TA_IS_ZERO); reference
 * NATR.cu:42-231 and docs/IMPLEMENTATION_NOTES.md:46-58.  pinned: derived
 * from the pinned ATR by an independent division in the test. */
API int orc_natr(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  if (p == 1) {
    trange_p1(h, l, c, out, n);
    return ORC_OK;
  }
  double* tmp = (double*)malloc(sizeof(double) * (size_t)n);
  atr_core(h, l, c, n, p, tmp);
  for (int i = 0; i < n; ++i) {
    if (i < p) {
      out[i] = NAN;
      continue;
    }
    const double cl = (double)c[i];
    out[i] = is_zero(cl) ? 0.0f : (float)((tmp[i] / cl) * 100.0);
  }
  free(tmp);
  return ORC_OK;
}

/* ------------------------------------------------------- VAR / STDDEV ---
 * Population variance of the window with the data shifted by the window's
 * oldest sample, E[d^2] - E[d]^2, floored to 0 below 1e-12*E[d^2] (the floor
 * of ta_BBANDS.c:236-239 / var.c).  Exactly 0 on a constant window.
 * pinned: _ref TA_VAR/TA_STDDEV within tolerance; two-pass brute force. */
static double window_var(const float* x, int t, int p) {
  const double shift = (double)x[t - p + 1];
  const double inv = 1.0 / (double)p;
  double s1 = 0.0, s2 = 0.0;
  for (int j = t - p + 1; j <= t; ++j) {
    double d = (double)x[j] - shift;
    s1 += d;
    s2 += d * d;
  }
  const double mean = s1 * inv;
  const double e2 = s2 * inv;
  double var = e2 - mean * mean;
  if (var < 1e-12 * e2) var = 0.0;
  if (!(var >= 0.0) && !isnan(var)) var = 0.0;
  return var;
}

API int orc_var(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) out[t] = (float)window_var(x, t, p);
  return ORC_OK;
}

API int orc_stddev(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) out[t] = (float)sqrt(window_var(x, t, p));
  return ORC_OK;
}

/* --------------------------------------------------------------- BBANDS ---
 * ta_BBANDS.c:37-397, SMA middle band: mid = SMA, sd = population std;
 * equal multipliers: mid +/- sd*k, else fma(sd, up, mid) and mid - sd*dn.
 * pinned: _ref TA_BBANDS within tolerance. */
API int orc_bbands(const float* x, float* up, float* mid, float* lo, int n, int p,
                   double kup, double kdn) {
  if (!x || !up || !mid || !lo || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(up, n);
  fill_nan(mid, n);
  fill_nan(lo, n);
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = t - p + 1; j <= t; ++j) s += (double)x[j];
    const double m = s / p;
    const double var = window_var(x, t, p);
    const double sd = (var != 0.0) ? sqrt(var) : 0.0;
    mid[t] = (float)m;
    if (kup == kdn) {
      const double d = sd * kup;
      up[t] = (float)(m + d);
      lo[t] = (float)(m - d);
    } else {
      up[t] = (float)fma(sd, kup, m);
      lo[t] = (float)(m - sd * kdn);
    }
  }
  return ORC_OK;
}

/* Rolling extrema with the comparison order of the reference WILLR.cu:21-27:
 * start from the newest bar, replace only on strict improvement. */
static void window_hl(const float* h, const float* l, int t, int p, float* hh, float* ll) {
  float hi = h[t], lw = l[t];
  for (int j = 1; j < p; ++j) {
    const float a = h[t - j], b = l[t - j];
    if (a > hi) hi = a;
    if (b < lw) lw = b;
  }
  *hh = hi;
  *ll = lw;
}

/* ---------------------------------------------------------------- STOCH ---
 * ta_STOCH.c:36-322 with SMA smoothing of both stages over dPeriod (the
 * reference ct_stochastic_d ABI: kPeriod, dPeriod): fastK = (C-LL)/(HH-LL)
 * *100, 0 when HH-LL is zero against |HH|+|LL| (ta_STOCH.c:262); both
 * outputs start at K-1 + 2(D-1).  pinned: _ref TA_STOCH within tolerance;
 * ramp closed form 100K/(K+1) (reference
 * tests/cpp/test_indicator_audit_regressions.cpp:64-100). */
API int orc_stoch(const float* h, const float* l, const float* c, float* k, float* d,
                  int n, int kp, int dp) {
  if (!h || !l || !c || !k || !d || n <= 0 || kp <= 0 || dp <= 0) return ORC_BAD;
  fill_nan(k, n);
  fill_nan(d, n);
  const int lb = kp - 1 + 2 * (dp - 1);
  if (lb >= n) return ORC_OK;
  double* fk = (double*)malloc(sizeof(double) * (size_t)n);
  double* sk = (double*)malloc(sizeof(double) * (size_t)n);
  for (int t = kp - 1; t < n; ++t) {
    float hh, ll;
    window_hl(h, l, t, kp, &hh, &ll);
    const double range = (double)hh - (double)ll;
    if (!is_zero_scaled(range, fabs((double)hh) + fabs((double)ll)))
      fk[t] = ((double)c[t] - (double)ll) / range * 100.0;
    else
      fk[t] = 0.0;
  }
  for (int t = kp - 1 + dp - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = t - dp + 1; j <= t; ++j) s += fk[j];
    sk[t] = s / dp;
  }
  for (int t = lb; t < n; ++t) {
    double s = 0.0;
    for (int j = t - dp + 1; j <= t; ++j) s += sk[j];
    k[t] = (float)sk[t];
    d[t] = (float)(s / dp);
  }
  free(fk);
  free(sk);
  return ORC_OK;
}

/* ---------------------------------------------------------------- WILLR ---
 * Reference WILLR.cu:12-30 (TA-Lib WILLR not vendored): extrema in float,
 * ratio in double, 0 on a zero range, clamped to [-100, 0].
 * pinned: fast %K - 100 of the vendored TA_STOCHF (_ref) within float
 * rounding; range, flat-window and brute-force checks. */
API int orc_willr(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    float hh, ll;
    window_hl(h, l, t, p, &hh, &ll);
    const double den = (double)hh - (double)ll;
    const double v = den == 0.0 ? 0.0 : (((double)hh - (double)c[t]) / den) * -100.0;
    out[t] = isnan(v) ? NAN : (float)fmin(0.0, fmax(-100.0, v));
  }
  return ORC_OK;
}

/* ------------------------------------------------------------------ CCI ---
 * Reference CCI.cu:11-37: typical price ((H+L)+C)/3 in double, window mean,
 * mean absolute deviation; 0 unless numerator and deviation are nonzero.
 * parity vs TA-Lib unpinned; pinned by constant-window zero and a
 * brute-force restatement in numpy. */
API int orc_cci(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = 0; j < p; ++j) {
      const int i = t - j;
      s += ((double)h[i] + (double)l[i] + (double)c[i]) / 3.0;
    }
    const double sma = s / p;
    double dev = 0.0;
    for (int j = 0; j < p; ++j) {
      const int i = t - j;
      const double tp = ((double)h[i] + (double)l[i] + (double)c[i]) / 3.0;
      dev += fabs(tp - sma);
    }
    const double md = dev / p;
    const double tpc = ((double)h[t] + (double)l[t] + (double)c[t]) / 3.0;
    const double num = tpc - sma;
    out[t] = (num != 0.0 && md != 0.0) ? (float)(num / (0.015 * md)) : 0.0f;
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------- AROON ---
 * TA-Lib AROON semantics as restated by reference Aroon.cu:12-50: window of
 * p+1 bars, newest extremum wins ties, value = (100/p)*(p - bars since).
 * parity vs TA-Lib unpinned; pinned by the {100k/p} value set and a
 * brute-force argmax restatement. */
API int orc_aroon(const float* h, const float* l, float* down, float* up, int n, int p) {
  if (!h || !l || !down || !up || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(down, n);
  fill_nan(up, n);
  const double factor = 100.0 / (double)p;
  for (int t = p; t < n; ++t) {
    int sh = 0, sl = 0;
    float mx = h[t], mn = l[t];
    for (int j = 1; j <= p; ++j) {
      if (h[t - j] > mx) {
        mx = h[t - j];
        sh = j;
      }
      if (l[t - j] < mn) {
        mn = l[t - j];
        sl = j;
      }
    }
    up[t] = (float)(factor * (double)(p - sh));
    down[t] = (float)(factor * (double)(p - sl));
  }
  return ORC_OK;
}

API int orc_aroonosc(const float* h, const float* l, float* out, int n, int p) {
  if (!h || !l || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  const double factor = 100.0 / (double)p;
  for (int t = p; t < n; ++t) {
    int sh = 0, sl = 0;
    float mx = h[t], mn = l[t];
    for (int j = 1; j <= p; ++j) {
      if (h[t - j] > mx) {
        mx = h[t - j];
        sh = j;
      }
      if (l[t - j] < mn) {
        mn = l[t - j];
        sl = j;
      }
    }
    /* TA-Lib AROONOSC: factor * (index of high - index of low) */
    out[t] = (float)(factor * (double)(sl - sh));
  }
  return ORC_OK;
}

/* ------------------------------------------------------------ MIN / MAX ---
 * Reference MAX.cu:12-30 / MIN.cu: NaN-ignoring window extremum (fmaxf),
 * except a window that starts inside the leading NaN prefix is NaN
 * (LeadingNaN.cuh:22-28).  pinned: bit-equal to the vendored TA_DONCHIAN
 * bands (_ref) on finite data; numpy nanmax brute force for NaN rules. */
static int in_lead_prefix(const float* x, int idx) {
  if (!isnan(x[0])) return 0;
  for (int i = idx - 1; i > 0; --i)
    if (!isnan(x[i])) return 0;
  return 1;
}

API int orc_max(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int s = 0; s + p - 1 < n; ++s) {
    float m = x[s];
    for (int k = 1; k < p; ++k) m = fmaxf(m, x[s + k]);
    if (!isnan(m) && isnan(x[s]) && in_lead_prefix(x, s)) m = NAN;
    out[s + p - 1] = m;
  }
  return ORC_OK;
}

API int orc_min(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int s = 0; s + p - 1 < n; ++s) {
    float m = x[s];
    for (int k = 1; k < p; ++k) m = fminf(m, x[s + k]);
    if (!isnan(m) && isnan(x[s]) && in_lead_prefix(x, s)) m = NAN;
    out[s + p - 1] = m;
  }
  return ORC_OK;
}

/* ------------------------------------------------------ directional move ---
 * TA-Lib DM/DI/DX/ADX (ta_PLUS_DM, ta_PLUS_DI, ta_DX, ta_ADX; not vendored).
 * Per bar t >= 1: up = H[t]-H[t-1], dn = L[t-1]-L[t]; +DM = up if up > 0 and
 * up > dn, -DM = dn if dn > 0 and dn > up (reference BatchShort.cu:181-185).
 * Wilder sums: seed = sum over bars 1..p-1, then prev - prev/p + x.
 * DI = 100*DM/TR (0 when TR is zero), DX = 100*|+DI - -DI|/(+DI + -DI)
 * (0 when undefined, reference BatchShort.cu:206-210), ADX = mean of the
 * first p DX values at 2p-1, then (ADX*(p-1)+DX)/p.
 * parity vs TA-Lib unpinned; pinned by DI/DX/ADX range invariants, DX
 * scale invariance, flat data, and a brute-force restatement. */
static void dm_bar(const float* h, const float* l, int t, double* pdm, double* mdm) {
  const double up = (double)h[t] - (double)h[t - 1];
  const double dn = (double)l[t - 1] - (double)l[t];
  *pdm = (up > 0.0 && up > dn) ? up : 0.0;
  *mdm = (dn > 0.0 && dn > up) ? dn : 0.0;
}

enum { DIR_PDM, DIR_MDM, DIR_PDI, DIR_MDI, DIR_DX, DIR_ADX };

static int directional(const float* h, const float* l, const float* c, float* out, int n,
                       int p, int what) {
  fill_nan(out, n);
  if (p == 1) {
    /* TA-Lib period-1 contracts: raw DM / DI / DX from bar 1. */
    for (int t = 1; t < n; ++t) {
      double pd, md;
      dm_bar(h, l, t, &pd, &md);
      if (what == DIR_PDM) {
        out[t] = (float)pd;
        continue;
      }
      if (what == DIR_MDM) {
        out[t] = (float)md;
        continue;
      }
      const double tr = true_range(h[t], l[t], c[t - 1]);
      const double pdi = is_zero(tr) ? 0.0 : 100.0 * (pd / tr);
      const double mdi = is_zero(tr) ? 0.0 : 100.0 * (md / tr);
      if (what == DIR_PDI) {
        out[t] = (float)pdi;
        continue;
      }
      if (what == DIR_MDI) {
        out[t] = (float)mdi;
        continue;
      }
      const double s = pdi + mdi;
      const double dx = is_zero(s) ? 0.0 : 100.0 * (fabs(mdi - pdi) / s);
      if (what == DIR_DX || what == DIR_ADX) out[t] = (float)dx;
    }
    return ORC_OK;
  }
  const int dmOnly = (what == DIR_PDM || what == DIR_MDM);
  const int lbAll = dmOnly ? p - 1 : (what == DIR_ADX ? 2 * p - 1 : p);
  if (lbAll >= n) return ORC_OK;
  const double pp = (double)p;
  double sp = 0.0, sm = 0.0, st = 0.0;
  for (int t = 1; t <= p - 1; ++t) {
    double pd, md;
    dm_bar(h, l, t, &pd, &md);
    sp += pd;
    sm += md;
    st += true_range(h[t], l[t], c[t - 1]);
  }
  if (dmOnly) out[p - 1] = (float)(what == DIR_PDM ? sp : sm);
  double adx = 0.0, sumdx = 0.0;
  for (int t = p; t < n; ++t) {
    double pd, md;
    dm_bar(h, l, t, &pd, &md);
    sp = sp - sp / pp + pd;
    sm = sm - sm / pp + md;
    st = st - st / pp + true_range(h[t], l[t], c[t - 1]);
    if (what == DIR_PDM) {
      out[t] = (float)sp;
      continue;
    }
    if (what == DIR_MDM) {
      out[t] = (float)sm;
      continue;
    }
    const double pdi = is_zero(st) ? 0.0 : 100.0 * (sp / st);
    const double mdi = is_zero(st) ? 0.0 : 100.0 * (sm / st);
    if (what == DIR_PDI) {
      out[t] = (float)pdi;
      continue;
    }
    if (what == DIR_MDI) {
      out[t] = (float)mdi;
      continue;
    }
    const double s = pdi + mdi;
    const double dx = is_zero(s) ? 0.0 : 100.0 * (fabs(mdi - pdi) / s);
    if (what == DIR_DX) {
      out[t] = (float)dx;
      continue;
    }
    if (t < 2 * p - 1) {
      sumdx += dx;
    } else if (t == 2 * p - 1) {
      sumdx += dx;
      adx = sumdx / pp;
      out[t] = (float)adx;
    } else {
      adx = ((adx * (pp - 1.0)) + dx) / pp;
      out[t] = (float)adx;
    }
  }
  return ORC_OK;
}

API int orc_plus_dm(const float* h, const float* l, float* out, int n, int p) {
  if (!h || !l || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, h, out, n, p, DIR_PDM);
}
API int orc_minus_dm(const float* h, const float* l, float* out, int n, int p) {
  if (!h || !l || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, h, out, n, p, DIR_MDM);
}
API int orc_plus_di(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, c, out, n, p, DIR_PDI);
}
API int orc_minus_di(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, c, out, n, p, DIR_MDI);
}
API int orc_dx(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, c, out, n, p, DIR_DX);
}
API int orc_adx(const float* h, const float* l, const float* c, float* out, int n, int p) {
  if (!h || !l || !c || !out || n <= 0 || p <= 0) return ORC_BAD;
  return directional(h, l, c, out, n, p, DIR_ADX);
}

/* ------------------------------------------------------------ OBV / AD ---
 * Reference OBV.cu / AD.cu semantics as restated by its stream states
 * (include/utils/StreamStates.h:442-505): OBV starts from bar 0's raw volume
 * and adds sign(close_t - close_{t-1}) * volume after; AD adds
 * (((c-l)-(h-c))/(h-l))*v for finite bars with a positive range, 0
 * otherwise.  Both accumulate in double from 0.0.  parity vs TA-Lib
 * unpinned (not vendored); pinned by brute force and sign invariants. */
API int orc_obv(const float* c, const float* v, float* out, int n) {
  if (!c || !v || !out || n <= 0) return ORC_BAD;
  double acc = 0.0;
  for (int t = 0; t < n; ++t) {
    if (t == 0) {
      acc += (double)v[0];
    } else {
      const double p = (double)c[t], prev = (double)c[t - 1];
      double s = 0.0;
      if (p > prev) s = 1.0;
      else if (p < prev) s = -1.0;
      acc += s * (double)v[t];
    }
    out[t] = (float)acc;
  }
  return ORC_OK;
}

API int orc_ad(const float* h, const float* l, const float* c, const float* v, float* out,
               int n) {
  if (!h || !l || !c || !v || !out || n <= 0) return ORC_BAD;
  double acc = 0.0;
  for (int t = 0; t < n; ++t) {
    const double hi = h[t], lo = l[t], cl = c[t], vo = v[t];
    double term = 0.0;
    if (isfinite(hi) && isfinite(lo) && isfinite(cl) && isfinite(vo)) {
      const double range = hi - lo;
      if (range > 0.0) term = (((cl - lo) - (hi - cl)) / range) * vo;
    }
    acc += term;
    out[t] = (float)acc;
  }
  return ORC_OK;
}

/* ------------------------------------------------- batch convenience ---
 * Row loop over a series-major [rows x len] matrix, used by the tests and
 * the CPU baseline to time the oracle over many series. */
API int orc_ema_rows(const float* x, float* out, int len, int rows, int p) {
  for (int r = 0; r < rows; ++r) {
    int rc = orc_ema(x + (size_t)r * len, out + (size_t)r * len, len, p);
    if (rc) return rc;
  }
  return ORC_OK;
}

/* ---------------------------------------------------------------- TRIMA ---
 * ta_TRIMA.c: triangular weights min(j+1, p-j) over the window, times
 * 1/((p/2+1)^2) for odd p and 1/((p/2)(p/2+1)) for even p.  TA-Lib keeps the
 * weighted total as add-then-subtract running sums of the float32 inputs
 * (exact in double), so the direct weighted sum gives the same double.
 * Window semantics for NaN.  pinned: _ref TA_MA(TRIMA) bit equality. */
API int orc_trima(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  const int i = p >> 1;
  const double factor = (p % 2 == 1) ? 1.0 / ((double)(i + 1) * (i + 1))
                                     : 1.0 / ((double)i * (i + 1));
  for (int t = p - 1; t < n; ++t) {
    double s = 0.0;
    for (int j = 0; j < p; ++j) {
      const int w = (j + 1 < p - j) ? j + 1 : p - j;
      s += (double)w * (double)x[t - p + 1 + j];
    }
    out[t] = (float)(s * factor);
  }
  return ORC_OK;
}

/* TA_MAType numbering: 0 SMA, 1 EMA, 2 WMA, 3 DEMA, 4 TEMA, 5 TRIMA, 6 KAMA,
 * 8 T3 (7 MAMA is not restated). */
static int ma_lookback(int p, int type) {
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
/* Moving averages in double over in[s..n) (the bars before s are not
 * part of the series), for MACDEXT, whose macd line is the difference of
 * two of them.  Each follows the float32 oracle function of its type above
 * (TA-Lib's arithmetic), with double outputs; out[t] is NaN before s +
 * lookback. */
static void md_ma(const double* in, double* out, int s, int n, int p, int type) {
  const double qnan = NAN;
  const int lb = ma_lookback(p, type);
  for (int t = 0; t < n; ++t) out[t] = qnan;
  if (s + lb >= n) return;
  if (p == 1) { /* TA_MA copies the input for a one-bar period */
    for (int t = s; t < n; ++t) out[t] = in[t];
    return;
  }
  double beta, k;
  ema_coef(p, &beta, &k);
  switch (type) {
    case 0: /* SMA */
      for (int t = s + p - 1; t < n; ++t) {
        double a = 0.0;
        for (int j = t - p + 1; j <= t; ++j) a += in[j];
        out[t] = a / p;
      }
      return;
    case 2: { /* WMA */
      const double divider = (double)p * (p + 1) / 2.0;
      for (int t = s + p - 1; t < n; ++t) {
        double a = 0.0;
        int w = 1;
        for (int j = t - p + 1; j <= t; ++j, ++w) a += in[j] * w;
        out[t] = a / divider;
      }
      return;
    }
    case 5: { /* TRIMA */
      const int i = p >> 1;
      const double factor = (p % 2 == 1) ? 1.0 / ((double)(i + 1) * (i + 1))
                                         : 1.0 / ((double)i * (i + 1));
      for (int t = s + p - 1; t < n; ++t) {
        double a = 0.0;
        for (int j = 0; j < p; ++j) {
          const int w = (j + 1 < p - j) ? j + 1 : p - j;
          a += (double)w * in[t - p + 1 + j];
        }
        out[t] = a * factor;
      }
      return;
    }
    case 1: case 3: case 4: { /* EMA, DEMA, TEMA: SMA-seeded EMA cascade */
      const int levels = type == 1 ? 1 : type == 3 ? 2 : 3;
      double e[3] = {0.0, 0.0, 0.0};
      int today = s;
      double tmp = 0.0;
      for (int i = 0; i < p; ++i) tmp += in[today++];
      e[0] = tmp / p;
      for (int lev = 1; lev < levels; ++lev) {
        tmp = e[lev - 1];
        for (int i = 0; i < p - 1; ++i) {
          const double xv = in[today++];
          e[0] = fma(beta, e[0], k * xv);
          for (int q = 1; q < lev; ++q) e[q] = fma(beta, e[q], k * e[q - 1]);
          tmp += e[lev - 1];
        }
        e[lev] = tmp / p;
      }
      out[today - 1] = levels == 1 ? e[0] : levels == 2 ? 2.0 * e[0] - e[1]
                                                        : e[2] + (3.0 * e[0] - 3.0 * e[1]);
      for (; today < n; ++today) {
        e[0] = fma(beta, e[0], k * in[today]);
        for (int q = 1; q < levels; ++q) e[q] = fma(beta, e[q], k * e[q - 1]);
        out[today] = levels == 1 ? e[0] : levels == 2 ? 2.0 * e[0] - e[1]
                                                      : e[2] + (3.0 * e[0] - 3.0 * e[1]);
      }
      return;
    }
    case 6: { /* KAMA 2/30 */
      const double cmax = 2.0 / 31.0, cdiff = 2.0 / 3.0 - cmax;
      double prev = in[s + p - 1];
      for (int t = s + p; t < n; ++t) {
        double sum = 0.0;
        for (int j = t - p + 1; j <= t; ++j) sum += fabs(in[j] - in[j - 1]);
        const double roc = in[t] - in[t - p];
        double er;
        if (sum <= 0.0 || sum <= roc) {
          er = 1.0;
        } else {
          er = fabs(roc / sum);
          if (er > 1.0) er = 1.0;
        }
        double sc = fma(er, cdiff, cmax);
        sc *= sc;
        prev = fma(in[t] - prev, sc, prev);
        out[t] = prev;
      }
      return;
    }
    case 8: { /* T3, volume factor 0.7 (as TA_MA passes it) */
      const double v = (double)0.7f;
      const double kk = 2.0 / (p + 1.0), omk = 1.0 - kk;
      double e[6];
      int today = s;
      for (int lev = 0; lev < 6; ++lev) {
        double tmp = 0.0;
        for (int i = 0; i < p; ++i) {
          if (lev == 0) {
            tmp += in[today++];
            continue;
          }
          if (i == 0) {
            tmp = e[lev - 1];
            continue;
          }
          const double xv = in[today++];
          e[0] = fma(omk, e[0], kk * xv);
          for (int q = 1; q < lev; ++q) e[q] = fma(omk, e[q], kk * e[q - 1]);
          tmp += e[lev - 1];
        }
        e[lev] = tmp / p;
      }
      const double v2 = v * v;
      const double c1 = -(v2 * v);
      const double c2 = 3.0 * (v2 - c1);
      const double c3 = -6.0 * v2 - 3.0 * (v - c1);
      const double c4 = fma(3.0, v2, fma(3.0, v, 1.0) - c1);
      out[today - 1] = fma(c4, e[2], fma(c3, e[3], fma(c1, e[5], c2 * e[4])));
      for (; today < n; ++today) {
        e[0] = fma(omk, e[0], kk * in[today]);
        for (int q = 1; q < 6; ++q) e[q] = fma(omk, e[q], kk * e[q - 1]);
        out[today] = fma(c4, e[2], fma(c3, e[3], fma(c1, e[5], c2 * e[4])));
      }
      return;
    }
  }
}

/* -------------------------------------------------------------- MACDEXT ---
 * ta_MACDEXT.c with the reference's formula backend convention
 * (src/compat/Talib08.cu:23-142: a leading NaN prefix is skipped):
 * reversed periods swap with their types; L = max of the two MA look-backs;
 * the fast and slow MAs are those TA_MA computes when asked to start at
 * bar L -- for the recursive types the seed sits at L, which is the
 * moving average of the bars from L - lookback(MA) on; macd = fast -
 * slow from L, in double; signal = MA of the macd line; every output
 * starts at L + lookback(signal MA).  pinned: _ref TA_MACDEXT. */
API int orc_macdext(const float* x, float* m, float* s, float* h, int n, int fast, int ft,
                    int slow, int st, int sig, int sgt) {
  /* TA_MACDEXT accepts fast and slow periods from 2, the signal from 1 */
  if (!x || !m || !s || !h || n <= 0 || fast < 2 || slow < 2 || sig < 1) return ORC_BAD;
  if (slow < fast) {
    int tp = fast;
    fast = slow;
    slow = tp;
    tp = ft;
    ft = st;
    st = tp;
  }
  const int lbF = ma_lookback(fast, ft), lbS = ma_lookback(slow, st), lbG = ma_lookback(sig, sgt);
  if (lbF < 0 || lbS < 0 || lbG < 0) return ORC_BAD;
  fill_nan(m, n);
  fill_nan(s, n);
  fill_nan(h, n);
  const int f = lead_nan(x, n);
  const int L = lbF > lbS ? lbF : lbS;
  if (f + L + lbG >= n) return ORC_OK;
  double* in = (double*)malloc(sizeof(double) * (size_t)n);
  double* a = (double*)malloc(sizeof(double) * (size_t)n);
  double* b = (double*)malloc(sizeof(double) * (size_t)n);
  for (int t = 0; t < n; ++t) in[t] = (double)x[t];
  md_ma(in, a, f + L - lbF, n, fast, ft);
  md_ma(in, b, f + L - lbS, n, slow, st);
  for (int t = f + L; t < n; ++t) a[t] -= b[t]; /* macd line from L */
  md_ma(a, b, f + L, n, sig, sgt);
  for (int t = f + L + lbG; t < n; ++t) {
    m[t] = (float)a[t];
    s[t] = (float)b[t];
    h[t] = (float)(a[t] - b[t]);
  }
  free(in);
  free(a);
  free(b);
  return ORC_OK;
}

/* ---------------------------------------------------- price transforms ---
 * MEDPRICE (h+l)/2, TYPPRICE (h+l+c)/3 in double (ta_MEDPRICE.c:70,
 * ta_TYPPRICE.c:68; pinned to _ref TA_MEDPRICE / TA_TYPPRICE); AVGPRICE
 * (o+h+l+c)/4 and WCLPRICE (h+l+2c)/4, the reference's AvgPrice.cu:18 and
 * WclPrice.cu:16 formulas in double (not vendored; pinned by closed forms on
 * dyadic prices, where float arithmetic is exact). */
API int orc_avgprice(const float* o, const float* h, const float* l, const float* c, float* out,
                     int n) {
  if (n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i)
    out[i] = (float)(((double)o[i] + (double)h[i] + (double)l[i] + (double)c[i]) / 4.0);
  return ORC_OK;
}
API int orc_medprice(const float* h, const float* l, float* out, int n) {
  if (n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i) out[i] = (float)(((double)h[i] + (double)l[i]) / 2.0);
  return ORC_OK;
}
API int orc_typprice(const float* h, const float* l, const float* c, float* out, int n) {
  if (n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i)
    out[i] = (float)(((double)h[i] + (double)l[i] + (double)c[i]) / 3.0);
  return ORC_OK;
}
API int orc_wclprice(const float* h, const float* l, const float* c, float* out, int n) {
  if (n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i)
    out[i] = (float)(((double)h[i] + (double)l[i] + (double)c[i] * 2.0) / 4.0);
  return ORC_OK;
}

/* ------------------------------------------------------------ momentum ---
 * y[t] = f(x[t], x[t-p]) from bar p on.  ROC ((x/prev - 1) * 100) and ROCP
 * ((x - prev)/prev), 0 for a zero earlier price: ta_ROC.c:111-124,
 * ta_ROCP.c:113-123 (pinned to _ref TA_ROC / TA_ROCP).  MOMENTUM (x - prev),
 * ROCR (x/prev) and ROCR100 (100 x/prev): the reference's Momentum.cu:16,
 * ROCR.cu:18, ROCR100.cu:19 definitions in double (not vendored; pinned by
 * brute force against numpy and by ROCR100 = 100 ROCR on exact ratios). */
static int mom_family(const float* x, float* out, int n, int p, int op) {
  if (n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p; t < n; ++t) {
    const double cur = x[t], prev = x[t - p];
    double v;
    switch (op) {
      case 0: v = cur - prev; break;
      case 1: v = prev != 0.0 ? (cur / prev - 1.0) * 100.0 : 0.0; break;
      case 2: v = prev != 0.0 ? (cur - prev) / prev : 0.0; break;
      case 3: v = prev != 0.0 ? cur / prev : 0.0; break;
      default: v = prev != 0.0 ? (cur / prev) * 100.0 : 0.0; break;
    }
    out[t] = (float)v;
  }
  return ORC_OK;
}
API int orc_momentum(const float* x, float* out, int n, int p) { return mom_family(x, out, n, p, 0); }
API int orc_roc(const float* x, float* out, int n, int p) { return mom_family(x, out, n, p, 1); }
API int orc_rocp(const float* x, float* out, int n, int p) { return mom_family(x, out, n, p, 2); }
API int orc_rocr(const float* x, float* out, int n, int p) { return mom_family(x, out, n, p, 3); }
API int orc_rocr100(const float* x, float* out, int n, int p) { return mom_family(x, out, n, p, 4); }

/* ---------------------------------------------------- window midpoints ---
 * (max + min)/2 over the p bars ending at t, from bar p-1 on; NaN samples
 * are skipped (fmaxf/fminf, reference MIDPRICE.cu:14-17, MIDPOINT.cu:19-24);
 * a MIDPOINT window that starts inside the row's leading NaN run is NaN
 * (MIDPOINT.cu:25-29, LeadingNaN.cuh:22-28).  MIDPRICE pinned to _ref
 * TA_MIDPRICE on finite data; MIDPOINT by brute force (numpy) and by
 * MIDPOINT(x) = MIDPRICE(x, x) on finite data. */
static int mid_family(const float* hi, const float* lo, float* out, int n, int p, int one) {
  if (n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    const int s = t - p + 1;
    float mx = hi[s], mn = lo[s];
    for (int k = s + 1; k <= t; ++k) {
      mx = fmaxf(mx, hi[k]);
      mn = fminf(mn, lo[k]);
    }
    if (one && !isnan(mx) && isnan(hi[s])) {
      int lead = 1;
      for (int q = s - 1; q >= 0 && lead; --q) lead = isnan(hi[q]);
      if (lead) continue;
    }
    out[t] = (float)(((double)mx + (double)mn) / 2.0);
  }
  return ORC_OK;
}
API int orc_midprice(const float* h, const float* l, float* out, int n, int p) {
  return mid_family(h, l, out, n, p, 0);
}
API int orc_midpoint(const float* x, float* out, int n, int p) {
  return mid_family(x, x, out, n, p, 1);
}

/* --------------------------------------------------- linear regression ---
 * ta_LINEARREG.c:105-131: x is the distance back from t (0..p-1), SumX,
 * SumXSqr and Divisor = SumX^2 - p SumXSqr of those, the window primed
 * oldest first; m = (p SumXY - SumX SumY)/Divisor, b = (SumY - m SumX)/p.
 * LINEARREG = fma(m, p-1, b) (pinned to _ref TA_LINEARREG); SLOPE m,
 * INTERCEPT b, ANGLE atan(m) 180/pi and TSF b + m p (TA-Lib's definitions,
 * ta_LINEARREG.c:90-96; not vendored: pinned by exact lines, where the
 * slope, intercept and forecast are known, and by LINEARREG = INTERCEPT +
 * (p-1) SLOPE).  A one-bar window has slope 0 and intercept x (the
 * reference's zero-denominator rule, LinearRegression.cuh:36-40). */
static int linreg_family(const float* x, float* out, int n, int p, int kind) {
  if (n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  const double sumX = (double)p * (p - 1) * 0.5;
  const double sumXSqr = (double)p * (p - 1) * (2 * p - 1) / 6.0;
  const double divisor = sumX * sumX - p * sumXSqr;
  for (int t = p - 1; t < n; ++t) {
    double sumY = 0.0, sumXY = 0.0;
    for (int i = p - 1; i >= 0; --i) {
      const double v = x[t - i];
      sumY += v;
      sumXY += (double)i * v;
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
    switch (kind) {
      case 0: v = fma(m, (double)(p - 1), b); break;
      case 1: v = m; break;
      case 2: v = b; break;
      case 3: v = atan(m) * (180.0 / 3.14159265358979323846); break;
      default: v = b + m * (double)p; break;
    }
    out[t] = (float)v;
  }
  return ORC_OK;
}
API int orc_linearreg(const float* x, float* out, int n, int p) { return linreg_family(x, out, n, p, 0); }
API int orc_linearreg_slope(const float* x, float* out, int n, int p) { return linreg_family(x, out, n, p, 1); }
API int orc_linearreg_intercept(const float* x, float* out, int n, int p) { return linreg_family(x, out, n, p, 2); }
API int orc_linearreg_angle(const float* x, float* out, int n, int p) { return linreg_family(x, out, n, p, 3); }
API int orc_tsf(const float* x, float* out, int n, int p) { return linreg_family(x, out, n, p, 4); }

/* ------------------------------------------------ element-wise math -----
 * Reference src/api/api_vector_math.cu:28-51: the float function of each
 * element (acosf ... tanhf) and the four float operators.  Here each is the
 * double function of the widened element rounded once to float: the
 * correctly rounded value up to the rare case double's rounding lands on a
 * float halfway point.  For CEIL, FLOOR, SQRT, +, -, * and / double's 53
 * bits are at least 2 x 24 + 2, so the double result rounded to float IS
 * the correctly rounded float result, which IEEE fixes bit for bit
 * (pinned against numpy's float32 operators).  The transcendental ones are
 * pinned against numpy's float64 functions and special values; the device
 * library is a few ulp from them, the tolerance the tests state. */
API int orc_math1(int op, const float* x, float* out, int n) {
  if (!x || !out || n <= 0 || op < 0 || op > 14) return ORC_BAD;
  for (int i = 0; i < n; ++i) {
    const double v = x[i];
    double r;
    switch (op) {
      case 0: r = acos(v); break;
      case 1: r = asin(v); break;
      case 2: r = atan(v); break;
      case 3: r = ceil(v); break;
      case 4: r = cos(v); break;
      case 5: r = cosh(v); break;
      case 6: r = exp(v); break;
      case 7: r = floor(v); break;
      case 8: r = log(v); break;
      case 9: r = log10(v); break;
      case 10: r = sin(v); break;
      case 11: r = sinh(v); break;
      case 12: r = sqrt(v); break;
      case 13: r = tan(v); break;
      default: r = tanh(v); break;
    }
    out[i] = (float)r;
  }
  return ORC_OK;
}
API int orc_math2(int op, const float* a, const float* b, float* out, int n) {
  if (!a || !b || !out || n <= 0 || op < 0 || op > 3) return ORC_BAD;
  for (int i = 0; i < n; ++i) {
    const double u = a[i], v = b[i];
    const double r = op == 0 ? u + v : op == 1 ? u - v : op == 2 ? u * v : u / v;
    out[i] = (float)r;
  }
  return ORC_OK;
}

/* ------------------------------------------------------------ BOP, IBS ---
 * BOP: reference BOP.cu:14-24, in float: d = high - low, 0 when d == 0,
 * else (close - open) / d (pinned by integer bars whose quotients are exact
 * binary fractions and the zero-range rule).  IBS: TA-Lib 0.8.3 ta_IBS
 * (vendored; pinned bit-identical to _ref TA_IBS): in double, range =
 * high - low, (close - low) / range, 0.5 when range <= 0. */
API int orc_bop(const float* o, const float* h, const float* l, const float* c, float* out,
                int n) {
  if (!o || !h || !l || !c || !out || n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i) {
    const float d = h[i] - l[i];
    out[i] = d == 0.0f ? 0.0f : (c[i] - o[i]) / d;
  }
  return ORC_OK;
}
API int orc_ibs(const float* h, const float* l, const float* c, float* out, int n) {
  if (!h || !l || !c || !out || n <= 0) return ORC_BAD;
  for (int i = 0; i < n; ++i) {
    const double range = (double)h[i] - (double)l[i];
    double v = ((double)c[i] - (double)l[i]) / range;
    if (range <= 0.0) v = 0.5;
    out[i] = (float)v;
  }
  return ORC_OK;
}

/* ------------------------------------------------ CHANGE, AVGDEV, MINMAX ---
 * CHANGE: reference Change.cu:24-29 and :37-42: x[t] - x[t-p] in float
 * from bar p, period >= size all NaN (pinned against numpy float32).
 * AVGDEV: reference AVGDEV.cu:18-33, in float: the window summed oldest
 * first, mean = sum / p, |x - mean| summed oldest first, / p (pinned by
 * windows whose float sums are exact, where the value is known in closed
 * form, and a float32 brute force over short windows).  MINMAX: reference
 * MINMAX.cu:17-31: fminf / fmaxf folded from the oldest bar (NaN skipped,
 * an all-NaN window NaN; pinned against numpy nanmin / nanmax). */
API int orc_change(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p; t < n; ++t) out[t] = x[t] - x[t - p];
  return ORC_OK;
}
API int orc_avgdev(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  const float fp = (float)p;
  for (int t = p - 1; t < n; ++t) {
    const float* w = x + t - (p - 1);
    float sum = 0.0f;
    for (int k = 0; k < p; ++k) sum += w[k];
    const float mean = sum / fp;
    float dev = 0.0f;
    for (int k = 0; k < p; ++k) dev += fabsf(w[k] - mean);
    out[t] = dev / fp;
  }
  return ORC_OK;
}
API int orc_minmax(const float* x, float* mn, float* mx, int n, int p) {
  if (!x || !mn || !mx || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(mn, n);
  fill_nan(mx, n);
  for (int t = p - 1; t < n; ++t) {
    const float* w = x + t - (p - 1);
    float a = w[0], b = w[0];
    for (int k = 1; k < p; ++k) {
      a = fminf(a, w[k]);
      b = fmaxf(b, w[k]);
    }
    mn[t] = a;
    mx[t] = b;
  }
  return ORC_OK;
}

/* ------------------------------------ TA-Lib 0.8.3 row functions ---------
 * The reference runs these through its 0.8.3 formula backend
 * (src/compat/Talib08.cu:47-103): per row, the leading bars where any input
 * is NaN are skipped (keeping at least the last bar), the TA_* routine runs
 * over the rest in double from index 0, and its outputs land as float at
 * prefix + outBegIdx; everything else is NaN.  Each routine below restates
 * the vendored ta_*.c loop of the same name (src/compat/talib08_formulas.inc)
 * in its order of operations and is pinned bit-identical to _ref's TA_*
 * (tests/test_oracle_rows.py).  Periods outside TA-Lib's ranges are invalid
 * arguments; a series no longer than the look-back is all NaN. */
typedef struct {
  const double* x[4];
  double* y[4];
  int n;
} RowsIO;

static int rows_prefix(const float* const* in, int nin, int n) {
  int f = 0;
  for (; f < n - 1; ++f) {
    int any = 0;
    for (int j = 0; j < nin; ++j) any |= isnan(in[j][f]);
    if (!any) break;
  }
  return f;
}

static int tafinite(double v) { return v <= 1.7976931348623157e308 && v >= -1.7976931348623157e308; }

/* one routine: writes y[k][t] for t >= its look-back (NaN before); par
 * holds the integer periods, g_row_d the real parameter */
typedef void (*RowRoutine)(const RowsIO* io, const int* par);
static double g_row_d; /* the real parameter of the routine being run */

static void r_cumsum(const RowsIO* io, const int* par) {
  double total = 0.0;
  for (int t = 0; t < io->n; ++t) {
    total += io->x[0][t];
    io->y[0][t] = total;
  }
}
static void r_pvt(const RowsIO* io, const int* par) {
  const double *c = io->x[0], *v = io->x[1];
  double acc = 0.0, pc = c[0];
  for (int t = 0; t < io->n; ++t) {
    const double cc = c[t];
    if (pc != 0.0) acc += (cc - pc) / pc * v[t];
    io->y[0][t] = acc;
    pc = cc;
  }
}
static void r_rma(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p - 1, n = io->n;
  if (n <= lb) return;
  const double* x = io->x[0];
  const double wb = (double)(p - 1) / (double)p, wa = 1.0 - wb;
  double tot = 0.0;
  for (int t = 0; t < p; ++t) tot += x[t];
  double prev = tot / p;
  io->y[0][lb] = prev;
  for (int t = lb + 1; t < n; ++t) {
    prev = fma(wb, prev, wa * x[t]);
    io->y[0][t] = prev;
  }
}
static void r_zlema(const RowsIO* io, const int* par) {
  const int p = par[0], lag = (p - 1) / 2, lb = lag + p - 1, n = io->n;
  const double* x = io->x[0];
  if (n <= lb) return;
  if (p == 1) {
    for (int t = 0; t < n; ++t) io->y[0][t] = x[t];
    return;
  }
  double beta, k;
  ema_coef(p, &beta, &k);
  int tr = 0, today = lag;
  double tmp = 0.0;
  for (int i = 0; i < p; ++i, ++today, ++tr) tmp += 2.0 * x[today] - x[tr];
  double prev = tmp / p;
  io->y[0][lb] = prev;
  for (; today < n; ++today, ++tr) {
    prev = fma(beta, prev, k * (2.0 * x[today] - x[tr]));
    io->y[0][today] = prev;
  }
}
static void r_efi(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p, n = io->n;
  const double *c = io->x[0], *v = io->x[1];
  if (n <= lb) return;
  if (p == 1) {
    for (int t = 1; t < n; ++t) io->y[0][t] = (c[t] - c[t - 1]) * v[t];
    return;
  }
  double beta, k;
  ema_coef(p, &beta, &k);
  double pc = c[0], tmp = 0.0;
  int today = 1;
  for (int i = 0; i < p; ++i, ++today) {
    const double f = (c[today] - pc) * v[today];
    pc = c[today];
    tmp += f;
  }
  double prev = tmp / p;
  io->y[0][lb] = prev;
  for (; today < n; ++today) {
    const double f = (c[today] - pc) * v[today];
    pc = c[today];
    prev = fma(beta, prev, k * f);
    io->y[0][today] = prev;
  }
}
static void r_vwma(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p - 1, n = io->n;
  const double *x = io->x[0], *v = io->x[1];
  if (n <= lb) return;
  if (p == 1) {
    for (int t = 0; t < n; ++t) io->y[0][t] = x[t];
    return;
  }
  double spv = 0.0, sv = 0.0;
  int zc = 0, tr = 0, i = 0;
  for (; i < lb; ++i) {
    spv += x[i] * v[i];
    sv += v[i];
    zc += fabs(v[i]) <= 0.0;
  }
  for (; i < n; ++i, ++tr) {
    spv += x[i] * v[i];
    sv += v[i];
    zc += fabs(v[i]) <= 0.0;
    const double tv = v[tr], tpv = x[tr] * tv;
    double a, b;
    if (zc >= p) {
      a = 0.0;
      b = 0.0;
      spv = 0.0 - tpv;
      sv = 0.0 - tv;
    } else {
      a = spv;
      b = sv;
      spv -= tpv;
      sv -= tv;
    }
    io->y[0][i] = a / (double)p / (b / (double)p);
    zc -= fabs(tv) <= 0.0;
  }
}
static void r_er(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p, n = io->n;
  const double* x = io->x[0];
  if (n <= lb) return;
  double sum = 0.0;
  int run = 0, today = 0, tr = 0;
  for (int i = 0; i < p; ++i) {
    double d = x[today];
    ++today;
    d -= x[today];
    sum += fabs(d);
    run = (run + 1) & ((fabs(d) <= 0.0 ? 0 : 1) - 1);
  }
  double cur = x[today], old = x[tr++];
  double roc = cur - old, trailing = old;
  if (sum <= 0.0 || sum <= roc) {
    io->y[0][today] = 1.0;
  } else {
    double r = fabs(roc / sum);
    io->y[0][today] = r > 1.0 ? 1.0 : r;
  }
  for (++today; today < n; ++today) {
    cur = x[today];
    old = x[tr++];
    roc = cur - old;
    sum -= fabs(trailing - old);
    sum += fabs(cur - x[today - 1]);
    run = (run + 1) & ((fabs(cur - x[today - 1]) <= 0.0 ? 0 : 1) - 1);
    double out;
    if (run >= p) {
      run = p;
      sum = 0.0;
      out = 1.0;
    } else if (sum <= 0.0 || sum <= roc) {
      out = 1.0;
    } else {
      out = fabs(roc / sum);
      if (out > 1.0) out = 1.0;
    }
    io->y[0][today] = out;
    trailing = old;
  }
}
static void r_vhf(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double* x = io->x[0];
  for (int t = p; t < n; ++t) {
    double hi = x[t], lo = hi, prev = hi, chg = 0.0;
    for (int i = p; i >= 0; --i) {
      const double v = x[t - i];
      if (i < p) {
        chg += fabs(v - prev);
        if (v > hi) hi = v;
        if (v < lo) lo = v;
      }
      prev = v;
    }
    io->y[0][t] = chg > 0.0 ? (hi - lo) / chg : 0.0;
  }
}
static void r_dpo(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const int lb = (p - 1) > (p / 2 + 1) ? (p - 1) : (p / 2 + 1);
  const double* x = io->x[0];
  if (n <= lb) return;
  double tot = 0.0;
  int tr = lb - (p - 1), disp = lb - (p / 2 + 1), i = tr;
  for (; i < lb; ++i) tot += x[i];
  for (; i < n; ++i) {
    tot += x[i];
    const double tmp = tot;
    tot -= x[tr++];
    io->y[0][i] = x[disp++] - tmp / (double)p;
  }
}
static void r_chop(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double *h = io->x[0], *l = io->x[1], *c = io->x[2];
  const double lp = log10((double)p);
  for (int t = p; t < n; ++t) {
    double hi = h[t], lo = l[t], str = 0.0, pc = 0.0;
    for (int i = p; i >= 0; --i) {
      if (i < p) {
        const double th = h[t - i], tl = l[t - i];
        const double thi = pc > th ? pc : th, tlo = pc < tl ? pc : tl;
        str += thi - tlo;
        if (th > hi) hi = th;
        if (tl < lo) lo = tl;
      }
      pc = c[t - i];
    }
    io->y[0][t] = (hi - lo <= 0.0 || str <= 0.0) ? 100.0 : 100.0 * (log10(str / (hi - lo)) / lp);
  }
}
/* ADR (high - low) and QSTICK (close - open): the running SMA of a difference */
static void r_sma_diff(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p - 1, n = io->n;
  const double *a = io->x[0], *b = io->x[1]; /* value a - b */
  if (n <= lb) return;
  double tot = 0.0;
  int tr = 0, i = 0;
  if (p > 1)
    for (; i < lb; ++i) tot += a[i] - b[i];
  for (; i < n; ++i, ++tr) {
    tot += a[i] - b[i];
    const double tmp = tot;
    tot -= a[tr] - b[tr];
    io->y[0][i] = tmp / (double)p;
  }
}
static void r_adr(const RowsIO* io, const int* par) { r_sma_diff(io, par); }
static void r_qstick(const RowsIO* io, const int* par) {
  RowsIO s = *io; /* inputs (open, close): close - open */
  s.x[0] = io->x[1];
  s.x[1] = io->x[0];
  r_sma_diff(&s, par);
}
static void r_ao(const RowsIO* io, const int* par) {
  const int f = par[0], s = par[1], n = io->n;
  const int lb = (f > s ? f : s) - 1;
  const double *h = io->x[0], *l = io->x[1];
  if (n <= lb) return;
  double sf = 0.0, ss = 0.0;
  int tf = lb - (f - 1), ts = lb - (s - 1), i = 0;
  for (; i < lb; ++i) {
    const double m = (h[i] + l[i]) / 2.0;
    if (i >= tf) sf += m;
    if (i >= ts) ss += m;
  }
  for (; i < n; ++i, ++tf, ++ts) {
    const double m = (h[i] + l[i]) / 2.0;
    sf += m;
    ss += m;
    const double out = sf / (double)f - ss / (double)s;
    sf -= (h[tf] + l[tf]) / 2.0;
    ss -= (h[ts] + l[ts]) / 2.0;
    io->y[0][i] = out;
  }
}
static void r_donchian(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p - 1, n = io->n;
  const double *h = io->x[0], *l = io->x[1];
  int hiIdx = -1, loIdx = -1, tr = 0;
  double hi = 0.0, lo = 0.0;
  for (int t = lb; t < n; ++t, ++tr) {
    if (hiIdx < tr) {
      hiIdx = tr;
      hi = h[tr];
      for (int i = tr + 1; i <= t; ++i)
        if (h[i] > hi) {
          hiIdx = i;
          hi = h[i];
        }
    } else if (h[t] >= hi) {
      hiIdx = t;
      hi = h[t];
    }
    if (loIdx < tr) {
      loIdx = tr;
      lo = l[tr];
      for (int i = tr + 1; i <= t; ++i)
        if (l[i] < lo) {
          loIdx = i;
          lo = l[i];
        }
    } else if (l[t] <= lo) {
      loIdx = t;
      lo = l[t];
    }
    io->y[0][t] = hi;
    io->y[1][t] = (hi + lo) / 2.0;
    io->y[2][t] = lo;
  }
}
static void r_marketfi(const RowsIO* io, const int* par) {
  const double *h = io->x[0], *l = io->x[1], *v = io->x[2];
  for (int t = 0; t < io->n; ++t) io->y[0][t] = v[t] != 0.0 ? (h[t] - l[t]) / v[t] : 0.0;
}
static void r_wad(const RowsIO* io, const int* par) {
  const double *h = io->x[0], *l = io->x[1], *c = io->x[2];
  double sum = 0.0, pc = c[0];
  for (int t = 0; t < io->n; ++t) {
    const double cc = c[t];
    if (cc > pc) sum += cc - (pc < l[t] ? pc : l[t]);
    else if (cc < pc) sum += cc - (pc > h[t] ? pc : h[t]);
    io->y[0][t] = sum;
    pc = cc;
  }
}
static void r_mcgd(const RowsIO* io, const int* par) {
  const int p = par[0], lb = p - 1, n = io->n;
  const double* x = io->x[0];
  if (n <= lb) return;
  const double per = (double)p;
  double md = x[0];
  for (int t = 1; t < n; ++t) {
    const double r = x[t] / md;
    const double nm = md + (x[t] - md) / (per * (r * r * (r * r)));
    if (tafinite(nm)) md = nm;
    if (t >= lb) io->y[0][t] = md;
  }
}
static void r_rvol(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double* v = io->x[0];
  if (n <= p) return;
  double tot = 0.0;
  int run = 0, tr = 0, i = 0;
  for (; i < p; ++i) {
    tot += v[i];
    run = (run + 1) & ((fabs(v[i]) <= 0.0 ? 0 : 1) - 1);
  }
  for (; i < n; ++i) {
    const double base = tot / (double)p, tv = v[tr++], cv = v[i];
    io->y[0][i] = cv / base;
    tot -= tv;
    tot += cv;
    run = (run + 1) & ((fabs(cv) <= 0.0 ? 0 : 1) - 1);
    if (run >= p) tot = 0.0;
  }
}


/* ta_CMF: money-flow volume (close - low - (high - close)) / (high - low) *
 * volume (0 on a non-positive range) and volume in running p-bar sums; the
 * ratio (0 when the volume sum is not positive), a full window of zero
 * volumes restarts the sums and gives 0. */
static double cmf_mfv(const RowsIO* io, int t) {
  const double h = io->x[0][t], l = io->x[1][t], c = io->x[2][t];
  const double r = h - l;
  return r > 0.0 ? (c - l - (h - c)) / r * io->x[3][t] : 0.0;
}
static void r_cmf(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  if (n < p) return;
  double sm = 0.0, sv = 0.0;
  int run = 0, t = 0;
  for (; t < p; ++t) {
    sm += cmf_mfv(io, t);
    sv += io->x[3][t];
    run = (run + 1) & ((fabs(io->x[3][t]) <= 0.0 ? 0 : 1) - 1);
  }
  io->y[0][p - 1] = sv > 0.0 ? sm / sv : 0.0;
  for (; t < n; ++t) {
    sm -= cmf_mfv(io, t - p);
    sv -= io->x[3][t - p];
    sm += cmf_mfv(io, t);
    sv += io->x[3][t];
    run = (run + 1) & ((fabs(io->x[3][t]) <= 0.0 ? 0 : 1) - 1);
    if (run >= p) {
      run = p;
      sm = 0.0;
      sv = 0.0;
      io->y[0][t] = 0.0;
    } else {
      io->y[0][t] = sv > 0.0 ? sm / sv : 0.0;
    }
  }
}
/* ta_EMV: raw = (mid - previous mid) / (volume / divisor / range) (0 on a
 * zero range or ratio), its running p-bar mean from bar p */
static double emv_raw(const RowsIO* io, int t, double div) {
  const double h = io->x[0][t], l = io->x[1][t];
  const double mid = (h + l) / 2.0, pm = (io->x[0][t - 1] + io->x[1][t - 1]) / 2.0;
  const double range = h - l;
  double raw = 0.0;
  if (range != 0.0) {
    const double box = io->x[2][t] / div / range;
    if (box != 0.0) raw = (mid - pm) / box;
  }
  return raw;
}
static void r_emv(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double div = g_row_d;
  if (n <= p) return;
  double tot = 0.0;
  for (int t = 1; t < p; ++t) tot += emv_raw(io, t, div);
  for (int t = p; t < n; ++t) {
    tot += emv_raw(io, t, div);
    const double tmp = tot;
    tot -= emv_raw(io, t - p + 1, div);
    io->y[0][t] = tmp / (double)p;
  }
}
/* ta_FOSC: 100 (x - forecast) / x with the least-squares line of the p
 * bars before t (x the distance back, as ta_LINEARREG) projected to
 * fma(m, p, b); sums updated incrementally and re-formed every 32 p bars
 * or when the leaving term dwarfs the window's magnitude; 0 when x is 0 */
static void r_fosc(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double* x = io->x[0];
  if (n <= p) return;
  const double sumX = (double)p * (p - 1) * 0.5;
  const double sumXSqr = (double)p * (p - 1) * (2 * p - 1) / 6.0;
  const double div = sumX * sumX - p * sumXSqr;
  double sxy = 0, sy = 0, sa = 0;
  for (int i = p - 1; i >= 0; --i) {
    const double v = x[p - 1 - i];
    sy += v;
    sxy += (double)i * v;
    sa += fabs(v);
  }
  int reseed = 32 * p, tr = 0;
  double trailing = x[tr++];
  for (int t = p; t < n; ++t) {
    if (t > p) {
      const double wt = (double)p * trailing;
      sxy = sxy + sy - wt;
      sy = sy - trailing + x[t - 1];
      sa = sa - fabs(trailing) + fabs(x[t - 1]);
      if (--reseed <= 0 || fabs(wt) > 100.0 * sa) {
        reseed = 32 * p;
        sy = 0;
        sxy = 0;
        sa = 0;
        double w = (double)(p - 1);
        for (int j = t - p; j < t; ++j, w -= 1.0) {
          sy += x[j];
          sxy += w * x[j];
          sa += fabs(x[j]);
        }
      }
      trailing = x[tr++];
    }
    const double m = (p * sxy - sumX * sy) / div;
    const double b = (sy - m * sumX) / (double)p;
    const double c = x[t];
    io->y[0][t] = c != 0.0 ? 100.0 * (c - fma(m, (double)p, b)) / c : 0.0;
  }
}
/* ta_ERI: bull = high - EMA(close), bear = low - EMA(close), the EMA
 * seeded with the mean of the first p closes; p = 1 uses the close */
static void r_eri(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double *h = io->x[0], *l = io->x[1], *c = io->x[2];
  if (n < p) return;
  if (p == 1) {
    for (int t = 0; t < n; ++t) {
      io->y[0][t] = h[t] - c[t];
      io->y[1][t] = l[t] - c[t];
    }
    return;
  }
  double beta, k;
  ema_coef(p, &beta, &k);
  double s = 0.0;
  for (int t = 0; t < p; ++t) s += c[t];
  double ma = s / p;
  io->y[0][p - 1] = h[p - 1] - ma;
  io->y[1][p - 1] = l[p - 1] - ma;
  for (int t = p; t < n; ++t) {
    ma = fma(beta, ma, k * c[t]);
    io->y[0][t] = h[t] - ma;
    io->y[1][t] = l[t] - ma;
  }
}
/* ta_PERCENTRANK: the share (x100) of the p bars before t below x[t] */
static void r_percentrank(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double* x = io->x[0];
  for (int t = p; t < n; ++t) {
    int cnt = 0;
    for (int i = p; i >= 1; --i) cnt += x[t - i] < x[t];
    io->y[0][t] = (double)cnt / (double)p * 100.0;
  }
}
/* Wilder's swing index of bar t against bar t-1 (ta_SI / ta_ASI):
 * inputs (open, high, low, close), limit move L */
static double swing(const RowsIO* io, int t, double lim) {
  const double *o = io->x[0], *h = io->x[1], *l = io->x[2], *c = io->x[3];
  const double pc = c[t - 1], pb = pc - o[t - 1], body = c[t] - o[t];
  double nn = c[t] - pc;
  nn += 0.5 * body;
  nn += 0.25 * pb;
  const double up = fabs(h[t] - pc), dn = fabs(l[t] - pc), rg = fabs(h[t] - l[t]);
  const double k = up > dn ? up : dn;
  double r = (up - 0.5 * dn) > (dn - 0.5 * up) ? (up - 0.5 * dn) : (dn - 0.5 * up);
  r = r > rg ? r : rg;
  r += 0.25 * fabs(pb);
  return r == 0.0 ? 0.0 : 50.0 * (nn / r) * (k / lim);
}
static void r_si(const RowsIO* io, const int* par) {
  for (int t = 1; t < io->n; ++t) io->y[0][t] = swing(io, t, g_row_d);
}
static void r_asi_wilder(const RowsIO* io, const int* par) {
  double sum = 0.0;
  io->y[0][0] = sum;
  for (int t = 1; t < io->n; ++t) {
    sum += swing(io, t, g_row_d);
    io->y[0][t] = sum;
  }
}

/* ta_VORTEX: running sums of the true range (against the previous close),
 * |high - previous low| and |low - previous high| over the p bars ending at
 * t; VI+ and VI- their ratios to the true-range sum (0 when it is not
 * positive); a full window of zero true ranges zeroes the true-range sum. */
static double vx_tr(const RowsIO* io, int i) {
  const double h = io->x[0][i], l = io->x[1][i], pc = io->x[2][i - 1];
  double tr = h - l, d = fabs(pc - h);
  if (d > tr) tr = d;
  d = fabs(pc - l);
  if (d > tr) tr = d;
  return tr;
}
static void r_vortex(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double *h = io->x[0], *l = io->x[1];
  if (n <= p) return;
  double st = 0, sp = 0, sm = 0;
  int run = 0;
  for (int i = 1; i < p; ++i) {
    const double tr = vx_tr(io, i);
    st += tr;
    sp += fabs(h[i] - l[i - 1]);
    sm += fabs(l[i] - h[i - 1]);
    run = (run + 1) & ((tr <= 0.0 ? 0 : 1) - 1);
  }
  for (int t = p, tr0 = 1; t < n; ++t, ++tr0) {
    const double tr = vx_tr(io, t);
    st += tr;
    sp += fabs(h[t] - l[t - 1]);
    sm += fabs(l[t] - h[t - 1]);
    run = (run + 1) & ((tr <= 0.0 ? 0 : 1) - 1);
    if (run >= p) {
      run = p;
      st = 0.0;
    }
    const double ct = st, cp = sp, cm = sm;
    st -= vx_tr(io, tr0);
    sp -= fabs(h[tr0] - l[tr0 - 1]);
    sm -= fabs(l[tr0] - h[tr0 - 1]);
    io->y[0][t] = ct > 0.0 ? cp / ct : 0.0;
    io->y[1][t] = ct > 0.0 ? cm / ct : 0.0;
  }
}
/* ta_CMOU: Chande's momentum oscillator with unsmoothed running sums of the
 * up and down moves over the p changes ending at t: 100 (up - down) /
 * (up + down), 0 when the sum is not positive; a full window of unchanged
 * bars restarts the sums and gives 0. */
static void r_cmou(const RowsIO* io, const int* par) {
  const int p = par[0], n = io->n;
  const double* x = io->x[0];
  if (n <= p) return;
  double us = 0.0, ds = 0.0;
  int run = 0;
  for (int t = 1; t <= p; ++t) {
    const double d = x[t] - x[t - 1], up = (d + fabs(d)) * 0.5;
    us += up;
    ds += up - d;
    run = (run + 1) & ((fabs(d) <= 0.0 ? 0 : 1) - 1);
  }
  double sum = us + ds;
  io->y[0][p] = sum > 0.0 ? 100.0 * (us - ds) / sum : 0.0;
  {
    const double d = x[1] - x[0], up = (d + fabs(d)) * 0.5;
    us -= up;
    ds -= up - d;
  }
  for (int t = p + 1; t < n; ++t) {
    const double d = x[t] - x[t - 1], up = (d + fabs(d)) * 0.5;
    us += up;
    ds += up - d;
    run = (run + 1) & ((fabs(d) <= 0.0 ? 0 : 1) - 1);
    if (run >= p) {
      run = p;
      us = 0.0;
      ds = 0.0;
      io->y[0][t] = 0.0;
    } else {
      sum = us + ds;
      io->y[0][t] = sum > 0.0 ? 100.0 * (us - ds) / sum : 0.0;
    }
    const int o = t - p + 1;
    const double e = x[o] - x[o - 1], ue = (e + fabs(e)) * 0.5;
    us -= ue;
    ds -= ue - e;
  }
}

typedef struct {
  const char* name;
  RowRoutine fn;
  int nin, nout, npar;
  int lo[2], hi[2]; /* valid period ranges */
  double dlo, dhi;  /* the real parameter's range (dlo > dhi: none) */
} RowSpec;

static const RowSpec kRows[] = {
    {"cumsum", r_cumsum, 1, 1, 0, {0, 0}, {0, 0}, 1, 0},
    {"pvt", r_pvt, 2, 1, 0, {0, 0}, {0, 0}, 1, 0},
    {"rma", r_rma, 1, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"zlema", r_zlema, 1, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"efi", r_efi, 2, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"vwma", r_vwma, 2, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"er", r_er, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"vhf", r_vhf, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"dpo", r_dpo, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"chop", r_chop, 3, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"adr", r_adr, 2, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"ao", r_ao, 2, 1, 2, {2, 2}, {100000, 100000}, 1, 0},
    {"qstick", r_qstick, 2, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"donchian", r_donchian, 2, 3, 1, {2, 0}, {100000, 0}, 1, 0},
    {"marketfi", r_marketfi, 3, 1, 0, {0, 0}, {0, 0}, 1, 0},
    {"wad", r_wad, 3, 1, 0, {0, 0}, {0, 0}, 1, 0},
    {"mcgd", r_mcgd, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"rvol", r_rvol, 1, 1, 1, {1, 0}, {100000, 0}, 1, 0},
    {"cmf", r_cmf, 4, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"emv", r_emv, 3, 1, 1, {1, 0}, {100000, 0}, 1.0, 3e37},
    {"fosc", r_fosc, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
    {"eri", r_eri, 3, 2, 1, {1, 0}, {100000, 0}, 1, 0},
    {"percentrank", r_percentrank, 1, 1, 1, {2, 0}, {10000, 0}, 1, 0},
    {"si", r_si, 4, 1, 0, {0, 0}, {0, 0}, 1e-8, 3e37},
    {"asi_wilder", r_asi_wilder, 4, 1, 0, {0, 0}, {0, 0}, 1e-8, 3e37},
    {"vortex", r_vortex, 3, 2, 1, {1, 0}, {100000, 0}, 1, 0},
    {"cmou", r_cmou, 1, 1, 1, {2, 0}, {100000, 0}, 1, 0},
};

/* orc_rows(name, inputs, outputs, n, p0, p1, d0): one row, the backend's
 * leading-NaN rule, float outputs; d0 is the real parameter (EMV's volume
 * divisor, SI's and ASI's limit move). */
API int orc_rows(const char* name, const float* const* in, float* const* out, int n, int p0,
                 int p1, double d0) {
  const RowSpec* s = NULL;
  for (size_t i = 0; i < sizeof(kRows) / sizeof(kRows[0]); ++i)
    if (!strcmp(kRows[i].name, name)) s = &kRows[i];
  if (!s || n <= 0) return ORC_BAD;
  const int par[2] = {p0, p1};
  for (int k = 0; k < s->npar; ++k)
    if (par[k] < s->lo[k] || par[k] > s->hi[k]) return ORC_BAD;
  if (s->dlo <= s->dhi && !(d0 >= s->dlo && d0 <= s->dhi)) return ORC_BAD;
  g_row_d = d0;
  for (int j = 0; j < s->nin; ++j)
    if (!in[j]) return ORC_BAD;
  for (int k = 0; k < s->nout; ++k) {
    if (!out[k]) return ORC_BAD;
    fill_nan(out[k], n);
  }
  const int f = rows_prefix(in, s->nin, n);
  const int m = n - f;
  RowsIO io;
  io.n = m;
  double* buf = (double*)malloc(sizeof(double) * (size_t)m * (size_t)(s->nin + s->nout));
  if (!buf) return ORC_BAD;
  for (int j = 0; j < s->nin; ++j) {
    double* d = buf + (size_t)j * m;
    for (int t = 0; t < m; ++t) d[t] = in[j][f + t];
    io.x[j] = d;
  }
  for (int k = 0; k < s->nout; ++k) {
    double* d = buf + (size_t)(s->nin + k) * m;
    for (int t = 0; t < m; ++t) d[t] = NAN;
    io.y[k] = d;
  }
  s->fn(&io, par);
  for (int k = 0; k < s->nout; ++k)
    for (int t = 0; t < m; ++t) out[k][f + t] = (float)io.y[k][t];
  free(buf);
  return ORC_OK;
}

/* ------------------------------------------------------- ULTOSC / MFI ---
 * ULTOSC: reference ULTOSC.cu:15-39 (not TA-Lib's running sums): float
 * buying pressure and true range summed newest first over each of the three
 * (sorted) windows, 100 (4 a1 + 2 a2 + a3) / 7 in float.  pinned: integer
 * bars where the float sums are exact, so the value is the exact formula;
 * its identity BP + (high' - close) = TR per bar; and numpy brute force.
 * MFI: reference MFI.cu:15-95 with the window summed directly (the value its
 * prefix differences approximate): tp = (h+l+c)/3 in double, flow tp x v
 * positive / negative by the tp change, 0 below a total of 1, else
 * 100 pos / total.  pinned: numpy brute force and monotone closed forms. */
API int orc_ultosc(const float* h, const float* l, const float* c, float* out, int n, int p1,
                   int p2, int p3) {
  if (!h || !l || !c || !out || n <= 0 || p1 <= 0 || p2 <= 0 || p3 <= 0) return ORC_BAD;
  int q;
  if (p1 > p2) { q = p1; p1 = p2; p2 = q; }
  if (p2 > p3) { q = p2; p2 = p3; p3 = q; }
  if (p1 > p2) { q = p1; p1 = p2; p2 = q; }
  fill_nan(out, n);
  for (int t = p3; t < n; ++t) {
    float b1 = 0, r1 = 0, b2 = 0, r2 = 0, b3 = 0, r3 = 0;
    for (int j = 0; j < p3; ++j) {
      const int i = t - j;
      const float pc = i > 0 ? c[i - 1] : c[i];
      const float bp = c[i] - fminf(l[i], pc);
      const float tr = fmaxf(h[i], pc) - fminf(l[i], pc);
      if (j < p1) { b1 += bp; r1 += tr; }
      if (j < p2) { b2 += bp; r2 += tr; }
      b3 += bp;
      r3 += tr;
    }
    const float a1 = r1 == 0.0f ? 0.0f : b1 / r1, a2 = r2 == 0.0f ? 0.0f : b2 / r2;
    const float a3 = r3 == 0.0f ? 0.0f : b3 / r3;
    out[t] = 100.0f * (4.0f * a1 + 2.0f * a2 + a3) / 7.0f;
  }
  return ORC_OK;
}
API int orc_mfi(const float* h, const float* l, const float* c, const float* v, float* out, int n,
                int p) {
  if (!h || !l || !c || !v || !out || n <= 0 || p < 2) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p; t < n; ++t) {
    double pos = 0.0, neg = 0.0;
    for (int k = t - p + 1; k <= t; ++k) {
      const double cur = ((double)h[k] + (double)l[k] + (double)c[k]) / 3.0;
      const double prv = ((double)h[k - 1] + (double)l[k - 1] + (double)c[k - 1]) / 3.0;
      if (!isfinite(cur) || !isfinite(prv)) continue;
      const double f = cur * (double)v[k];
      if (cur > prv) pos += f;
      else if (cur < prv) neg += f;
    }
    const double tot = pos + neg;
    out[t] = tot < 1.0 ? 0.0f : (float)(100.0 * (pos / tot));
  }
  return ORC_OK;
}

/* ------------------------------------------------------ strict rolling ---
 * ROLLING_MEAN / _MAX / _MIN / _STD: reference ClubPack.cu:377-413.  The p
 * bars ending at t, newest first; a non-finite bar anywhere in the window
 * leaves NaN; double sum and sum of squares, mean = sum / p, std =
 * sqrt(max(0, sq / p - mean^2)); float fmaxf / fminf from x[t]; period <= 0
 * or > size is invalid.  pinned: a float64 brute force in the same order,
 * windows of integers (exact), constant windows (std 0). */
API int orc_rolling(int mode, const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0 || p > n || mode < 0 || mode > 3) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    int ok = 1;
    double sum = 0.0, sq = 0.0;
    float mx = x[t], mn = x[t];
    for (int k = 0; k < p; ++k) {
      const float v = x[t - k];
      if (!isfinite(v)) {
        ok = 0;
        break;
      }
      sum += v;
      sq += (double)v * v;
      mx = fmaxf(mx, v);
      mn = fminf(mn, v);
    }
    if (!ok) continue;
    if (mode == 0) out[t] = (float)(sum / p);
    else if (mode == 1) out[t] = mx;
    else if (mode == 2) out[t] = mn;
    else {
      const double mean = sum / p;
      out[t] = (float)sqrt(fmax(0.0, sq / p - mean * mean));
    }
  }
  return ORC_OK;
}

/* ------------------------------------------------------------- pivots ---
 * PIVOT_HIGH / PIVOT_LOW: reference ClubPack.cu:357-374 and :635-657.  Bar
 * b = t - s is a pivot high when x[b] > x[b-k] and x[b] >= x[b+k] for
 * k = 1..s (low: < and <=); its value is reported at t, NaN elsewhere;
 * strength <= 0 or 2 s >= size is invalid.  pinned: numpy brute force and
 * hand-built peaks, plateaus and troughs. */
API int orc_pivot(int high, const float* x, float* out, int n, int s) {
  if (!x || !out || n <= 0 || s <= 0 || 2 * s >= n) return ORC_BAD;
  fill_nan(out, n);
  for (int t = 2 * s; t < n; ++t) {
    const int b = t - s;
    int piv = 1;
    for (int k = 1; k <= s && piv; ++k)
      piv = high ? (x[b] > x[b - k] && x[b] >= x[b + k]) : (x[b] < x[b - k] && x[b] <= x[b + k]);
    if (piv) out[t] = x[b];
  }
  return ORC_OK;
}

/* --------------------------------------------------------------- SKEW ---
 * Reference ClubPack.cu:327-355 (RollingSkew): over the p bars ending at t,
 * newest first, the double mean, central moments m2, m3 (/ p); 0 when
 * m2 <= 0, else m3 / m2^1.5; period <= 0 or > size invalid.  pinned: the
 * numpy population skewness, symmetric windows (0), constant windows (0). */
API int orc_skew(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0 || p > n) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p - 1; t < n; ++t) {
    double mean = 0.0;
    for (int k = 0; k < p; ++k) mean += x[t - k];
    mean /= p;
    double m2 = 0.0, m3 = 0.0;
    for (int k = 0; k < p; ++k) {
      const double d = (double)x[t - k] - mean, d2 = d * d;
      m2 += d2;
      m3 += d2 * d;
    }
    m2 /= p;
    m3 /= p;
    out[t] = m2 <= 0.0 ? 0.0f : (float)(m3 / pow(m2, 1.5));
  }
  return ORC_OK;
}

/* --------------------------------------- TRENDSCORE / MARKET_STRENGTH ---
 * Reference ClubPack.cu:274-306 and :550-590.  TRENDSCORE: sum over k =
 * 1..10 of +1 if x[t] >= x[t-lb-k] else -1, from bar lb + 10 (lb < 0 or
 * size <= lb + 10 invalid).  MARKET_STRENGTH: the p one-bar changes ending
 * at t in double: 100 total / ups when total >= 0, 100 total / downs
 * otherwise, 0 on a zero divisor; p >= size all NaN.  pinned: numpy brute
 * force and monotone closed forms (+10 / -10, +100 / -100). */
API int orc_trendscore(const float* x, float* out, int n, int lb) {
  if (!x || !out || n <= 0 || lb < 0 || n <= lb + 10) return ORC_BAD;
  fill_nan(out, n);
  for (int t = lb + 10; t < n; ++t) {
    int sc = 0;
    for (int k = 1; k <= 10; ++k) sc += x[t] >= x[t - lb - k] ? 1 : -1;
    out[t] = (float)sc;
  }
  return ORC_OK;
}
API int orc_market_strength(const float* x, float* out, int n, int p) {
  if (!x || !out || n <= 0 || p <= 0) return ORC_BAD;
  fill_nan(out, n);
  for (int t = p; t < n; ++t) {
    double tot = 0.0, ups = 0.0, dns = 0.0;
    for (int k = 0; k < p; ++k) {
      const double d = (double)x[t - k] - x[t - k - 1];
      tot += d;
      if (d > 0.0) ups += d;
      else dns += -d;
    }
    if (tot >= 0.0) out[t] = ups > 0.0 ? (float)(100.0 * tot / ups) : 0.0f;
    else out[t] = dns > 0.0 ? (float)(100.0 * tot / dns) : 0.0f;
  }
  return ORC_OK;
}
