// Incremental (streaming) state machines, one per family, shared by the CPU
// single-series streams, the CPU batch path and the GPU batch kernel.  Each
// step is the sequential TA-Lib loop of the full-series definition cut at a
// bar boundary, with the same arithmetic in the same order, so appending bars
// one at a time reproduces the full-series result exactly (reference
// semantics: include/tacuda_stream.h:8-20, include/utils/StreamStates.h).
// Window families keep their last `period` samples in an external ring.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace tca {
namespace st {

#define TCA_HD __host__ __device__ __forceinline__

enum Kind { kSma = 0, kEma, kWma, kRsi, kAtr, kNatr, kObv, kAd, kMacd, kBbands, kKinds };

struct Bar {
  float h, l, c, v;
};

// One plain-old-data record for every family (fields reused per family).
struct State {
  int kind;
  int p1, p2, p3;  // periods (MACD: fast, slow, signal after the swap)
  float up, dn;    // BBANDS multipliers
  int head;        // ring write position
  int started;     // a valid sample has been seen (families that skip NaNs)
  long long n;     // bars appended
  long long k;     // bars since the first valid sample
  float prev;      // previous close
  float pad;
  double a, b, c, d, e;  // running sums / smoothed values
};

TCA_HD float nanf_() { return __builtin_nanf(""); }

TCA_HD void ema_coef(int p, double& beta, double& k) {
  double b = (double)(p - 1) / (double)(p + 1);
  const double kk = 1.0 - b;
  if (b < 0.5) b = 1.0 - kk;
  beta = b;
  k = kk;
}

TCA_HD double true_range(double h, double l, double pc) {
  double g = h - l;
  const double v2 = fabs(pc - h);
  if (v2 > g) g = v2;
  const double v3 = fabs(pc - l);
  if (v3 > g) g = v3;
  return g;
}

TCA_HD bool is_zero(double v) { return (-1e-14 < v) && (v < 1e-14); }

TCA_HD int ring_size(const State& s) {
  return (s.kind == kSma || s.kind == kWma || s.kind == kBbands) ? s.p1 : 0;
}

TCA_HD int lookback(const State& s) {
  switch (s.kind) {
    case kSma: case kEma: case kWma: case kBbands: return s.p1 - 1;
    case kRsi: case kAtr: case kNatr: return s.p1;
    case kMacd: return s.p2 - 1 + s.p3 - 1;
    default: return 0;
  }
}

TCA_HD int outputs(const State& s) { return (s.kind == kMacd || s.kind == kBbands) ? 3 : 1; }

TCA_HD void reset(State& s) {
  s.head = 0;
  s.started = 0;
  s.n = 0;
  s.k = 0;
  s.prev = 0.0f;
  s.a = s.b = s.c = s.d = s.e = 0.0;
}

// Window of the last p samples, oldest first, starting at the ring head.
TCA_HD float ring_at(const float* r, int head, int p, int j) {
  int i = head + j;
  if (i >= p) i -= p;
  return r[i];
}

TCA_HD void step(State& s, float* ring, const Bar& bar, float* out) {
  const float nan = nanf_();
  switch (s.kind) {
    case kSma: case kWma: case kBbands: {
      const int p = s.p1;
      ring[s.head] = bar.c;
      s.head = (s.head + 1 == p) ? 0 : s.head + 1;
      const long long n = ++s.n;
      if (s.kind == kWma && p == 1) {
        out[0] = bar.c;
        return;
      }
      if (n < p) {
        out[0] = nan;
        if (s.kind == kBbands) out[1] = out[2] = nan;
        return;
      }
      if (s.kind == kSma) {  // ta_SMA.c window total, oldest first
        double sum = 0.0;
        for (int j = 0; j < p; ++j) sum += (double)ring_at(ring, s.head, p, j);
        out[0] = (float)(sum / (double)p);
      } else if (s.kind == kWma) {  // ta_WMA.c weights 1..p, oldest first
        const double divider = (double)p * (p + 1) / 2.0;
        double sum = 0.0;
        for (int j = 0; j < p; ++j) sum += (double)ring_at(ring, s.head, p, j) * (j + 1);
        out[0] = (float)(sum / divider);
      } else {  // BBANDS: SMA middle, population deviation of the window
        const double shift = (double)ring_at(ring, s.head, p, 0);
        double sum = 0.0, s1 = 0.0, s2 = 0.0;
        for (int j = 0; j < p; ++j) {
          const double x = (double)ring_at(ring, s.head, p, j);
          sum += x;
          const double dd = x - shift;
          s1 += dd;
          s2 += dd * dd;
        }
        const double m = sum / p;
        const double inv = 1.0 / (double)p;
        const double mean = s1 * inv, e2 = s2 * inv;
        double var = e2 - mean * mean;
        if (var < 1e-12 * e2) var = 0.0;
        if (!(var >= 0.0) && !__builtin_isnan(var)) var = 0.0;
        const double sd = (var != 0.0) ? sqrt(var) : 0.0;
        const double ku = s.up, kd = s.dn;
        out[1] = (float)m;
        if (ku == kd) {
          const double dv = sd * ku;
          out[0] = (float)(m + dv);
          out[2] = (float)(m - dv);
        } else {
          out[0] = (float)fma(sd, ku, m);
          out[2] = (float)(m - sd * kd);
        }
      }
      return;
    }
    case kEma: {  // ta_EMA.c after the leading NaNs
      const int p = s.p1;
      ++s.n;
      if (p == 1) {
        out[0] = bar.c;
        return;
      }
      if (!s.started) {
        if (__builtin_isnan(bar.c)) {
          out[0] = nan;
          return;
        }
        s.started = 1;
      }
      const double x = (double)bar.c;
      const long long k = s.k++;
      if (k < p) {
        s.a += x;
        if (k < p - 1) {
          out[0] = nan;
          return;
        }
        s.e = s.a / p;
      } else {
        double beta, kk;
        ema_coef(p, beta, kk);
        s.e = fma(beta, s.e, kk * x);
      }
      out[0] = (float)s.e;
      return;
    }
    case kRsi: {  // ta_RSI.c after the leading NaNs
      const int p = s.p1;
      ++s.n;
      if (!s.started) {
        out[0] = nan;
        if (__builtin_isnan(bar.c)) return;
        s.started = 1;
        s.prev = bar.c;
        s.k = 0;
        return;
      }
      const long long k = ++s.k;
      const double v = (double)bar.c, d = v - (double)s.prev;
      s.prev = bar.c;
      const double gd = d > 0.0 ? d : 0.0;
      const double inv = 1.0 / (double)p;
      if (k <= p) {
        s.a += gd;
        s.b += gd - d;
        if (k < p) {
          out[0] = nan;
          return;
        }
        s.b *= inv;
        s.a *= inv;
      } else {
        s.b *= (double)(p - 1);
        s.a *= (double)(p - 1);
        s.a += gd;
        s.b += gd - d;
        s.b *= inv;
        s.a *= inv;
      }
      const double sum = s.a + s.b;
      out[0] = (float)(sum > 0.0 ? 100.0 * (s.a / sum) : 50.0);
      return;
    }
    case kAtr: case kNatr: {  // ta_ATR.c; period 1 is TRANGE
      const int p = s.p1;
      const long long n = s.n++;
      const double pc = (double)s.prev;
      s.prev = bar.c;
      if (n == 0) {
        out[0] = nan;
        return;
      }
      const double tr = true_range((double)bar.h, (double)bar.l, pc);
      if (p == 1) {
        out[0] = (float)tr;
        return;
      }
      if (n <= p) {
        s.a += tr;
        if (n < p) {
          out[0] = nan;
          return;
        }
        s.e = s.a / p;
      } else {
        const double wb = (double)(p - 1) / (double)p, wa = 1.0 - wb;
        s.e = fma(wb, s.e, wa * tr);
      }
      if (s.kind == kAtr) {
        out[0] = (float)s.e;
      } else {
        const double cl = (double)bar.c;
        out[0] = is_zero(cl) ? 0.0f : (float)((s.e / cl) * 100.0);
      }
      return;
    }
    case kObv: {
      const long long n = s.n++;
      if (n == 0) {
        s.a += (double)bar.v;
      } else {
        const double pp = (double)bar.c, pv = (double)s.prev;
        double sg = 0.0;
        if (pp > pv) sg = 1.0;
        else if (pp < pv) sg = -1.0;
        s.a += sg * (double)bar.v;
      }
      s.prev = bar.c;
      out[0] = (float)s.a;
      return;
    }
    case kAd: {
      ++s.n;
      const double hi = bar.h, lo = bar.l, cl = bar.c, vo = bar.v;
      double term = 0.0;
      if (__builtin_isfinite(hi) && __builtin_isfinite(lo) && __builtin_isfinite(cl) && __builtin_isfinite(vo)) {
        const double range = hi - lo;
        if (range > 0.0) term = (((cl - lo) - (hi - cl)) / range) * vo;
      }
      s.a += term;
      out[0] = (float)s.a;
      return;
    }
    case kMacd: {  // ta_MACD.c after the leading NaNs: a, b seed sums,
                   // c fast, d slow, e signal
      const int fast = s.p1, slow = s.p2, sig = s.p3;
      ++s.n;
      out[0] = out[1] = out[2] = nan;
      if (!s.started) {
        if (__builtin_isnan(bar.c)) return;
        s.started = 1;
      }
      const double v = (double)bar.c;
      const long long k = s.k++;
      const long long lb = slow - 1 + sig - 1;
      double fb, fk, sb, sk, gb, gk;
      ema_coef(fast, fb, fk);
      ema_coef(slow, sb, sk);
      ema_coef(sig, gb, gk);
      double mv;
      if (k < slow) {
        s.b += v;
        if (k >= slow - fast) s.a += v;
        if (k < slow - 1) return;
        s.d = s.b / slow;
        s.c = s.a / fast;
        mv = s.c - s.d;
        s.e = 0.0;
        s.e += mv;
      } else {
        s.c = fma(fb, s.c, fk * v);
        s.d = fma(sb, s.d, sk * v);
        mv = s.c - s.d;
        if (k <= lb) {
          s.e += mv;
        } else {
          s.e = (sig == 1) ? mv : fma(gb, s.e, gk * mv);
        }
      }
      if (k < lb) return;
      if (k == lb) s.e = s.e / sig;
      out[0] = (float)mv;
      out[1] = (float)s.e;
      out[2] = (float)(mv - s.e);
      return;
    }
    default:
      out[0] = nan;
  }
}

#undef TCA_HD

}  // namespace st
}  // namespace tca
