// Fused multi-indicator launch ("pack"): one wave per segment computes any
// subset of SMA, EMA, WMA, DEMA, TEMA, RSI, MACD, BBANDS, ATR, ADX, STOCH,
// WILLR and CCI from one staging of high/low/close.
//
// Unlike the per-family kernels, no lane runs a sequential warm-up.  Every
// recurrence is an affine map per bar:
//   bars before its first input      identity
//   its first input bar              structural reset to the input
//   warm-up bars                     running sum (exact for float32 inputs)
//   the seed bar                     (sum + input) / period  (TA-Lib's seed)
//   later bars                       the family's TA-Lib update
// so seeds, leading-NaN skips and cascades all come out of the same scan,
// and all level-1 recurrences of all families (EMA, RSI gain/loss, MACD
// fast/slow, ATR, +DM/-DM/TR) share one wave scan and one look-back;
// level 2 (DEMA/TEMA e2, MACD signal, ADX) and level 3 (TEMA e3) follow.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/tacuda_amd.h"
#include "families.hpp"
#include "host.hpp"
#include "series.hpp"
#include "windows.hpp"

namespace tca {

enum PackOut {
  kOSma, kOEma, kOWma, kODema, kOTema, kORsi, kOMacd, kOMacdSig, kOMacdHist, kOBbU, kOBbM,
  kOBbL, kOAtr, kOAdx, kOStK, kOStD, kOWillr, kOCci, kONum
};

struct PackP {
  int smaP, emaP, wmaP, rsiP, mFast, mSlow, mSig, bbP, atrP, adxP, stK, stD, willrP, cciP;
  double bbUp, bbDn;
  unsigned mask;  // bit i = output i requested
};

struct PackCfg {
  int K;  // bars per lane of the recurrence kernels
  int W;  // segments (waves) per workgroup
};

// Tiles of the two recurrence kernels, chosen independently, and whether the
// pass runs as one launch (one = bars per lane of that launch, 0 = three).
struct PackTiles {
  PackCfg close, hlc;
  int one;
};

struct PackIO {
  const float* h;
  const float* l;
  const float* c;
  float* out[kONum];
};


template <class T>
__device__ __forceinline__ void pack_store(const Seg& c, float* lo, const float (&v)[T::K], float* dst) {
  if (dst) store_lane<T>(lo, v, dst, c);
}

// Seeds are TA-Lib's warm-up sequences, evaluated by every lane of the wave
// that holds the seed bar in lockstep (same bar, broadcast reads), so all
// lanes know them without a broadcast.  Bars before a seed contribute 0 to
// the offset scans (bcarry), which keeps the scans free of reset logic.

template <class T>
__device__ __forceinline__ double seed_sum(const float* lds, const float* src, const Seg& c,
                                           int t0, int t1) {
  double s = 0.0;
  for (int t = t0; t <= t1; ++t) s += (double)at<T>(lds, src, c, t);
  return s;
}

__device__ __forceinline__ bool here(const Seg& c, int L) { return L >= c.s0 && L < c.s1; }

// Close recurrences: EMA and its DEMA/TEMA cascade, RSI, MACD.
template <class T>
__device__ __forceinline__ void pack_close(const Seg& c, float* li, float* lo, const PackP& P,
                                           const PackIO& io) {
  constexpr int K = T::K;
  const unsigned m = P.mask;
  const bool wantE = m & ((1u << kOEma) | (1u << kODema) | (1u << kOTema));
  const bool wantR = m & (1u << kORsi);
  const bool wantM = m & ((1u << kOMacd) | (1u << kOMacdSig) | (1u << kOMacdHist));
  const int pe = P.emaP;
  const bool casc = (m & ((1u << kODema) | (1u << kOTema))) && pe > 1;
  const bool wantT = (m & (1u << kOTema)) && pe > 1;
  float* lc = li;
  stage<T>(lc, io.c, c);
  wave_sync();
  float cv[K], ov[K];
  lane_vals<T>(lc, cv);
  const int a = c.s0 + lane_id() * K;
  const float* x = io.c;
  if (pe == 1) {
    if (m & (1u << kOEma)) pack_store<T>(c, lo, cv, io.out[kOEma]);
    if (m & (1u << kODema)) pack_store<T>(c, lo, cv, io.out[kODema]);
    if (m & (1u << kOTema)) pack_store<T>(c, lo, cv, io.out[kOTema]);
  }
  const int f = first_valid(c, x);
  const int pr = P.rsiP, slow = P.mSlow, fast = P.mFast, sig = P.mSig;
  const int Le = sat_add(f, pe - 1), L2 = sat_add(Le, pe - 1), L3 = sat_add(L2, pe - 1);
  const int Lg = sat_add(f, pr);
  const int Lm = sat_add(f, slow - 1), Ls = sat_add(Lm, sig - 1);
  double eb, ek, fb, fk, sb, sk, gb, gk;
  ema_coef(pe, eb, ek);
  ema_coef(fast, fb, fk);
  ema_coef(slow, sb, sk);
  ema_coef(sig, gb, gk);
  const double inv = 1.0 / (double)pr, pm1 = (double)(pr - 1);

  // ---- level-1 seeds (ta_EMA.c, ta_RSI.c, ta_MACD.c warm-ups).  A window
  // inside this segment is summed across the wave (exact for float inputs
  // of ordinary range, where any order gives TA-Lib's sum); one that starts
  // in an earlier segment is summed in TA-Lib's order from the inputs.
  double sE = 0, sG = 0, sL = 0, sF = 0, sS = 0, s2 = 0, sSig = 0, s3 = 0;
  double dlt[K];
  {
    float pc = at<T>(lc, x, c, a - 1);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      dlt[i] = (double)cv[i] - (double)pc;
      pc = cv[i];
    }
  }
  const bool hE = wantE && pe > 1 && here(c, Le), hR = wantR && here(c, Lg);
  const bool hM = wantM && here(c, Lm);
  const bool inE = hE && f >= c.s0, inR = hR && f >= c.s0, inM = hM && f >= c.s0;
  if (inE || inR || inM) {
    double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      const double v = (double)cv[i];
      if (inE && t >= f && t <= Le) part[0] += v;
      if (inR && t > f && t <= Lg) {
        const double gd = dlt[i] > 0.0 ? dlt[i] : 0.0;
        part[1] += gd;
        part[2] += gd - dlt[i];
      }
      if (inM && t >= f && t <= Lm) part[3] += v;
      if (inM && t > Lm - fast && t <= Lm) part[4] += v;
    }
    wave_total<5>(part);
    sE = part[0] / pe;
    sG = part[1] * inv;
    sL = part[2] * inv;
    sS = part[3] / slow;
    sF = part[4] / fast;
  }
  if (hE && !inE) sE = seed_sum<T>(lc, x, c, f, Le) / pe;
  if (hR && !inR) {
    double prev = (double)at<T>(lc, x, c, f), g = 0.0, l = 0.0;
    for (int t = f + 1; t <= Lg; ++t) {
      const double v = (double)at<T>(lc, x, c, t);
      const double d = v - prev;
      prev = v;
      const double gd = d > 0.0 ? d : 0.0;
      g += gd;
      l += gd - d;
    }
    l *= inv;
    g *= inv;
    sG = g;
    sL = l;
  }
  if (hM && !inM) {
    sS = seed_sum<T>(lc, x, c, f, Lm) / slow;
    sF = seed_sum<T>(lc, x, c, Lm - fast + 1, Lm) / fast;
  }

  // ---- level 1: EMA, RSI gain / loss, MACD fast / slow
  constexpr int R1 = 5;
  const double beta1[R1] = {eb, pm1 * inv, pm1 * inv, fb, sb};
  double P1[R1], b1[R1], y1[R1];
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    P1[r] = pow_n(beta1[r], K);
    b1[r] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    const double v = (double)cv[i];
    if (wantE) {
      if (t == Le) b1[0] = sE;
      else if (t > Le) b1[0] = fma(eb, b1[0], ek * v);
    }
    if (wantR) {
      if (t == Lg) {
        b1[1] = sG;
        b1[2] = sL;
      } else if (t > Lg) {
        const double gd = dlt[i] > 0.0 ? dlt[i] : 0.0;
        b1[1] = (b1[1] * pm1 + gd) * inv;
        b1[2] = (b1[2] * pm1 + (gd - dlt[i])) * inv;
      }
    }
    if (wantM) {
      if (t == Lm) {
        b1[3] = sF;
        b1[4] = sS;
      } else if (t > Lm) {
        b1[3] = fma(fb, b1[3], fk * v);
        b1[4] = fma(sb, b1[4], sk * v);
      }
    }
  }
  bcarry<R1>(c, 0, b1, P1, beta1, y1);

  double e1v[K], mv[K];
  {
    float ob[K];
    double ye = y1[0], yg = y1[1], yl = y1[2], yf = y1[3], ys = y1[4];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      const double v = (double)cv[i];
      e1v[i] = mv[i] = qnan_();
      ov[i] = ob[i] = qnanf_();
      if (t == Le) ye = sE;
      else if (t > Le) ye = fma(eb, ye, ek * v);
      if (t >= Le) {
        e1v[i] = ye;
        ov[i] = (float)ye;
      }
      if (t == Lg) {
        yg = sG;
        yl = sL;
      } else if (t > Lg) {
        const double gd = dlt[i] > 0.0 ? dlt[i] : 0.0;
        yl *= pm1;
        yg *= pm1;
        yg += gd;
        yl += gd - dlt[i];
        yl *= inv;
        yg *= inv;
      }
      if (t >= Lg) {
        const double s = yg + yl;
        ob[i] = (float)(s > 0.0 ? 100.0 * (yg / s) : 50.0);
      }
      if (t == Lm) {
        yf = sF;
        ys = sS;
      } else if (t > Lm) {
        yf = fma(fb, yf, fk * v);
        ys = fma(sb, ys, sk * v);
      }
      if (t >= Lm) mv[i] = yf - ys;
    }
    if ((m & (1u << kOEma)) && pe > 1) pack_store<T>(c, lo, ov, io.out[kOEma]);
    if (wantR) pack_store<T>(c, lo, ob, io.out[kORsi]);
  }

  // ---- level 2: DEMA/TEMA e2 (input e1), MACD signal (input MACD line)
  if (!(casc || wantM)) return;
  // ---- level-2 seeds: mean of the first p e1 values (ta_DEMA.c) and of
  // the first `signal` MACD values (ta_MACD.c), from this wave's level-1
  // values when the window starts here, else TA-Lib's warm-up from inputs
  {
    const bool h2 = casc && here(c, L2), hS = wantM && here(c, Ls);
    const bool in2 = h2 && Le >= c.s0, inS = hS && Lm >= c.s0;
    if (in2 || inS) {
      double part[2] = {0.0, 0.0};
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (in2 && t >= Le && t <= L2) part[0] += e1v[i];
        if (inS && t >= Lm && t <= Ls) part[1] += mv[i];
      }
      wave_total<2>(part);
      s2 = part[0] / pe;
      sSig = part[1] / sig;
    }
    if (h2 && !in2) {
      double e1 = seed_sum<T>(lc, x, c, f, Le) / pe;
      double tmp = 0.0;
      tmp += e1;
      for (int t = Le + 1; t <= L2; ++t) {
        e1 = fma(eb, e1, ek * (double)at<T>(lc, x, c, t));
        tmp += e1;
      }
      s2 = tmp / pe;
    }
    if (hS && !inS) {
      double ps = seed_sum<T>(lc, x, c, f, Lm) / slow;
      double pf = seed_sum<T>(lc, x, c, Lm - fast + 1, Lm) / fast;
      double psig = 0.0;
      psig += pf - ps;
      for (int t = Lm + 1; t <= Ls; ++t) {
        const double v = (double)at<T>(lc, x, c, t);
        pf = fma(fb, pf, fk * v);
        ps = fma(sb, ps, sk * v);
        psig += pf - ps;
      }
      sSig = psig / sig;
    }
  }
  constexpr int R2 = 2;
  const double beta2[R2] = {eb, gb};
  double P2[R2], b2[R2] = {0.0, 0.0}, y2[R2];
#pragma unroll
  for (int r = 0; r < R2; ++r) P2[r] = pow_n(beta2[r], K);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    if (casc) {
      if (t == L2) b2[0] = s2;
      else if (t > L2) b2[0] = fma(eb, b2[0], ek * e1v[i]);
    }
    if (wantM) {
      if (t == Ls) b2[1] = sSig;
      else if (t > Ls) b2[1] = (sig == 1) ? mv[i] : fma(gb, b2[1], gk * mv[i]);
    }
  }
  bcarry<R2>(c, 1, b2, P2, beta2, y2);
  double e2v[K];
  {
    float om[K], os[K], oh[K];
    double ye = y2[0], yg = y2[1];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      e2v[i] = qnan_();
      ov[i] = om[i] = os[i] = oh[i] = qnanf_();
      if (t == L2) ye = s2;
      else if (t > L2) ye = fma(eb, ye, ek * e1v[i]);
      if (t >= L2) {
        e2v[i] = ye;
        ov[i] = (float)(2.0 * e1v[i] - ye);
      }
      if (t == Ls) yg = sSig;
      else if (t > Ls) yg = (sig == 1) ? mv[i] : fma(gb, yg, gk * mv[i]);
      if (t >= Ls) {
        om[i] = (float)mv[i];
        os[i] = (float)yg;
        oh[i] = (float)(mv[i] - yg);
      }
    }
    if ((m & (1u << kODema)) && pe > 1) pack_store<T>(c, lo, ov, io.out[kODema]);
    if (wantM) {
      pack_store<T>(c, lo, om, io.out[kOMacd]);
      pack_store<T>(c, lo, os, io.out[kOMacdSig]);
      pack_store<T>(c, lo, oh, io.out[kOMacdHist]);
    }
  }

  // ---- level 3: TEMA e3 (input e2)
  if (!wantT) return;
  // ---- level-3 seed: mean of the first p e2 values (ta_TEMA.c)
  {
    const bool h3 = here(c, L3), in3 = h3 && L2 >= c.s0;
    if (in3) {
      double part[1] = {0.0};
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (t >= L2 && t <= L3) part[0] += e2v[i];
      }
      wave_total<1>(part);
      s3 = part[0] / pe;
    }
    if (h3 && !in3) {
      double e1 = seed_sum<T>(lc, x, c, f, Le) / pe;
      double tmp = 0.0;
      tmp += e1;
      for (int t = Le + 1; t <= L2; ++t) {
        e1 = fma(eb, e1, ek * (double)at<T>(lc, x, c, t));
        tmp += e1;
      }
      double e2 = tmp / pe;
      tmp = 0.0;
      tmp += e2;
      for (int t = L2 + 1; t <= L3; ++t) {
        e1 = fma(eb, e1, ek * (double)at<T>(lc, x, c, t));
        e2 = fma(eb, e2, ek * e1);
        tmp += e2;
      }
      s3 = tmp / pe;
    }
  }
  const double beta3[1] = {eb};
  double P3[1] = {pow_n(eb, K)}, b3[1] = {0.0}, y3[1];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    if (t == L3) b3[0] = s3;
    else if (t > L3) b3[0] = fma(eb, b3[0], ek * e2v[i]);
  }
  bcarry<1>(c, 2, b3, P3, beta3, y3);
  {
    double ye = y3[0];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      ov[i] = qnanf_();
      if (t == L3) ye = s3;
      else if (t > L3) ye = fma(eb, ye, ek * e2v[i]);
      if (t >= L3) ov[i] = (float)(ye + (3.0 * e1v[i] - 3.0 * e2v[i]));
    }
    pack_store<T>(c, lo, ov, io.out[kOTema]);
  }
}

// High/low/close recurrences: ATR (ta_ATR.c) and ADX (Wilder sums of +DM,
// -DM and TR, DX, Wilder mean of DX; the reference ADX.cu/DX.cu semantics).
template <class T>
__device__ __forceinline__ void pack_hlc(const Seg& c, float* li, float* lo, const PackP& P,
                                         const PackIO& io) {
  constexpr int K = T::K;
  const unsigned m = P.mask;
  const bool wantA = m & (1u << kOAtr);
  const bool wantX = m & (1u << kOAdx);
  float* lh = li;
  float* ll = li + T::PLEN;
  float* lc = li + 2 * T::PLEN;
  stage<T>(lh, io.h, c);
  stage<T>(ll, io.l, c);
  stage<T>(lc, io.c, c);
  wave_sync();
  float hv[K], lv[K], cv[K], ov[K];
  lane_vals<T>(lh, hv);
  lane_vals<T>(ll, lv);
  lane_vals<T>(lc, cv);
  const int a = c.s0 + lane_id() * K;
  const int pa = P.atrP, px = P.adxP;
  const int La = pa, Lx = px - 1, Lax = 2 * px - 1;
  const double wb = (double)(pa - 1) / (double)pa, wa = 1.0 - wb;
  const double pp = (double)px, am = 1.0 - 1.0 / pp, bx = (pp - 1.0) / pp;
  auto tr_at = [&](int t) {
    return true_range(at<T>(lh, io.h, c, t), at<T>(ll, io.l, c, t), at<T>(lc, io.c, c, t - 1));
  };
  auto dm_at = [&](int t, double& pd, double& md) {
    dm_pair(at<T>(lh, io.h, c, t - 1), at<T>(lh, io.h, c, t), at<T>(ll, io.l, c, t - 1),
            at<T>(ll, io.l, c, t), pd, md);
  };

  // ---- level 1: ATR, +DM, -DM, TR sums
  double tr[K], pdm[K], mdm[K];
  {
    double ph = at<T>(lh, io.h, c, a - 1), pl = at<T>(ll, io.l, c, a - 1),
           pc = at<T>(lc, io.c, c, a - 1);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      tr[i] = true_range(hv[i], lv[i], pc);
      dm_pair(ph, hv[i], pl, lv[i], pdm[i], mdm[i]);
      ph = hv[i];
      pl = lv[i];
      pc = cv[i];
    }
  }
  // ---- level-1 seeds: mean TR over bars 1..p (ta_ATR.c) and the Wilder
  // sums of +DM, -DM, TR over bars 1..p-1
  double sA = 0, sP = 0, sM = 0, sT = 0, sX = 0;
  const bool hA = wantA && here(c, La), hX = wantX && here(c, Lx);
  const bool inA = hA && c.s0 <= 1, inX = hX && c.s0 <= 1;
  if (inA || inX) {
    double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      if (inA && t >= 1 && t <= La) part[0] += tr[i];
      if (inX && t >= 1 && t <= Lx) {
        part[1] += pdm[i];
        part[2] += mdm[i];
        part[3] += tr[i];
      }
    }
    wave_total<4>(part);
    sA = part[0] / pa;
    sP = part[1];
    sM = part[2];
    sT = part[3];
  }
  if (hA && !inA) {
    double tot = 0.0;
    for (int t = 1; t <= pa; ++t) tot += tr_at(t);
    sA = tot / pa;
  }
  if (hX && !inX) {
    double sp = 0.0, sm = 0.0, st = 0.0;
    for (int t = 1; t <= px - 1; ++t) {
      double pd, md;
      dm_at(t, pd, md);
      sp += pd;
      sm += md;
      st += tr_at(t);
    }
    sP = sp;
    sM = sm;
    sT = st;
  }
  constexpr int R1 = 4;
  const double beta1[R1] = {wb, am, am, am};
  double P1[R1], b1[R1], y1[R1];
#pragma unroll
  for (int r = 0; r < R1; ++r) {
    P1[r] = pow_n(beta1[r], K);
    b1[r] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    if (wantA) {
      if (t == La) b1[0] = sA;
      else if (t > La) b1[0] = fma(wb, b1[0], wa * tr[i]);
    }
    if (wantX) {
      if (t == Lx) {
        b1[1] = sP;
        b1[2] = sM;
        b1[3] = sT;
      } else if (t > Lx) {
        b1[1] = b1[1] - b1[1] / pp + pdm[i];
        b1[2] = b1[2] - b1[2] / pp + mdm[i];
        b1[3] = b1[3] - b1[3] / pp + tr[i];
      }
    }
  }
  bcarry<R1>(c, 0, b1, P1, beta1, y1);
  double dxv[K];
  {
    double ya = y1[0], yp = y1[1], yq = y1[2], yt = y1[3];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      ov[i] = qnanf_();
      dxv[i] = qnan_();
      if (t == La) ya = sA;
      else if (t > La) ya = fma(wb, ya, wa * tr[i]);
      if (t >= La) ov[i] = (float)(pa == 1 ? tr[i] : ya);
      if (t == Lx) {
        yp = sP;
        yq = sM;
        yt = sT;
      } else if (t > Lx) {
        yp = yp - yp / pp + pdm[i];
        yq = yq - yq / pp + mdm[i];
        yt = yt - yt / pp + tr[i];
        double pdi, mdi, dx;
        dir_outputs(yp, yq, yt, pdi, mdi, dx);
        dxv[i] = dx;
      }
    }
    if (wantA) pack_store<T>(c, lo, ov, io.out[kOAtr]);
  }

  // ---- level 2: ADX (input DX), seeded by the mean of DX over bars p..2p-1
  if (!wantX) return;
  {
    const bool hAx = here(c, Lax), inAx = hAx && px >= c.s0;
    if (inAx) {
      double part[1] = {0.0};
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (t >= px && t <= Lax) part[0] += dxv[i];
      }
      wave_total<1>(part);
      sX = part[0] / pp;
    }
    if (hAx && !inAx) {
      double sp = 0.0, sm = 0.0, st = 0.0, sum = 0.0;
      for (int t = 1; t <= px - 1; ++t) {
        double pd, md;
        dm_at(t, pd, md);
        sp += pd;
        sm += md;
        st += tr_at(t);
      }
      for (int t = px; t <= Lax; ++t) {
        double pd, md;
        dm_at(t, pd, md);
        sp = sp - sp / pp + pd;
        sm = sm - sm / pp + md;
        st = st - st / pp + tr_at(t);
        double pdi, mdi, dx;
        dir_outputs(sp, sm, st, pdi, mdi, dx);
        sum += dx;
      }
      sX = sum / pp;
    }
  }
  const double beta2[1] = {bx};
  double P2[1] = {pow_n(bx, K)}, b2[1] = {0.0}, y2[1];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    if (t == Lax) b2[0] = sX;
    else if (t > Lax) b2[0] = ((b2[0] * (pp - 1.0)) + dxv[i]) / pp;
  }
  bcarry<1>(c, 1, b2, P2, beta2, y2);
  {
    double yx = y2[0];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      ov[i] = qnanf_();
      if (t == Lax) yx = sX;
      else if (t > Lax) yx = ((yx * (pp - 1.0)) + dxv[i]) / pp;
      if (t >= Lax) ov[i] = (float)yx;
    }
    pack_store<T>(c, lo, ov, io.out[kOAdx]);
  }
}

// ================================================================ windows ==
// SMA, WMA, BBANDS, STOCH, WILLR and CCI over a column tile (one bar per lane
// per 64-bar chunk), each window evaluated directly in the oracle's order.
template <class C>
__device__ __forceinline__ void pack_windows(const Seg& c, ColLds<C::E>& s, const PackP& P,
                                             const PackIO& io) {
  const unsigned m = P.mask;
  const int lane = lane_id();
  const int base = c.s0 - C::H;  // bar of e = 0
  const bool needHL = m & ((1u << kOStK) | (1u << kOStD) | (1u << kOWillr) | (1u << kOCci));
  if (needHL) {
    col_stage<C>(s.h, io.h, c);
    col_stage<C>(s.l, io.l, c);
  }
  col_stage<C>(s.c, io.c, c);
  col_sync();

  // SMA and the BBANDS middle band of one period are the same window mean:
  // the BBANDS loop then writes both
  const bool bbAll = m & ((1u << kOBbU) | (1u << kOBbM) | (1u << kOBbL));
  const bool smaInBb = (m & (1u << kOSma)) && bbAll && P.smaP == P.bbP;
  if ((m & (1u << kOSma)) && !smaInBb) {
    const int p = P.smaP;
    for (int j = C::NH; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      double sum = 0.0;
      for (int q = e - p + 1; q <= e; ++q) sum += (double)lds_f(s.c, q);
      const float v = (base + e >= p - 1) ? (float)(sum / (double)p) : qnanf_();
      col_store<C>(io.out[kOSma], c, j, v);
    }
  }
  if (m & (1u << kOWma)) {
    const int p = P.wmaP;
    const double divider = (double)p * (p + 1) / 2.0;
    for (int j = C::NH; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      float v;
      if (p == 1) {
        v = lds_f(s.c, e);
      } else {
        double sum = 0.0;
        double w = 1.0;
        for (int q = e - p + 1; q <= e; ++q, w += 1.0) sum += (double)lds_f(s.c, q) * w;
        v = (float)(sum / divider);
      }
      col_store<C>(io.out[kOWma], c, j, (base + e >= p - 1) ? v : qnanf_());
    }
  }
  if (bbAll) {
    const int p = P.bbP;
    const double inv = 1.0 / (double)p;
    for (int j = C::NH; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      const double shift = (double)lds_f(s.c, e - p + 1);
      double sum = 0.0, s1 = 0.0, s2 = 0.0;
      for (int q = e - p + 1; q <= e; ++q) {
        const double x = (double)lds_f(s.c, q);
        sum += x;
        const double d = x - shift;
        s1 += d;
        s2 += d * d;
      }
      const double mean = sum / p;
      const double var = col_var_finish(s1, s2, inv);
      const double sd = (var != 0.0) ? sqrt(var) : 0.0;
      float up, lo;
      if (P.bbUp == P.bbDn) {
        const double d = sd * P.bbUp;
        up = (float)(mean + d);
        lo = (float)(mean - d);
      } else {
        up = (float)fma(sd, P.bbUp, mean);
        lo = (float)(mean - sd * P.bbDn);
      }
      const bool ok = base + e >= p - 1;
      if (smaInBb) col_store<C>(io.out[kOSma], c, j, ok ? (float)(sum / (double)p) : qnanf_());
      col_store<C>(io.out[kOBbU], c, j, ok ? up : qnanf_());
      col_store<C>(io.out[kOBbM], c, j, ok ? (float)mean : qnanf_());
      col_store<C>(io.out[kOBbL], c, j, ok ? lo : qnanf_());
    }
  }
  const bool wantSt = m & ((1u << kOStK) | (1u << kOStD));
  const bool wantW = m & (1u << kOWillr);
  if (wantSt || wantW) {
    // WILLR first when it shares the extrema table with STOCH, whose
    // smoothing then reuses the table's LDS
    const int kp = wantSt ? P.stK : P.willrP;
    int k = col_extrema_table<C>(s, kp);
    auto willr = [&]() {
      const int p = P.willrP;
      for (int j = C::NH; j < C::NC; ++j) {
        const int e = j * kWave + lane;
        float hh, ll;
        col_hl<C>(s, e, p, k, hh, ll);
        const double den = (double)hh - (double)ll;
        const double v =
            den == 0.0 ? 0.0 : (((double)hh - (double)lds_f(s.c, e)) / den) * -100.0;
        const float o = isnan(v) ? qnanf_() : (float)fmin(0.0, fmax(-100.0, v));
        col_store<C>(io.out[kOWillr], c, j, (base + e >= p - 1) ? o : qnanf_());
      }
    };
    if (wantW && P.willrP == kp) willr();
    if (wantSt) {
      const int dp = P.stD;
      const int lb = kp - 1 + 2 * (dp - 1);
      // fast %K from bar s0 - 2(d-1), slow %K from s0 - (d-1)
      const int eK = C::H - 2 * (dp - 1), eS = C::H - (dp - 1);
      for (int j = eK / kWave; j < C::NC; ++j) {
        const int e = j * kWave + lane;
        float hh, ll;
        col_hl<C>(s, e < eK ? eK : e, kp, k, hh, ll);
        const double range = (double)hh - (double)ll;
        const double cl = (double)lds_f(s.c, e);
        const double fk = (fabs(range) <= 1e-14 * (fabs((double)hh) + fabs((double)ll)))
                              ? 0.0
                              : (cl - (double)ll) / range * 100.0;
        if (e >= eK) lds_wd(s.d0, e, fk);
      }
      col_sync();
      double* d1 = s.d1();  // the extrema table is no longer read
      for (int j = eS / kWave; j < C::NC; ++j) {
        const int e = j * kWave + lane;
        double sum = 0.0;
        if (e >= eS)
          for (int q = e - dp + 1; q <= e; ++q) sum += lds_d(s.d0, q);
        if (e >= eS) lds_wd(d1, e, sum / dp);
      }
      col_sync();
      for (int j = C::NH; j < C::NC; ++j) {
        const int e = j * kWave + lane;
        double sum = 0.0;
        for (int q = e - dp + 1; q <= e; ++q) sum += lds_d(d1, q);
        const bool ok = base + e >= lb;
        col_store<C>(io.out[kOStK], c, j, ok ? (float)lds_d(d1, e) : qnanf_());
        col_store<C>(io.out[kOStD], c, j, ok ? (float)(sum / dp) : qnanf_());
      }
      col_sync();
      if (wantW && P.willrP != kp) {
        k = col_extrema_table<C>(s, P.willrP);
        willr();
      }
    } else {
      willr();
    }
  }
  if (m & (1u << kOCci)) {
    const int p = P.cciP;
    col_sync();
    const int e0 = C::H - (p - 1);
    for (int j = e0 / kWave; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      const double tp =
          ((double)lds_f(s.h, e) + (double)lds_f(s.l, e) + (double)lds_f(s.c, e)) / 3.0;
      lds_wd(s.d0, e, tp);
    }
    col_sync();
    for (int j = C::NH; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      double sum = 0.0;
      for (int q = e; q > e - p; --q) sum += lds_d(s.d0, q);
      const double sma = sum / p;
      double dev = 0.0;
      for (int q = e; q > e - p; --q) dev += fabs(lds_d(s.d0, q) - sma);
      const double md = dev / p;
      const double num = lds_d(s.d0, e) - sma;
      const float v = (num != 0.0 && md != 0.0) ? (float)(num / (0.015 * md)) : 0.0f;
      col_store<C>(io.out[kOCci], c, j, (base + e >= p - 1) ? v : qnanf_());
    }
  }
}

template <int H, int W>
__global__ __launch_bounds__(W* kWave) void k_pack_win(Geometry g, PackP prm, PackIO io) {
  using C = ColTile<H>;
  __shared__ ColLds<C::E> s_lds[W];
  const int w = threadIdx.x / kWave;
  Seg c;
  c.tree = false;
  c.epoch = 0;
  if (resolve(g, C::SEG, blockIdx.x * W + w, c)) pack_windows<C>(c, s_lds[w], prm, io);
}

template <int K, int H, int W, int PART>
__global__ __launch_bounds__(W* kWave) void k_pack(Geometry g, Ws ws, int useLB, PackP prm,
                                                    PackIO io) {
  using T = Tile<K, H>;
  __shared__ __align__(16) float s_in[W][3 * T::PLEN];
  __shared__ __align__(16) float s_out[W][T::POUT];
  const int w = threadIdx.x / kWave;
  Seg c;
  c.tree = useLB != 0;
  c.epoch = useLB ? read_epoch(ws) : 0u;
  if (resolve(g, T::SEG, blockIdx.x * W + w, c)) {
    c.ws = ws;
    if (PART == 1)
      pack_close<T>(c, s_in[w], s_out[w], prm, io);
    else
      pack_hlc<T>(c, s_in[w], s_out[w], prm, io);
  }
  if (useLB) {
    __syncthreads();
    if (threadIdx.x == 0) finish_block(ws, gridDim.x);
  }
}

constexpr unsigned kWinMask = (1u << kOSma) | (1u << kOWma) | (1u << kOBbU) | (1u << kOBbM) |
                              (1u << kOBbL) | (1u << kOStK) | (1u << kOStD) | (1u << kOWillr) |
                              (1u << kOCci);

constexpr unsigned kHlcRecMask = (1u << kOAtr) | (1u << kOAdx);
constexpr unsigned kCloseRecMask = ~(kWinMask | kHlcRecMask) & ((1u << kONum) - 1);

// Longest window look-back the fused launch stages (bars before the segment).
constexpr int kPackMaxWindow = 1024;


// Halo of the window kernel: the longest look-back of the requested window
// outputs, rounded up to whole 64-bar chunks.
static int window_halo(const PackP& P) {
  const unsigned m = P.mask;
  int need = 0;
  auto upd = [&](unsigned bits, int v) {
    if ((m & bits) && v > need) need = v;
  };
  upd(1u << kOSma, P.smaP - 1);
  upd(1u << kOWma, P.wmaP - 1);
  upd((1u << kOBbU) | (1u << kOBbM) | (1u << kOBbL), P.bbP - 1);
  upd((1u << kOStK) | (1u << kOStD), P.stK - 1 + 2 * (P.stD - 1));
  upd(1u << kOWillr, P.willrP - 1);
  upd(1u << kOCci, P.cciP - 1);
  return need;
}

template <int H, int W>
static ctStatus_t launch_win_h(int len, int rows, hipStream_t s, const PackP& prm,
                               const PackIO& io) {
  Geometry g = dense_geometry(len, rows, ColTile<H>::SEG);
  int maxSegs = 0;
  if (t_ragged && !ragged_geometry(t_ragged, ColTile<H>::SEG, &g, &maxSegs))
    return fail(CT_STATUS_UNSUPPORTED, "pack: no ragged table for the window tile");
  if (g.totalSegs <= 0) return CT_STATUS_SUCCESS;
  const int blocks = div_up(g.totalSegs, W);
  hipLaunchKernelGGL((k_pack_win<H, W>), dim3(blocks), dim3(W * kWave), 0, s, g, prm, io);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

static ctStatus_t launch_windows(int len, int rows, hipStream_t s, const PackP& prm,
                                 const PackIO& io) {
  const int need = window_halo(prm);
  if (need <= 64) return launch_win_h<64, 4>(len, rows, s, prm, io);
  if (need <= 192) return launch_win_h<192, 2>(len, rows, s, prm, io);
  if (need <= 512) return launch_win_h<512, 1>(len, rows, s, prm, io);
  return launch_win_h<kPackMaxWindow, 1>(len, rows, s, prm, io);
}

// One recurrence kernel (PART 1 close, 2 high/low/close) of one tile
// configuration: K bars per lane (the scan chunk), W waves (segments) per
// workgroup.
template <int K, int W, int PART>
static ctStatus_t launch_rec(int len, int rows, hipStream_t s, const PackP& prm,
                             const PackIO& io) {
  constexpr int SEG = K * kWave;
  Geometry g = dense_geometry(len, rows, SEG);
  int maxSegs = g.segsPerRow;
  if (t_ragged && !ragged_geometry(t_ragged, SEG, &g, &maxSegs))
    return fail(CT_STATUS_UNSUPPORTED, "pack: no ragged table for the recurrence tile");
  if (g.totalSegs <= 0) return CT_STATUS_SUCCESS;
  const int blocks = div_up(g.totalSegs, W);
  const bool useLB = maxSegs > 1;
  // 3 levels x (segments + groups) slots of a flag and up to 5 offsets
  Ws ws{nullptr, nullptr, 1 + 5};
  if (useLB) {
    const size_t bytes = (size_t)3 * ((size_t)g.totalSegs + g.totalSegs / kGroup) * (1 + 5) * 8;
    ctStatus_t st = workspace_for(s, bytes, &ws);
    if (st != CT_STATUS_SUCCESS) return st;
  }
  hipLaunchKernelGGL((k_pack<K, K, W, PART>), dim3(blocks), dim3(W * kWave), 0, s, g, ws,
                     useLB ? 1 : 0, prm, io);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// Tile configurations the autotuner chooses from.
const PackCfg kPackCfgs[] = {{4, 4}, {4, 2}, {4, 1}, {4, 8}, {8, 2}, {8, 1}, {8, 4}};

template <int PART>
static ctStatus_t launch_part(PackCfg cfg, int len, int rows, hipStream_t s, const PackP& p,
                              const PackIO& io) {
#define TCA_CFG(KK, WW) \
  if (cfg.K == KK && cfg.W == WW) return launch_rec<KK, WW, PART>(len, rows, s, p, io);
  TCA_CFG(4, 4)
  TCA_CFG(4, 2)
  TCA_CFG(4, 1)
  TCA_CFG(4, 8)
  TCA_CFG(8, 2)
  TCA_CFG(8, 1)
  TCA_CFG(8, 4)
#undef TCA_CFG
  return fail(CT_STATUS_INVALID_ARGUMENT, "pack: unknown tile configuration %dx%d", cfg.K, cfg.W);
}

// ------------------------------------------------------- one launch --
// The window kernel's and both recurrence kernels' workgroups in one grid,
// kFuseW waves each.  Rows of several segments put the recurrence blocks
// first, [close | high/low/close | windows]: their look-back chains set the
// pass's length, and their waves waiting on look-back leave the ALUs to the
// window waves.  Rows of one segment (no look-back) put the heavier window
// blocks first, [windows | close | high/low/close], which shortens the
// tail.  Each
// recurrence part scans through its own workspace part; tickets within a
// part follow block order, so every predecessor a wave waits for was
// dispatched before it.  Window look-backs up to 64 bars.
constexpr int kFuseW = 4;

template <int K>
struct FuseLds {
  using C = ColTile<64>;
  using T = Tile<K, K>;
  union {
    ColLds<C::E> win[kFuseW];
    struct {
      float in[kFuseW][3 * T::PLEN];
      float out[kFuseW][T::POUT];
    } rec;
  };
};

template <int K>
__global__ __launch_bounds__(kFuseW* kWave) void k_pack_one(Geometry gw, Geometry gr, Ws wsC,
                                                            Ws wsH, int useLB, int nbW, int nbC,
                                                            int nbH, PackP prm, PackIO io) {
  using C = ColTile<64>;
  using T = Tile<K, K>;
  __shared__ __align__(16) FuseLds<K> sm;
  const int w = threadIdx.x / kWave;
  int b = blockIdx.x;
  // windows-first grids rotate the index so the same tests below apply
  if (!useLB) b = b >= nbW ? b - nbW : b + nbC + nbH;
  if (b >= nbC + nbH) {
    b -= nbC + nbH;
    Seg c;
    c.tree = false;
    c.epoch = 0;
    if (resolve(gw, C::SEG, b * kFuseW + w, c)) pack_windows<C>(c, sm.win[w], prm, io);
    return;
  }
  const bool cl = b < nbC;
  if (!cl) b -= nbC;
  const Ws ws = cl ? wsC : wsH;
  Seg c;
  c.tree = useLB != 0;
  c.epoch = useLB ? read_epoch(ws) : 0u;
  if (resolve(gr, T::SEG, b * kFuseW + w, c)) {
    c.ws = ws;
    if (cl)
      pack_close<T>(c, sm.rec.in[w], sm.rec.out[w], prm, io);
    else
      pack_hlc<T>(c, sm.rec.in[w], sm.rec.out[w], prm, io);
  }
  if (useLB) {
    __syncthreads();
    if (threadIdx.x == 0) finish_block(ws, cl ? nbC : nbH, b);
  }
}

template <int K>
static ctStatus_t launch_one(int len, int rows, hipStream_t s, const PackP& prm,
                             const PackIO& io) {
  using C = ColTile<64>;
  constexpr int SEG = K * kWave;
  const bool wW = prm.mask & kWinMask, wC = prm.mask & kCloseRecMask, wH = prm.mask & kHlcRecMask;
  Geometry gw = dense_geometry(len, rows, C::SEG);
  Geometry gr = dense_geometry(len, rows, SEG);
  int maxW = 0, maxSegs = gr.segsPerRow;
  if (t_ragged && (!ragged_geometry(t_ragged, C::SEG, &gw, &maxW) ||
                   !ragged_geometry(t_ragged, SEG, &gr, &maxSegs)))
    return fail(CT_STATUS_UNSUPPORTED, "pack: no ragged table for the tile");
  const int nbW = wW ? div_up(gw.totalSegs, kFuseW) : 0;
  const int nbR = gr.totalSegs > 0 ? div_up(gr.totalSegs, kFuseW) : 0;
  const int nbC = wC ? nbR : 0, nbH = wH ? nbR : 0;
  if (nbW + nbC + nbH == 0) return CT_STATUS_SUCCESS;
  const bool useLB = maxSegs > 1;
  Ws wsC{nullptr, nullptr, 1 + 5}, wsH{nullptr, nullptr, 1 + 5};
  if (useLB) {
    const size_t bytes = (size_t)3 * ((size_t)gr.totalSegs + gr.totalSegs / kGroup) * (1 + 5) * 8;
    ctStatus_t st = CT_STATUS_SUCCESS;
    if (wC) st = workspace_for(s, bytes, &wsC, 0);
    if (st == CT_STATUS_SUCCESS && wH) st = workspace_for(s, bytes, &wsH, 1);
    if (st != CT_STATUS_SUCCESS) return st;
  }
  hipLaunchKernelGGL((k_pack_one<K>), dim3(nbW + nbC + nbH), dim3(kFuseW * kWave), 0, s, gw, gr,
                     wsC, wsH, useLB ? 1 : 0, nbW, nbC, nbH, prm, io);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// The fused pass: the window kernel and the two recurrence kernels read the
// same inputs and write disjoint outputs -- as one launch (tiles.one) when
// the window look-backs fit 64 bars and at least two parts are requested,
// else as up to three launches in stream order.
ctStatus_t pack_launch(int len, int rows, hipStream_t s, const PackP& p, const PackIO& io,
                       PackTiles t) {
  const int parts = ((p.mask & kWinMask) ? 1 : 0) + ((p.mask & kCloseRecMask) ? 1 : 0) +
                    ((p.mask & kHlcRecMask) ? 1 : 0);
  if (t.one && parts >= 2 && window_halo(p) <= 64) {
    if (t.one == 4) return launch_one<4>(len, rows, s, p, io);
    if (t.one == 8) return launch_one<8>(len, rows, s, p, io);
    return fail(CT_STATUS_INVALID_ARGUMENT, "pack: unknown single-launch tile %d", t.one);
  }
  ctStatus_t st = CT_STATUS_SUCCESS;
  if (p.mask & kWinMask) st = launch_windows(len, rows, s, p, io);
  if (st == CT_STATUS_SUCCESS && (p.mask & kCloseRecMask))
    st = launch_part<1>(t.close, len, rows, s, p, io);
  if (st == CT_STATUS_SUCCESS && (p.mask & kHlcRecMask))
    st = launch_part<2>(t.hlc, len, rows, s, p, io);
  return st;
}

// Default tiles by row length: rows that fit one 256-bar wave use 4 bars per
// lane (most waves for short rows); longer rows use 8 bars per lane, which
// halves the look-back segments -- one wave per workgroup for the close
// kernel (fewest waves waiting on one workgroup's slowest segment), four
// for the high/low/close kernel.  The pass runs as one launch.
PackTiles default_tiles(int len) {
  if (len <= 4 * kWave) return PackTiles{{4, 4}, {4, 4}, 4};
  return PackTiles{{8, 1}, {8, 4}, 8};
}

// Dense, single-series or ragged (maximum row length) pack with the default
// tiles.
ctStatus_t pack_dense(int len, int rows, hipStream_t s, const PackP& p, const PackIO& io) {
  return pack_launch(len, rows, s, p, io, default_tiles(len));
}

}  // namespace tca

// =================================================================== C ABI ==
using namespace tca;

namespace {

ctStatus_t pack_validate(const ctxPackParams* p, const ctxPackOutputs* o, PackP& P, PackIO& io,
                         const float* h, const float* l, const float* c) {
  if (!p || !o) return fail(CT_STATUS_INVALID_ARGUMENT, "pack: null params/outputs");
  float* outs[kONum] = {o->sma, o->ema, o->wma, o->dema, o->tema, o->rsi, o->macd,
                        o->macd_signal, o->macd_hist, o->bb_upper, o->bb_middle, o->bb_lower,
                        o->atr, o->adx, o->stoch_k, o->stoch_d, o->willr, o->cci};
  P = PackP{};
  P.mask = 0;
  for (int i = 0; i < kONum; ++i) {
    io.out[i] = outs[i];
    if (outs[i]) P.mask |= 1u << i;
  }
  if (P.mask == 0) return fail(CT_STATUS_INVALID_ARGUMENT, "pack: no output requested");
  io.h = h;
  io.l = l;
  io.c = c;
  const bool needHL = P.mask & ((1u << kOAtr) | (1u << kOAdx) | (1u << kOStK) | (1u << kOStD) |
                                (1u << kOWillr) | (1u << kOCci));
  if (!c || (needHL && (!h || !l))) return fail(CT_STATUS_INVALID_ARGUMENT, "pack: null input");
  if (!needHL) {
    io.h = c;
    io.l = c;
  }
  auto need = [&](int bit, int v, int lo, const char* what) -> bool {
    if (!(P.mask & (1u << bit))) return true;
    if (v < lo) {
      fail(CT_STATUS_INVALID_ARGUMENT, "pack: invalid %s period", what);
      return false;
    }
    return true;
  };
  // EMA, DEMA and TEMA share one cascade: requested ones must agree.
  int pe = 0;
  for (int b : {kOEma, kODema, kOTema}) {
    if (!(P.mask & (1u << b))) continue;
    const int v = b == kOEma ? p->ema_period : b == kODema ? p->dema_period : p->tema_period;
    if (pe && v != pe)
      return fail(CT_STATUS_UNSUPPORTED, "pack: EMA/DEMA/TEMA share one period in a pack");
    pe = v;
  }
  P.smaP = p->sma_period;
  P.emaP = pe > 0 ? pe : 2;
  P.wmaP = p->wma_period;
  P.rsiP = p->rsi_period;
  P.mFast = p->macd_fast;
  P.mSlow = p->macd_slow;
  P.mSig = p->macd_signal;
  if (P.mSlow < P.mFast) std::swap(P.mSlow, P.mFast);
  P.bbP = p->bb_period;
  P.bbUp = p->bb_up;
  P.bbDn = p->bb_dn;
  P.atrP = p->atr_period;
  P.adxP = p->adx_period;
  P.stK = p->stoch_k;
  P.stD = p->stoch_d;
  P.willrP = p->willr_period;
  P.cciP = p->cci_period;
  if (!need(kOSma, P.smaP, 1, "SMA") || !need(kOWma, P.wmaP, 1, "WMA") ||
      !need(kOEma, pe, 1, "EMA") || !need(kODema, pe, 1, "DEMA") ||
      !need(kOTema, pe, 1, "TEMA") || !need(kORsi, P.rsiP, 1, "RSI") ||
      !need(kOBbU, P.bbP, 1, "BBANDS") || !need(kOBbM, P.bbP, 1, "BBANDS") ||
      !need(kOBbL, P.bbP, 1, "BBANDS") || !need(kOAtr, P.atrP, 1, "ATR") ||
      !need(kOAdx, P.adxP, 2, "ADX") || !need(kOStK, P.stK, 1, "STOCH") ||
      !need(kOStD, P.stD, 1, "STOCH") || !need(kOWillr, P.willrP, 1, "WILLR") ||
      !need(kOCci, P.cciP, 1, "CCI"))
    return CT_STATUS_INVALID_ARGUMENT;
  const bool wantM = P.mask & ((1u << kOMacd) | (1u << kOMacdSig) | (1u << kOMacdHist));
  if (wantM && (P.mFast < 1 || P.mSlow < 1 || P.mSig < 1))
    return fail(CT_STATUS_INVALID_ARGUMENT, "pack: invalid MACD periods");
  if (!wantM) {
    P.mFast = 2;
    P.mSlow = 3;
    P.mSig = 2;
  }
  if ((P.mask & ((1u << kOStK) | (1u << kOStD))) && P.stD > kStochMaxD)
    return fail(CT_STATUS_UNSUPPORTED, "pack: STOCH dPeriod above %d is not provided", kStochMaxD);
  if (window_halo(P) > kPackMaxWindow)
    return fail(CT_STATUS_UNSUPPORTED, "pack: window look-back above %d bars is not provided",
                kPackMaxWindow);
  if (!(P.mask & (1u << kORsi))) P.rsiP = 2;
  if (!(P.mask & (1u << kOAtr))) P.atrP = 2;
  if (!(P.mask & (1u << kOAdx))) P.adxP = 2;
  return CT_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_pack_batch_d(const float* h, const float* l, const float* c,
                                         int seriesLen, int numSeries, const ctxPackParams* p,
                                         const ctxPackOutputs* o, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (seriesLen <= 0 || numSeries <= 0)
      return fail(CT_STATUS_INVALID_ARGUMENT, "pack: seriesLen and numSeries must be positive");
    PackP P;
    PackIO io;
    ctStatus_t st = pack_validate(p, o, P, io, h, l, c);
    if (st != CT_STATUS_SUCCESS) return st;
    st = pack_dense(seriesLen, numSeries, s, P, io);
    if (st != CT_STATUS_SUCCESS) return st;
    return finish_batch(s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_pack_ragged_d(const float* h, const float* l, const float* c,
                                          ctxRagged layout, const ctxPackParams* p,
                                          const ctxPackOutputs* o, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    const Ragged* r = ragged_of(layout);
    if (!r) return fail(CT_STATUS_INVALID_ARGUMENT, "pack: null layout");
    PackP P;
    PackIO io;
    ctStatus_t st = pack_validate(p, o, P, io, h, l, c);
    if (st != CT_STATUS_SUCCESS) return st;
    RaggedScope scope(r);
    st = pack_dense(r->maxLen, r->rows, s, P, io);
    if (st != CT_STATUS_SUCCESS) return st;
    return finish_batch(s);
  });
}

}  // extern "C"

// ================================================================== plans ==
// A plan is one fused pass over fixed buffers, tuned once and captured into
// a hipGraph: later passes are a single graph launch.
struct ctxPlan_st {
  hipStream_t stream = nullptr;  // private stream the graph was captured on
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PackTiles tiles{{4, 4}, {4, 4}, 4};
  unsigned mask = 0;
  int runK = 4, runW = 4;  // tile of the launch that runs (the close kernel's if several)
  float tunedUs = 0.0f, defaultUs = 0.0f;
};

namespace {

// Tuned tiles per shape class, so a second plan of the same shape skips the
// measurement.
struct TuneKey {
  int device, rows, len, ragged;
  unsigned mask;
  bool operator<(const TuneKey& o) const {
    return std::tie(device, rows, len, ragged, mask) <
           std::tie(o.device, o.rows, o.len, o.ragged, o.mask);
  }
};
struct Tuned {
  PackTiles tiles;
  float tunedUs, defaultUs;
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuned> g_tuned;

// Average microseconds of `reps` passes with tiles `t` on `st`.
ctStatus_t time_tiles(int len, int rows, hipStream_t st, const PackP& P, const PackIO& io,
                      PackTiles t, int reps, float* us) {
  for (int i = 0; i < 2; ++i) {
    ctStatus_t r = pack_launch(len, rows, st, P, io, t);
    if (r != CT_STATUS_SUCCESS) return r;
  }
  hipEvent_t e0, e1;
  TCA_HIP(hipEventCreate(&e0));
  TCA_HIP(hipEventCreate(&e1));
  ctStatus_t r = CT_STATUS_SUCCESS;
  (void)hipEventRecord(e0, st);
  for (int i = 0; i < reps && r == CT_STATUS_SUCCESS; ++i) r = pack_launch(len, rows, st, P, io, t);
  (void)hipEventRecord(e1, st);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (r != CT_STATUS_SUCCESS) return r;
  if (e != hipSuccess) return hip_fail(e, "plan: timing");
  *us = ms * 1000.0f / reps;
  return CT_STATUS_SUCCESS;
}

// Best tile of one recurrence kernel (outputs `part` of P), each candidate
// timed alone.
ctStatus_t tune_part(int len, int rows, hipStream_t st, const PackP& P, const PackIO& io,
                     unsigned part, PackCfg* best) {
  PackP Q = P;
  Q.mask = P.mask & part;
  float bestUs = 0.0f;
  for (const PackCfg& cf : kPackCfgs) {
    float us = 0.0f;
    ctStatus_t r = time_tiles(len, rows, st, Q, io, PackTiles{cf, cf, 0}, 10, &us);
    if (r != CT_STATUS_SUCCESS) return r;
    if (bestUs == 0.0f || us < bestUs) {
      bestUs = us;
      *best = cf;
    }
  }
  return CT_STATUS_SUCCESS;
}

void plan_free(ctxPlan p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->stream) {
    workspace_release(p->stream);
    (void)hipStreamDestroy(p->stream);
  }
  delete p;
}

}  // namespace

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_pack_plan_create(const float* h, const float* l, const float* c,
                                             int seriesLen, int numSeries, ctxRagged layout,
                                             const ctxPackParams* params,
                                             const ctxPackOutputs* outputs, int tune,
                                             ctxPlan* out) {
  return guarded([&]() -> ctStatus_t {
    if (!out) return fail(CT_STATUS_INVALID_ARGUMENT, "plan: null output handle");
    *out = nullptr;
    const Ragged* rg = ragged_of(layout);
    if (!rg && (seriesLen <= 0 || numSeries <= 0))
      return fail(CT_STATUS_INVALID_ARGUMENT, "plan: seriesLen and numSeries must be positive");
    PackP P;
    PackIO io;
    ctStatus_t st = pack_validate(params, outputs, P, io, h, l, c);
    if (st != CT_STATUS_SUCCESS) return st;
    const int len = rg ? rg->maxLen : seriesLen, rows = rg ? rg->rows : numSeries;
    RaggedScope scope(rg);
    ctxPlan p = new ctxPlan_st;
    auto bail = [&](ctStatus_t r) {
      plan_free(p);
      return r;
    };
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: stream"));
    int dev = 0;
    (void)hipGetDevice(&dev);
    const TuneKey key{dev, rows, rg ? (int)(rg->total & 0x7fffffff) : len, rg ? 1 : 0, P.mask};
    bool cached = false;
    const PackTiles dflt = default_tiles(len);
    p->tiles = dflt;
    p->mask = P.mask;
    {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tuned.find(key);
      if (it != g_tuned.end()) {
        p->tiles = it->second.tiles;
        p->tunedUs = it->second.tunedUs;
        p->defaultUs = it->second.defaultUs;
        cached = true;
      }
    }
    // every candidate runs eagerly before capture: each reserves the
    // look-back workspace its tile needs on the plan's stream
    if (tune && !cached) {
      PackTiles t = dflt;
      t.one = 0;
      ctStatus_t r = CT_STATUS_SUCCESS;
      if (P.mask & kCloseRecMask) r = tune_part(len, rows, p->stream, P, io, kCloseRecMask, &t.close);
      if (r == CT_STATUS_SUCCESS && (P.mask & kHlcRecMask))
        r = tune_part(len, rows, p->stream, P, io, kHlcRecMask, &t.hlc);
      // the whole pass decides: three launches with the tuned tiles, or one
      // launch of 4 or 8 bars per lane
      float du = 0.0f;
      if (r == CT_STATUS_SUCCESS) r = time_tiles(len, rows, p->stream, P, io, dflt, 10, &du);
      PackTiles best = dflt;
      float bu = du;
      for (int one : {0, 4, 8}) {
        if (r != CT_STATUS_SUCCESS) break;
        PackTiles c = t;
        c.one = one;
        float us = 0.0f;
        r = time_tiles(len, rows, p->stream, P, io, c, 10, &us);
        if (r == CT_STATUS_SUCCESS && us < bu) {
          bu = us;
          best = c;
        }
      }
      if (r != CT_STATUS_SUCCESS) return bail(r);
      p->tiles = best;
      p->tunedUs = bu;
      p->defaultUs = du;
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tuned[key] = Tuned{p->tiles, p->tunedUs, p->defaultUs};
    } else {
      float us = 0.0f;
      ctStatus_t r = time_tiles(len, rows, p->stream, P, io, p->tiles, 1, &us);
      if (r != CT_STATUS_SUCCESS) return bail(r);
      if (!cached) p->tunedUs = p->defaultUs = us;
    }
    {
      const int parts = ((P.mask & kWinMask) ? 1 : 0) + ((P.mask & kCloseRecMask) ? 1 : 0) +
                        ((P.mask & kHlcRecMask) ? 1 : 0);
      const PackCfg& t = (P.mask & kCloseRecMask) ? p->tiles.close : p->tiles.hlc;
      const bool one = p->tiles.one && parts >= 2 && window_halo(P) <= 64;
      p->runK = one ? p->tiles.one : t.K;
      p->runW = one ? kFuseW : t.W;
    }
    e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: begin capture"));
    ctStatus_t r = pack_launch(len, rows, p->stream, P, io, p->tiles);
    e = hipStreamEndCapture(p->stream, &p->graph);
    if (r != CT_STATUS_SUCCESS) return bail(r);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: end capture"));
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: instantiate"));
    *out = p;
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ctx_plan_launch(ctxPlan plan, hipStream_t stream) {
  return guarded([&]() -> ctStatus_t {
    if (!plan || !plan->exec) return fail(CT_STATUS_INVALID_ARGUMENT, "plan: null plan");
    TCA_HIP(hipGraphLaunch(plan->exec, stream));
    return finish_batch(stream);
  });
}

CTAPI_EXPORT ctStatus_t ctx_plan_info(ctxPlan plan, int* barsPerLane, int* wavesPerGroup,
                                      float* tunedUs, float* defaultUs) {
  return guarded([&]() -> ctStatus_t {
    if (!plan) return fail(CT_STATUS_INVALID_ARGUMENT, "plan: null plan");
    if (barsPerLane) *barsPerLane = plan->runK;
    if (wavesPerGroup) *wavesPerGroup = plan->runW;
    if (tunedUs) *tunedUs = plan->tunedUs;
    if (defaultUs) *defaultUs = plan->defaultUs;
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ctx_plan_destroy(ctxPlan plan) {
  return guarded([&]() -> ctStatus_t {
    plan_free(plan);
    return CT_STATUS_SUCCESS;
  });
}

}  // extern "C"
