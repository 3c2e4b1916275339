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

#include <array>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/tacuda_amd.h"
#include "divide.hpp"
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
  int twoPass;  // long dense rows may take the two-pass recurrences
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
  const DivBy byPP = div_by_of(pp);
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
        b1[1] = b1[1] - div_by(b1[1], byPP) + pdm[i];
        b1[2] = b1[2] - div_by(b1[2], byPP) + mdm[i];
        b1[3] = b1[3] - div_by(b1[3], byPP) + tr[i];
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
        yp = yp - div_by(yp, byPP) + pdm[i];
        yq = yq - div_by(yq, byPP) + mdm[i];
        yt = yt - div_by(yt, byPP) + tr[i];
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
      sX = div_by(part[0], byPP);
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
        sp = sp - div_by(sp, byPP) + pd;
        sm = sm - div_by(sm, byPP) + md;
        st = st - div_by(st, byPP) + tr_at(t);
        double pdi, mdi, dx;
        dir_outputs(sp, sm, st, pdi, mdi, dx);
        sum += dx;
      }
      sX = div_by(sum, byPP);
    }
  }
  const double beta2[1] = {bx};
  double P2[1] = {pow_n(bx, K)}, b2[1] = {0.0}, y2[1];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    if (t == Lax) b2[0] = sX;
    else if (t > Lax) b2[0] = div_by((b2[0] * (pp - 1.0)) + dxv[i], byPP);
  }
  bcarry<1>(c, 1, b2, P2, beta2, y2);
  {
    double yx = y2[0];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      ov[i] = qnanf_();
      if (t == Lax) yx = sX;
      else if (t > Lax) yx = div_by((yx * (pp - 1.0)) + dxv[i], byPP);
      if (t >= Lax) ov[i] = (float)yx;
    }
    pack_store<T>(c, lo, ov, io.out[kOAdx]);
  }
}

// ================================================================ windows ==
// Window families over a column tile (one bar per lane per 64-bar chunk),
// each window evaluated directly in the oracle's order.  The sections read
// the staged h, l, c; the extrema section uses d0 and w, CCI its own d.

// SMA and the BBANDS middle band of one period are the same window mean:
// the BBANDS loop then writes both.
__device__ __forceinline__ bool win_sma_in_bb(const PackP& P) {
  const unsigned m = P.mask;
  const bool bbAll = m & ((1u << kOBbU) | (1u << kOBbM) | (1u << kOBbL));
  return (m & (1u << kOSma)) && bbAll && P.smaP == P.bbP;
}

template <class C, int IL = 4>
__device__ __forceinline__ void win_means(const Seg& c, ColLds<C::E>& s, const PackP& P,
                                          const PackIO& io) {
  const unsigned m = P.mask;
  const int lane = lane_id();
  const int base = c.s0 - C::H;  // bar of e = 0
  const bool smaInBb = win_sma_in_bb(P);
  // the output chunks in groups of IL, evaluated side by side
  constexpr int NJ = C::NC - C::NH;
  static_assert(NJ % IL == 0, "chunk groups");
  if ((m & (1u << kOSma)) && !smaInBb) {
    const int p = P.smaP;
    const DivBy byP = div_by_of((double)p);
    for (int g = 0; g < NJ; g += IL) {
      double sum[IL];
#pragma unroll
      for (int u = 0; u < IL; ++u) sum[u] = 0.0;
      const int q0 = (C::NH + g) * kWave + lane - p + 1;
      for (int k = 0; k < p; ++k) {
#pragma unroll
        for (int u = 0; u < IL; ++u) sum[u] += (double)lds_f(s.c, q0 + u * kWave + k);
      }
#pragma unroll
      for (int u = 0; u < IL; ++u) {
        const int j = C::NH + g + u, e = j * kWave + lane;
        const float v = (base + e >= p - 1) ? (float)div_by(sum[u], byP) : qnanf_();
        col_store<C>(io.out[kOSma], c, j, v);
      }
    }
  }
  if (m & (1u << kOWma)) {
    const int p = P.wmaP;
    const double divider = (double)p * (p + 1) / 2.0;
    const DivBy byDiv = div_by_of(divider);
    for (int g = 0; g < NJ; g += IL) {
      double sum[IL];
#pragma unroll
      for (int u = 0; u < IL; ++u) sum[u] = 0.0;
      if (p > 1) {
        const int q0 = (C::NH + g) * kWave + lane - p + 1;
        double w = 1.0;
        for (int k = 0; k < p; ++k, w += 1.0) {
#pragma unroll
          for (int u = 0; u < IL; ++u) sum[u] += (double)lds_f(s.c, q0 + u * kWave + k) * w;
        }
      }
#pragma unroll
      for (int u = 0; u < IL; ++u) {
        const int j = C::NH + g + u, e = j * kWave + lane;
        const float v = p == 1 ? lds_f(s.c, e) : (float)div_by(sum[u], byDiv);
        col_store<C>(io.out[kOWma], c, j, (base + e >= p - 1) ? v : qnanf_());
      }
    }
  }
}

template <class C, int IL = 4>
__device__ __forceinline__ void win_bbands(const Seg& c, ColLds<C::E>& s, const PackP& P,
                                           const PackIO& io) {
  const unsigned m = P.mask;
  const int lane = lane_id();
  const int base = c.s0 - C::H;
  const bool bbAll = m & ((1u << kOBbU) | (1u << kOBbM) | (1u << kOBbL));
  const bool smaInBb = win_sma_in_bb(P);
  if (bbAll) {
    const int p = P.bbP;
    const double inv = 1.0 / (double)p;
    const DivBy byP = div_by_of((double)p);
    // the output chunks in groups of IL, evaluated side by side
    constexpr int NJ = C::NC - C::NH;
    static_assert(NJ % IL == 0, "chunk groups");
    for (int g = 0; g < NJ; g += IL) {
    const int q0 = (C::NH + g) * kWave + lane - p + 1;
    double shift[IL], sum[IL], s1[IL], s2[IL];
#pragma unroll
    for (int u = 0; u < IL; ++u) {
      shift[u] = (double)lds_f(s.c, q0 + u * kWave);
      sum[u] = s1[u] = s2[u] = 0.0;
    }
    for (int k = 0; k < p; ++k) {
#pragma unroll
      for (int u = 0; u < IL; ++u) {
        const double x = (double)lds_f(s.c, q0 + u * kWave + k);
        sum[u] += x;
        const double d = x - shift[u];
        s1[u] += d;
        s2[u] += d * d;
      }
    }
#pragma unroll
    for (int u = 0; u < IL; ++u) {
      const int j = C::NH + g + u, e = j * kWave + lane;
      const double mean = div_by(sum[u], byP);
      const double var = col_var_finish(s1[u], s2[u], inv);
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
      if (smaInBb) col_store<C>(io.out[kOSma], c, j, ok ? (float)mean : qnanf_());
      col_store<C>(io.out[kOBbU], c, j, ok ? up : qnanf_());
      col_store<C>(io.out[kOBbM], c, j, ok ? (float)mean : qnanf_());
      col_store<C>(io.out[kOBbL], c, j, ok ? lo : qnanf_());
    }
    }
  }
}

template <class C, int IL = 4>
__device__ __forceinline__ void win_extrema(const Seg& c, ColLds<C::E>& s, const PackP& P,
                                            const PackIO& io) {
  const unsigned m = P.mask;
  const int lane = lane_id();
  const int base = c.s0 - C::H;
  const bool wantSt = m & ((1u << kOStK) | (1u << kOStD));
  const bool wantW = m & (1u << kOWillr);
  if (wantSt || wantW) {
    // WILLR first when it shares the extrema table with STOCH, whose
    // smoothing then reuses the table's LDS
    const int kp = wantSt ? P.stK : P.willrP;
    int k = col_extrema_table<C>(s, kp);
    auto willr = [&]() {
      const int p = P.willrP;
#pragma unroll
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
      const DivBy byD = div_by_of((double)dp);
      const int lb = kp - 1 + 2 * (dp - 1);
      // fast %K from bar s0 - 2(d-1), slow %K from s0 - (d-1)
      const int eK = C::H - 2 * (dp - 1), eS = C::H - (dp - 1);
      if constexpr (IL == 1) {
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
          if (e >= eS) lds_wd(d1, e, div_by(sum, byD));
        }
        col_sync();
        for (int j = C::NH; j < C::NC; ++j) {
          const int e = j * kWave + lane;
          double sum = 0.0;
          for (int q = e - dp + 1; q <= e; ++q) sum += lds_d(d1, q);
          const bool ok = base + e >= lb;
          col_store<C>(io.out[kOStK], c, j, ok ? (float)lds_d(d1, e) : qnanf_());
          col_store<C>(io.out[kOStD], c, j, ok ? (float)div_by(sum, byD) : qnanf_());
        }
        col_sync();
      } else {
        // every chunk side by side (chunks before eK or eS compute values
        // that are not stored)
  #pragma unroll
        for (int j = 0; j < C::NC; ++j) {
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
        {
          double sum[C::NC];
  #pragma unroll
          for (int j = 0; j < C::NC; ++j) sum[j] = 0.0;
          for (int q = 0; q < dp; ++q) {
  #pragma unroll
            for (int j = 0; j < C::NC; ++j) {
              const int i = j * kWave + lane - dp + 1 + q;
              sum[j] += lds_d(s.d0, i < 0 ? 0 : i);
            }
          }
  #pragma unroll
          for (int j = 0; j < C::NC; ++j) {
            const int e = j * kWave + lane;
            if (e >= eS) lds_wd(d1, e, div_by(sum[j], byD));
          }
        }
        col_sync();
        {
          constexpr int NJ = C::NC - C::NH;
          double sum[NJ];
  #pragma unroll
          for (int u = 0; u < NJ; ++u) sum[u] = 0.0;
          for (int q = 0; q < dp; ++q) {
  #pragma unroll
            for (int u = 0; u < NJ; ++u) sum[u] += lds_d(d1, (C::NH + u) * kWave + lane - dp + 1 + q);
          }
  #pragma unroll
          for (int u = 0; u < NJ; ++u) {
            const int j = C::NH + u, e = j * kWave + lane;
            const bool ok = base + e >= lb;
            col_store<C>(io.out[kOStK], c, j, ok ? (float)lds_d(d1, e) : qnanf_());
            col_store<C>(io.out[kOStD], c, j, ok ? (float)div_by(sum[u], byD) : qnanf_());
          }
        }
        col_sync();
      }
      if (wantW && P.willrP != kp) {
        k = col_extrema_table<C>(s, P.willrP);
        willr();
      }
    } else {
      willr();
    }
  }
}

template <class C, int IL = 4>
__device__ __forceinline__ void win_cci(const Seg& c, ColLds<C::E>& s, double* d, const PackP& P,
                                        const PackIO& io) {
  const unsigned m = P.mask;
  const int lane = lane_id();
  const int base = c.s0 - C::H;
  if (m & (1u << kOCci)) {
    const int p = P.cciP;
    const DivBy byP = div_by_of((double)p), by3 = div_by_of(3.0);
    const int e0 = C::H - (p - 1);
    for (int j = e0 / kWave; j < C::NC; ++j) {
      const int e = j * kWave + lane;
      const double tp =
          div_by((double)lds_f(s.h, e) + (double)lds_f(s.l, e) + (double)lds_f(s.c, e), by3);
      lds_wd(d, e, tp);
    }
    col_sync();
    // the output chunks in groups of IL, evaluated side by side
    constexpr int NJ = C::NC - C::NH;
    static_assert(NJ % IL == 0, "chunk groups");
    for (int g = 0; g < NJ; g += IL) {
      const int eo = (C::NH + g) * kWave + lane;
      double sum[IL], sma[IL], dev[IL];
#pragma unroll
      for (int u = 0; u < IL; ++u) sum[u] = dev[u] = 0.0;
      for (int k = 0; k < p; ++k) {
#pragma unroll
        for (int u = 0; u < IL; ++u) sum[u] += lds_d(d, eo + u * kWave - k);
      }
#pragma unroll
      for (int u = 0; u < IL; ++u) sma[u] = div_by(sum[u], byP);
      for (int k = 0; k < p; ++k) {
#pragma unroll
        for (int u = 0; u < IL; ++u) dev[u] += fabs(lds_d(d, eo + u * kWave - k) - sma[u]);
      }
#pragma unroll
      for (int u = 0; u < IL; ++u) {
        const int j = C::NH + g + u, e = j * kWave + lane;
        const double md = div_by(dev[u], byP);
        const double num = lds_d(d, e) - sma[u];
        const float v = (num != 0.0 && md != 0.0) ? (float)(num / (0.015 * md)) : 0.0f;
        col_store<C>(io.out[kOCci], c, j, (base + e >= p - 1) ? v : qnanf_());
      }
    }
  }
}

template <class C>
__device__ __forceinline__ bool win_need_hl(const PackP& P) {
  return P.mask & ((1u << kOStK) | (1u << kOStD) | (1u << kOWillr) | (1u << kOCci));
}

// Window chunks evaluated side by side when one wave computes every window
// family (four independent accumulator chains per family loop).
constexpr int kWinIL = 4;

// Every window family of the pass, one wave.
template <class C>
__device__ __forceinline__ void pack_windows(const Seg& c, ColLds<C::E>& s, const PackP& P,
                                             const PackIO& io) {
  if (win_need_hl<C>(P)) {
    col_stage<C>(s.h, io.h, c);
    col_stage<C>(s.l, io.l, c);
  }
  col_stage<C>(s.c, io.c, c);
  col_sync();
  win_means<C, kWinIL>(c, s, P, io);
  win_bbands<C, kWinIL>(c, s, P, io);
  win_extrema<C, kWinIL>(c, s, P, io);
  col_sync();
  win_cci<C, kWinIL>(c, s, s.d0, P, io);
}

// Per-block LDS of the split window pass: one staged tile, the extrema
// section's d0 and w, CCI's own array.
template <int E>
struct ColSplitLds {
  ColLds<E> t;
  double d[E];
};

// The window families of one segment over a workgroup of four waves, one
// family group per wave: BBANDS (and SMA of its period); STOCH and WILLR;
// CCI; SMA and WMA.
template <class C, int IL>
__device__ __forceinline__ void pack_windows_split(const Seg& c, ColSplitLds<C::E>& L,
                                                   const PackP& P, const PackIO& io, int w) {
  const bool hl = win_need_hl<C>(P);
  if (w == 0) col_stage<C>(L.t.c, io.c, c);
  else if (w == 1 && hl) col_stage<C>(L.t.h, io.h, c);
  else if (w == 2 && hl) col_stage<C>(L.t.l, io.l, c);
  __syncthreads();
  if (w == 0) win_bbands<C, IL>(c, L.t, P, io);
  else if (w == 1) win_extrema<C, IL>(c, L.t, P, io);
  else if (w == 2) win_cci<C, IL>(c, L.t, L.d, P, io);
  else win_means<C, IL>(c, L.t, P, io);
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
  KTime kt(s, K == 4 ? "k_pack_one<4>" : "k_pack_one<8>");
  hipLaunchKernelGGL((k_pack_one<K>), dim3(nbW + nbC + nbH), dim3(kFuseW * kWave), 0, s, gw, gr,
                     wsC, wsH, useLB ? 1 : 0, nbW, nbC, nbH, prm, io);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// ---------------------------------------------------------- two passes --
// Long dense rows: the fused pass's recurrences -- EMA, RSI, MACD with its
// signal, ATR -- in two streaming passes instead of look-back chains, like
// the fused long scan: a fold pass writes every 512-bar segment's composite
// (the affine map of its bars, from which the state after the segment
// follows from the state before it), a wave per 64 segments scans them into
// in-group prefixes and group composites, a wave per row scans those into
// the state entering every group, and an emit pass replays the bars from
// there with the families' TA-Lib arithmetic.  Fold and emit give each
// 512-bar segment a workgroup with one wave per recurrence group; the
// window part of the pass runs in the fold launch.
//
// Every recurrence maps y -> A y + b over a run of bars: identity before its
// seed bar, a reset to the seed value at it, the family's step after it.
// The MACD signal g also reads the MACD line, so its row is
// g -> agf f + ags s + agg g + bgg over the fast and slow EMAs f and s.
// Products of nonzero multipliers that underflow are clamped to the
// smallest normal double, so a carried Inf survives them as in the
// sequential recurrence.
constexpr int kPK = 8;               // bars per lane
constexpr int kPSeg = kPK * kWave;   // bars per segment
constexpr int kPW = 4;               // waves per workgroup
constexpr int kPMinLen = 16 * kPSeg; // shortest row taking the two passes
// Passes whose segments x groups-per-row stay within this many scan the
// row's group composites inside the emit launch (no k_pipe_rows launch).
constexpr long long kPipeRowsInEmit = 1ll << 17;
constexpr unsigned kPUnsupported = (1u << kODema) | (1u << kOTema) | (1u << kOAdx);
constexpr int kPWinIL = 1;  // window chunks side by side in the fold launch's window waves

// Multiplier fields of a composite, in the order of PipeP::q: Ae, Ag, Al,
// Af, As, agf, ags, agg, Aa.
constexpr int kQF = 9;

struct PipeP {
  int n, rows, segsPerRow;
  int wE, wR, wM, wA;
  int pe, pr, fast, slow, sig, pa;
  double eb, ek, inv, pm1, br, fb, fk, sb, sk, gb, gk, wb, wa;
  // Multipliers of a steady run of 8 * 2^s bars (every bar past every seed):
  // they do not depend on the bars, so steady segments fold and scan only
  // the offsets.  q[0] is one lane's run, q[s + 1] = q[s] after q[s].
  double q[7][kQF];
  // Per-lane multipliers of the two row-broadcast scan steps (lane l of a
  // steady wave after the row-shift steps, then after the first broadcast):
  // qt[(k * kQF + field) * 64 + lane]: the host table or k_pipe_qt's.
  double* qt;
};

struct PC {
  double Ae, be, Ag, bg, Al, bl, Af, bf, As, bs, agf, ags, agg, bgg, Aa, ba;
};

// Recurrence groups of a composite: EMA (e), RSI (G, L), MACD (f, s, g), ATR (a).
constexpr unsigned kGE = 1, kGR = 2, kGM = 4, kGA = 8, kGAll = 15;

__device__ __forceinline__ void pc_id(PC& p) {
  p.Ae = p.Ag = p.Al = p.Af = p.As = p.Aa = p.agg = 1.0;
  p.be = p.bg = p.bl = p.bf = p.bs = p.ba = p.agf = p.ags = p.bgg = 0.0;
}

__device__ __forceinline__ double mulc(double a, double m) {
  const double r = a * m;
  return (r == 0.0 && a != 0.0 && m != 0.0) ? 2.2250738585072014e-308 : r;
}

// l o e: the run e followed by the run l.
__device__ __forceinline__ PC pc_after(const PC& e, const PC& l) {
  PC r;
  r.be = fma(l.Ae, e.be, l.be);
  r.Ae = mulc(l.Ae, e.Ae);
  r.bg = fma(l.Ag, e.bg, l.bg);
  r.Ag = mulc(l.Ag, e.Ag);
  r.bl = fma(l.Al, e.bl, l.bl);
  r.Al = mulc(l.Al, e.Al);
  r.bf = fma(l.Af, e.bf, l.bf);
  r.Af = mulc(l.Af, e.Af);
  r.bs = fma(l.As, e.bs, l.bs);
  r.As = mulc(l.As, e.As);
  r.ba = fma(l.Aa, e.ba, l.ba);
  r.Aa = mulc(l.Aa, e.Aa);
  r.agf = fma(l.agf, e.Af, l.agg * e.agf);
  r.ags = fma(l.ags, e.As, l.agg * e.ags);
  r.agg = l.agg * e.agg;
  r.bgg = fma(l.agf, e.bf, fma(l.ags, e.bs, fma(l.agg, e.bgg, l.bgg)));
  return r;
}

// The composite of the lane before (identity in lane 0).
__device__ __forceinline__ PC pc_prev_lane(const PC& p) {
  PC r;
  r.Ae = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.Ae); r.be = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.be);
  r.Ag = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.Ag); r.bg = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.bg);
  r.Al = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.Al); r.bl = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.bl);
  r.Af = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.Af); r.bf = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.bf);
  r.As = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.As); r.bs = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.bs);
  r.agf = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.agf); r.ags = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.ags);
  r.agg = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.agg); r.bgg = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.bgg);
  r.Aa = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, p.Aa); r.ba = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, p.ba);
  return r;
}

// One DPP step of the wave scan, one recurrence group at a time (the same
// composition as pc_after).  Lanes the step does not reach read the identity,
// whose composition is exact, so no lane branches.
template <int CTRL, int RM, int BM, unsigned GM>
__device__ __forceinline__ void pc_step(PC& F) {
  if constexpr (GM & kGE) {
    const double A = dpp_d<CTRL, RM, BM>(1.0, F.Ae), b = dpp_d<CTRL, RM, BM>(0.0, F.be);
    F.be = fma(F.Ae, b, F.be);
    F.Ae = mulc(F.Ae, A);
  }
  if constexpr (GM & kGR) {
    const double Ag = dpp_d<CTRL, RM, BM>(1.0, F.Ag), bg = dpp_d<CTRL, RM, BM>(0.0, F.bg);
    const double Al = dpp_d<CTRL, RM, BM>(1.0, F.Al), bl = dpp_d<CTRL, RM, BM>(0.0, F.bl);
    F.bg = fma(F.Ag, bg, F.bg);
    F.Ag = mulc(F.Ag, Ag);
    F.bl = fma(F.Al, bl, F.bl);
    F.Al = mulc(F.Al, Al);
  }
  if constexpr (GM & kGM) {
    const double Af = dpp_d<CTRL, RM, BM>(1.0, F.Af), bf = dpp_d<CTRL, RM, BM>(0.0, F.bf);
    const double As = dpp_d<CTRL, RM, BM>(1.0, F.As), bs = dpp_d<CTRL, RM, BM>(0.0, F.bs);
    const double agf = dpp_d<CTRL, RM, BM>(0.0, F.agf), ags = dpp_d<CTRL, RM, BM>(0.0, F.ags);
    const double agg = dpp_d<CTRL, RM, BM>(1.0, F.agg), bgg = dpp_d<CTRL, RM, BM>(0.0, F.bgg);
    F.bgg = fma(F.agf, bf, fma(F.ags, bs, fma(F.agg, bgg, F.bgg)));
    F.agf = fma(F.agf, Af, F.agg * agf);
    F.ags = fma(F.ags, As, F.agg * ags);
    F.agg = F.agg * agg;
    F.bf = fma(F.Af, bf, F.bf);
    F.Af = mulc(F.Af, Af);
    F.bs = fma(F.As, bs, F.bs);
    F.As = mulc(F.As, As);
  }
  if constexpr (GM & kGA) {
    const double A = dpp_d<CTRL, RM, BM>(1.0, F.Aa), b = dpp_d<CTRL, RM, BM>(0.0, F.ba);
    F.ba = fma(F.Aa, b, F.ba);
    F.Aa = mulc(F.Aa, A);
  }
}

// Inclusive scan over the wave's lanes (lane 0 earliest).
template <unsigned GM = kGAll>
__device__ __forceinline__ void pc_scan(PC& F) {
  pc_step<kRowShr1, 0xf, 0xf, GM>(F);
  pc_step<kRowShr2, 0xf, 0xf, GM>(F);
  pc_step<kRowShr4, 0xf, 0xf, GM>(F);
  pc_step<kRowShr8, 0xf, 0xf, GM>(F);
  pc_step<kRowBcast15, 0xa, 0xf, GM>(F);
  pc_step<kRowBcast31, 0xc, 0xf, GM>(F);
}

__device__ __forceinline__ void pc_set_a(PC& F, const double (&q)[kQF]) {
  F.Ae = q[0]; F.Ag = q[1]; F.Al = q[2]; F.Af = q[3]; F.As = q[4];
  F.agf = q[5]; F.ags = q[6]; F.agg = q[7]; F.Aa = q[8];
}

// One scan step of a steady wave: the later run's multipliers A are known,
// so only the offsets move (the same arithmetic as pc_step).
template <int CTRL, int RM, int BM, unsigned GM, class Q>
__device__ __forceinline__ void pc_bstep(PC& F, const Q& A) {
  if constexpr (GM & kGE) {
    const double b = dpp_d<CTRL, RM, BM>(0.0, F.be);
    F.be = fma(A[0], b, F.be);
  }
  if constexpr (GM & kGR) {
    const double bg = dpp_d<CTRL, RM, BM>(0.0, F.bg), bl = dpp_d<CTRL, RM, BM>(0.0, F.bl);
    F.bg = fma(A[1], bg, F.bg);
    F.bl = fma(A[2], bl, F.bl);
  }
  if constexpr (GM & kGM) {
    const double bf = dpp_d<CTRL, RM, BM>(0.0, F.bf), bs = dpp_d<CTRL, RM, BM>(0.0, F.bs);
    const double bgg = dpp_d<CTRL, RM, BM>(0.0, F.bgg);
    F.bgg = fma(A[5], bf, fma(A[6], bs, fma(A[7], bgg, F.bgg)));
    F.bf = fma(A[3], bf, F.bf);
    F.bs = fma(A[4], bs, F.bs);
  }
  if constexpr (GM & kGA) {
    const double b = dpp_d<CTRL, RM, BM>(0.0, F.ba);
    F.ba = fma(A[8], b, F.ba);
  }
}

// The lane's multipliers of the two broadcast steps (fields of GM only).
struct PQLane {
  double a4[kQF], a5[kQF];
};
template <unsigned GM>
__device__ __forceinline__ void pq_lane(const PipeP& P, PQLane& L) {
  const int lane = lane_id();
#pragma unroll
  for (int f = 0; f < kQF; ++f) {
    const unsigned g = f == 0 ? kGE : f <= 2 ? kGR : f <= 7 ? kGM : kGA;
    if (GM & g) {
      L.a4[f] = P.qt[f * kWave + lane];
      L.a5[f] = P.qt[(kQF + f) * kWave + lane];
    } else {
      L.a4[f] = L.a5[f] = 1.0;
    }
  }
}

// Inclusive scan of a steady wave's offsets.  Multipliers are left as they
// were.
template <unsigned GM>
__device__ __forceinline__ void pc_scan_steady(PC& F, const PipeP& P, const PQLane& L) {
  pc_bstep<kRowShr1, 0xf, 0xf, GM>(F, P.q[0]);
  pc_bstep<kRowShr2, 0xf, 0xf, GM>(F, P.q[1]);
  pc_bstep<kRowShr4, 0xf, 0xf, GM>(F, P.q[2]);
  pc_bstep<kRowShr8, 0xf, 0xf, GM>(F, P.q[3]);
  pc_bstep<kRowBcast15, 0xa, 0xf, GM>(F, L.a4);
  pc_bstep<kRowBcast31, 0xc, 0xf, GM>(F, L.a5);
}

// State (e, G, L, f, s, g, atr) after a run, from the state before it.
struct PState {
  double e, G, L, f, s, g, a;
};
__device__ __forceinline__ PState pc_apply(const PC& p, const PState& y) {
  PState r;
  r.e = fma(p.Ae, y.e, p.be);
  r.G = fma(p.Ag, y.G, p.bg);
  r.L = fma(p.Al, y.L, p.bl);
  r.f = fma(p.Af, y.f, p.bf);
  r.s = fma(p.As, y.s, p.bs);
  r.g = fma(p.agf, y.f, fma(p.ags, y.s, fma(p.agg, y.g, p.bgg)));
  r.a = fma(p.Aa, y.a, p.ba);
  return r;
}

struct PSeeds {
  int f, Le, Lg, Lm, Ls, La;
  double sE, sG, sL, sF, sS, sSig, sA;
};

__device__ __forceinline__ float lane_f(float x, int j) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), j));
}

// body(t, x_t) for t = t0..t1 in order, uniform over the wave: the lanes load
// 64 bars at a time, the sequence reads them back lane by lane.  Every t lies
// before the end of the staged segment.
template <class T, class Body>
__device__ __forceinline__ void pipe_walk(const float* lds, const float* src, const Seg& c, int t0,
                                          int t1, Body&& body) {
  for (int b = t0; b <= t1; b += kWave) {
    const int t = b + lane_id();
    const float x = t <= t1 ? at<T>(lds, src, c, t) : 0.0f;
    const int m = min(kWave, t1 - b + 1);
    for (int j = 0; j < m; ++j) body(b + j, lane_f(x, j));
  }
}

// Seeds of the recurrences whose seed bar lies in this segment, in the
// oracle's order (TA-Lib's warm-up sums), evaluated by every lane.
template <class T, unsigned GM = kGAll>
__device__ __forceinline__ PSeeds pipe_seeds(const float* lc, const float* lh, const float* ll,
                                             const PackIO& io, const Seg& c, const PipeP& P,
                                             int f) {
  PSeeds S;
  S.f = f;
  S.Le = sat_add(f, P.pe - 1);
  S.Lg = sat_add(f, P.pr);
  S.Lm = sat_add(f, P.slow - 1);
  S.Ls = sat_add(S.Lm, P.sig - 1);
  S.La = P.pa;
  S.sE = S.sG = S.sL = S.sF = S.sS = S.sSig = S.sA = 0.0;
  if ((GM & kGE) && P.wE && here(c, S.Le)) {
    double t0 = 0.0;
    pipe_walk<T>(lc, io.c, c, f, S.Le, [&](int, float x) { t0 += (double)x; });
    S.sE = t0 / P.pe;
  }
  if ((GM & kGR) && P.wR && here(c, S.Lg)) {
    double prev = (double)at<T>(lc, io.c, c, f), g = 0.0, l = 0.0;
    pipe_walk<T>(lc, io.c, c, f + 1, S.Lg, [&](int, float x) {
      const double v = (double)x;
      const double d = v - prev;
      prev = v;
      const double gd = d > 0.0 ? d : 0.0;
      g += gd;
      l += gd - d;
    });
    S.sG = g * P.inv;
    S.sL = l * P.inv;
  }
  // the signal's seed (at Ls) needs the fast and slow seeds (at Lm <= Ls);
  // every bar read lies at or before a seed bar of this segment
  if ((GM & kGM) && P.wM && (here(c, S.Lm) || here(c, S.Ls))) {
    double ts = 0.0, tf = 0.0;
    const int ff = S.Lm - P.fast + 1;
    pipe_walk<T>(lc, io.c, c, f, S.Lm, [&](int t, float x) {
      ts += (double)x;
      if (t >= ff) tf += (double)x;
    });
    S.sS = ts / P.slow;
    S.sF = tf / P.fast;
    if (here(c, S.Ls)) {
      double ps = S.sS, pf = S.sF, psig = 0.0;
      psig += pf - ps;
      pipe_walk<T>(lc, io.c, c, S.Lm + 1, S.Ls, [&](int, float x) {
        const double v = (double)x;
        pf = fma(P.fb, pf, P.fk * v);
        ps = fma(P.sb, ps, P.sk * v);
        psig += pf - ps;
      });
      S.sSig = psig / P.sig;
    }
  }
  if ((GM & kGA) && P.wA && here(c, S.La)) {
    double tot = 0.0;
    for (int b = 1; b <= P.pa; b += kWave) {
      const int t = b + lane_id();
      const double tr = t <= P.pa ? true_range(at<T>(lh, io.h, c, t), at<T>(ll, io.l, c, t),
                                               at<T>(lc, io.c, c, t - 1))
                                  : 0.0;
      const int m = min(kWave, P.pa - b + 1);
      for (int j = 0; j < m; ++j) tot += bcast_d(tr, j);
    }
    S.sA = tot / P.pa;
  }
  return S;
}

// The lane's bars (close, high, low) and the close before the first.
struct PBars {
  float cv[kPK], hv[kPK], lv[kPK];
  float pc;
  __device__ float prev(int i) const { return i == 0 ? pc : cv[i - 1]; }
  __device__ double d(int i) const { return (double)cv[i] - (double)prev(i); }
  __device__ double tr(int i) const { return true_range(hv[i], lv[i], prev(i)); }
};

template <class T>
__device__ __forceinline__ void pipe_bars(const float* lc, const float* lh, const float* ll,
                                          const PackIO& io, const Seg& c, const PipeP& P, int a,
                                          PBars& B) {
  lane_vals<T>(lc, B.cv);
  B.pc = at<T>(lc, io.c, c, a - 1);
  // high and low are staged only for ATR; without it these values are unread
  lane_vals<T>(lh, B.hv);
  lane_vals<T>(ll, B.lv);
}

// Composite of the lane's bars from a zero state.  STEADY: every bar is past
// every seed (no per-bar seed tests).
template <bool STEADY, unsigned GM = kGAll>
__device__ __forceinline__ PC pipe_fold(const PipeP& P, const PSeeds& S, const PBars& B, int a,
                                        int s1) {
  PC F;
  pc_id(F);
#pragma unroll
  for (int i = 0; i < kPK; ++i) {
    const int t = a + i;
    if (!STEADY && t >= s1) break;
    const double v = (double)B.cv[i];
    if ((GM & kGE) && P.wE) {
      if (!STEADY && t == S.Le) {
        F.Ae = 0.0;
        F.be = S.sE;
      } else if (STEADY || t > S.Le) {
        if (!STEADY) F.Ae = mulc(F.Ae, P.eb);
        F.be = fma(P.eb, F.be, P.ek * v);
      }
    }
    if ((GM & kGR) && P.wR) {
      if (!STEADY && t == S.Lg) {
        F.Ag = F.Al = 0.0;
        F.bg = S.sG;
        F.bl = S.sL;
      } else if (STEADY || t > S.Lg) {
        const double di = B.d(i), gd = di > 0.0 ? di : 0.0;
        if (!STEADY) {
          F.Ag = mulc(F.Ag, P.br);
          F.Al = mulc(F.Al, P.br);
        }
        F.bg = (F.bg * P.pm1 + gd) * P.inv;
        F.bl = (F.bl * P.pm1 + (gd - di)) * P.inv;
      }
    }
    if ((GM & kGM) && P.wM) {
      if (!STEADY && t == S.Lm) {
        F.Af = F.As = 0.0;
        F.bf = S.sF;
        F.bs = S.sS;
      } else if (STEADY || t > S.Lm) {
        if (!STEADY) {
          F.Af = mulc(F.Af, P.fb);
          F.As = mulc(F.As, P.sb);
        }
        F.bf = fma(P.fb, F.bf, P.fk * v);
        F.bs = fma(P.sb, F.bs, P.sk * v);
      }
      if (!STEADY && t == S.Ls) {
        F.agf = F.ags = F.agg = 0.0;
        F.bgg = S.sSig;
      } else if (STEADY || t > S.Ls) {
        if (!STEADY) {
          F.agf = fma(P.gb, F.agf, P.gk * F.Af);
          F.ags = fma(P.gb, F.ags, -P.gk * F.As);
          F.agg = P.gb * F.agg;
        }
        F.bgg = fma(P.gb, F.bgg, P.gk * (F.bf - F.bs));
      }
    }
    if ((GM & kGA) && P.wA) {
      if (!STEADY && t == S.La) {
        F.Aa = 0.0;
        F.ba = S.sA;
      } else if (STEADY || t > S.La) {
        if (!STEADY) F.Aa = mulc(F.Aa, P.wb);
        F.ba = fma(P.wb, F.ba, P.wa * B.tr(i));
      }
    }
  }
  if (STEADY) pc_set_a(F, P.q[0]);
  return F;
}

// Whether every bar of a full segment lies past the seeds of the groups GM.
template <unsigned GM = kGAll>
__device__ __forceinline__ bool pipe_steady(const Seg& c, const PipeP& P, const PSeeds& S) {
  if (c.s1 - c.s0 != kPSeg) return false;
  int last = -1;
  if ((GM & kGE) && P.wE) last = max(last, S.Le);
  if ((GM & kGR) && P.wR) last = max(last, S.Lg);
  if ((GM & kGM) && P.wM) last = max(last, S.Ls);
  if ((GM & kGA) && P.wA) last = max(last, S.La);
  return c.s0 > last;
}

__device__ __forceinline__ Seg pipe_seg(const PipeP& P, int seg, int& row) {
  Seg c;
  row = seg / P.segsPerRow;
  c.seg = seg - row * P.segsPerRow;
  c.nseg = P.segsPerRow;
  c.gseg = seg;
  c.rowOff = (long long)row * P.n;
  c.n = P.n;
  c.s0 = c.seg * kPSeg;
  c.s1 = min(P.n, c.s0 + kPSeg);
  c.totalSegs = P.rows * P.segsPerRow;
  c.tree = false;
  c.epoch = 0;
  return c;
}

// Groups of 64 consecutive segments in a row (the carry's first level).
__device__ __forceinline__ int pipe_groups(const PipeP& P) {
  return (P.segsPerRow + kWave - 1) / kWave;
}

// One wave: the per-lane multipliers of steady waves (the general scan run
// on q[0] in every lane, recorded before each broadcast step).  The host
// computes the same table (pipe_qt_host) and keeps it on the device per
// parameter set; this kernel writes it where the host cannot upload it
// (under graph capture before the parameters were seen eagerly, and in the
// allocation-free entry point).
__global__ __launch_bounds__(64) void k_pipe_qt(PipeP P) {
  PC F;
  pc_id(F);
  pc_set_a(F, P.q[0]);
  pc_step<kRowShr1, 0xf, 0xf, kGAll>(F);
  pc_step<kRowShr2, 0xf, 0xf, kGAll>(F);
  pc_step<kRowShr4, 0xf, 0xf, kGAll>(F);
  pc_step<kRowShr8, 0xf, 0xf, kGAll>(F);
  const int lane = lane_id();
  auto put = [&](int k) {
    const double v[kQF] = {F.Ae, F.Ag, F.Al, F.Af, F.As, F.agf, F.ags, F.agg, F.Aa};
#pragma unroll
    for (int q = 0; q < kQF; ++q) P.qt[(k * kQF + q) * kWave + lane] = v[q];
  };
  put(0);
  pc_step<kRowBcast15, 0xa, 0xf, kGAll>(F);
  put(1);
}

using PTile = Tile<kPK, kPK>;

struct PLds {
  float c[PTile::PLEN], h[PTile::PLEN], l[PTile::PLEN];
};

template <class T>
__device__ __forceinline__ void pipe_stage(PLds& L, const PackIO& io, const Seg& c,
                                           const PipeP& P) {
  stage<T>(L.c, io.c, c);
  if (P.wA) {
    stage<T>(L.h, io.h, c);
    stage<T>(L.l, io.l, c);
  }
  wave_sync();
}

// The fields of group GM of a composite.
template <unsigned GM>
__device__ __forceinline__ void pc_store_group(double* o, const PC& F) {
  if constexpr (GM == kGE) {
    o[0] = F.Ae; o[1] = F.be;
  } else if constexpr (GM == kGR) {
    o[2] = F.Ag; o[3] = F.bg; o[4] = F.Al; o[5] = F.bl;
  } else if constexpr (GM == kGM) {
    o[6] = F.Af; o[7] = F.bf; o[8] = F.As; o[9] = F.bs;
    o[10] = F.agf; o[11] = F.ags; o[12] = F.agg; o[13] = F.bgg;
  } else {
    o[14] = F.Aa; o[15] = F.ba;
  }
}

// One group's segment composite (lane 63's inclusive scan) into agg; a group
// the pass does not compute stores the identity.
template <unsigned GM>
__device__ __forceinline__ void pipe_fold_group(const PipeP& P, const PSeeds& S, const PBars& B,
                                                int a, int s1, bool steady, double* o) {
  PC F;
  if (steady) {
    PQLane L;
    pq_lane<GM>(P, L);
    F = pipe_fold<true, GM>(P, S, B, a, s1);
    pc_scan_steady<GM>(F, P, L);
    pc_set_a(F, P.q[6]);
  } else {
    F = pipe_fold<false, GM>(P, S, B, a, s1);
    pc_scan<GM>(F);
  }
  if (lane_id() == kWave - 1) pc_store_group<GM>(o, F);
}

// The block's segment staged by its waves: close, then high and low for ATR.
template <class T>
__device__ __forceinline__ void pipe_stage_block(PLds& L, const PackIO& io, const Seg& c,
                                                 const PipeP& P, int w) {
  if (w == 0) stage<T>(L.c, io.c, c);
  else if (w == 1 && P.wA) stage<T>(L.h, io.h, c);
  else if (w == 2 && P.wA) stage<T>(L.l, io.l, c);
  __syncthreads();
}

template <unsigned GM>
__device__ __forceinline__ bool pipe_wants(const PipeP& P) {
  return GM == kGE ? P.wE : GM == kGR ? P.wR : GM == kGM ? P.wM : P.wA;
}

// One wave of the fold: group GM of the block's segment.
template <unsigned GM>
__device__ __forceinline__ void pipe_fold_wave(const PipeP& P, const PLds& L, const PackIO& io,
                                               const Seg& c, int f, double* o) {
  if (!pipe_wants<GM>(P)) {
    if (lane_id() == kWave - 1) {
      PC F;
      pc_id(F);
      pc_store_group<GM>(o, F);
    }
    return;
  }
  using T = PTile;
  const int a = c.s0 + lane_id() * kPK;
  PBars B;
  pipe_bars<T>(L.c, L.h, L.l, io, c, P, a, B);
  const PSeeds S = pipe_seeds<T, GM>(L.c, L.h, L.l, io, c, P, f);
  pipe_fold_group<GM>(P, S, B, a, c.s1, pipe_steady<GM>(c, P, S), o);
}

__device__ __forceinline__ PC pc_load(const double* g) {
  PC p;
  p.Ae = g[0]; p.be = g[1]; p.Ag = g[2]; p.bg = g[3]; p.Al = g[4]; p.bl = g[5];
  p.Af = g[6]; p.bf = g[7]; p.As = g[8]; p.bs = g[9]; p.agf = g[10]; p.ags = g[11];
  p.agg = g[12]; p.bgg = g[13]; p.Aa = g[14]; p.ba = g[15];
  return p;
}

__device__ __forceinline__ void pc_store(double* g, const PC& p) {
  const double v[16] = {p.Ae, p.be, p.Ag, p.bg, p.Al, p.bl, p.Af, p.bf,
                        p.As, p.bs, p.agf, p.ags, p.agg, p.bgg, p.Aa, p.ba};
#pragma unroll
  for (int q = 0; q < 16; ++q) g[q] = v[q];
}

// Lane j's composite in every lane (j uniform over the wave).
__device__ __forceinline__ PC pc_lane(const PC& p, int j) {
  PC r;
  r.Ae = bcast_d(p.Ae, j); r.be = bcast_d(p.be, j);
  r.Ag = bcast_d(p.Ag, j); r.bg = bcast_d(p.bg, j);
  r.Al = bcast_d(p.Al, j); r.bl = bcast_d(p.bl, j);
  r.Af = bcast_d(p.Af, j); r.bf = bcast_d(p.bf, j);
  r.As = bcast_d(p.As, j); r.bs = bcast_d(p.bs, j);
  r.agf = bcast_d(p.agf, j); r.ags = bcast_d(p.ags, j);
  r.agg = bcast_d(p.agg, j); r.bgg = bcast_d(p.bgg, j);
  r.Aa = bcast_d(p.Aa, j); r.ba = bcast_d(p.ba, j);
  return r;
}

__device__ __forceinline__ PC pc_lane63(const PC& p) { return pc_lane(p, kWave - 1); }


// One wave per group of 64 consecutive segments of a row: every segment's
// composite becomes the composite of the group's segments before it (in
// place), and the group's composite goes to gtot.
__global__ __launch_bounds__(kPW* kWave) void k_pipe_group(PipeP P, double* agg, double* gtot) {
  const int G = pipe_groups(P);
  const int gid = blockIdx.x * kPW + threadIdx.x / kWave;
  if (gid >= P.rows * G) return;
  const int row = gid / G, g = gid - row * G;
  const int j = g * kWave + lane_id();
  const bool on = j < P.segsPerRow;
  double* slot = agg + ((size_t)row * P.segsPerRow + j) * 16;
  PC F;
  if (on) F = pc_load(slot);
  else pc_id(F);
  pc_scan(F);
  const PC E = pc_prev_lane(F);
  if (on) pc_store(slot, E);
  if (lane_id() == kWave - 1) pc_store(gtot + (size_t)gid * 16, F);
}

// One wave per row: the state entering every group (7 doubles), 64 groups
// at a time.
__global__ __launch_bounds__(kPW* kWave) void k_pipe_rows(PipeP P, const double* gtot,
                                                          double* genter) {
  const int G = pipe_groups(P);
  const int row = blockIdx.x * kPW + threadIdx.x / kWave;
  if (row >= P.rows) return;
  PState y{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int g0 = 0; g0 < G; g0 += kWave) {
    const int g = g0 + lane_id();
    const size_t k = (size_t)row * G + g;
    PC F;
    if (g < G) F = pc_load(gtot + k * 16);
    else pc_id(F);
    pc_scan(F);
    if (g < G) {
      const PState v = pc_apply(pc_prev_lane(F), y);
      double* o = genter + k * 7;
      o[0] = v.e; o[1] = v.G; o[2] = v.L; o[3] = v.f; o[4] = v.s; o[5] = v.g; o[6] = v.a;
    }
    y = pc_apply(pc_lane63(F), y);
  }
}

// The state entering group g of a row, for one recurrence group, computed
// by the emit wave itself from the row's group composites: the lane scans
// and compositions of k_pipe_rows restricted to the groups before g, so the
// same doubles come out.  Used where the row scans are short (few segments
// in all), which saves the k_pipe_rows launch.
template <unsigned GM>
__device__ __forceinline__ PState pipe_group_enter(const PipeP& P, const double* gtot, int row,
                                                   int g) {
  const int G = pipe_groups(P);
  PState y{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int g0 = 0;; g0 += kWave) {
    const int j = g0 + lane_id();
    PC F;
    if (j < g) F = pc_load(gtot + ((size_t)row * G + j) * 16);
    else pc_id(F);
    pc_scan<GM>(F);
    if (g < g0 + kWave) return pc_apply(pc_lane(pc_prev_lane(F), g - g0), y);
    y = pc_apply(pc_lane63(F), y);
  }
}

__device__ __forceinline__ void pipe_put(float* dst, const Seg& c, int a, const float (&v)[kPK]) {
  if (!dst) return;
  float* p = dst + c.rowOff + a;
  if (a + kPK <= c.s1 && ((reinterpret_cast<size_t>(p) & 15) == 0)) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
    return;
  }
#pragma unroll
  for (int i = 0; i < kPK; ++i)
    if (a + i < c.s1) p[i] = v[i];
}

// Replay of one group for one lane from the state entering its bars; the
// group's outputs are stored together (whole 32-byte runs per lane).
template <bool STEADY, unsigned GM>
__device__ __forceinline__ void pipe_replay(const PipeP& P, const PSeeds& S, const PBars& B,
                                            const Seg& c, int a, PState y, const PackIO& io) {
  const float qn = __builtin_nanf("");
  float o0[kPK], o1[kPK], o2[kPK];
#pragma unroll
  for (int i = 0; i < kPK; ++i) {
    const int t = a + i;
    o0[i] = o1[i] = o2[i] = qn;
    if (!STEADY && t >= c.s1) continue;
    const double v = (double)B.cv[i];
    if constexpr (GM == kGE) {
      if (!STEADY && t == S.Le) y.e = S.sE;
      else if (STEADY || t > S.Le) y.e = fma(P.eb, y.e, P.ek * v);
      if (STEADY || t >= S.Le) o0[i] = (float)y.e;
    } else if constexpr (GM == kGR) {
      if (!STEADY && t == S.Lg) {
        y.G = S.sG;
        y.L = S.sL;
      } else if (STEADY || t > S.Lg) {
        const double d = B.d(i), gd = d > 0.0 ? d : 0.0;
        y.L *= P.pm1;
        y.G *= P.pm1;
        y.G += gd;
        y.L += gd - d;
        y.L *= P.inv;
        y.G *= P.inv;
      }
      if (STEADY || t >= S.Lg) {
        const double sm = y.G + y.L;
        o0[i] = (float)(sm > 0.0 ? 100.0 * (y.G / sm) : 50.0);
      }
    } else if constexpr (GM == kGM) {
      if (!STEADY && t == S.Lm) {
        y.f = S.sF;
        y.s = S.sS;
      } else if (STEADY || t > S.Lm) {
        y.f = fma(P.fb, y.f, P.fk * v);
        y.s = fma(P.sb, y.s, P.sk * v);
      }
      const double mv = y.f - y.s;
      if (!STEADY && t == S.Ls) y.g = S.sSig;
      else if (STEADY || t > S.Ls) y.g = fma(P.gb, y.g, P.gk * mv);
      if (STEADY || t >= S.Ls) {
        o0[i] = (float)mv;
        o1[i] = (float)y.g;
        o2[i] = (float)(mv - y.g);
      }
    } else {
      if (!STEADY && t == S.La) y.a = S.sA;
      else if (STEADY || t > S.La) y.a = fma(P.wb, y.a, P.wa * B.tr(i));
      if (STEADY || t >= S.La) o0[i] = (float)y.a;
    }
  }
  if constexpr (GM == kGE) {
    pipe_put(io.out[kOEma], c, a, o0);
  } else if constexpr (GM == kGR) {
    pipe_put(io.out[kORsi], c, a, o0);
  } else if constexpr (GM == kGM) {
    pipe_put(io.out[kOMacd], c, a, o0);
    pipe_put(io.out[kOMacdSig], c, a, o1);
    pipe_put(io.out[kOMacdHist], c, a, o2);
  } else {
    pipe_put(io.out[kOAtr], c, a, o0);
  }
}

// One group of the emit pass: the lane's composite, the wave scan seeded by
// the state entering the segment (lane 0), then the replay.
template <unsigned GM>
__device__ __forceinline__ void pipe_emit_scan(const PipeP& P, const PSeeds& S, const PBars& B,
                                               const Seg& c, int a, bool steady,
                                               const PQLane& Lq, PC F, const PState& y0,
                                               const PackIO& io) {
  const int lane = lane_id();
  // lane 0 starts from y0 (a composite with zero multipliers); the scans
  // below read only the offsets of the lanes they combine it with
  if (lane == 0) {
    PC C;
    pc_id(C);
    C.Ae = C.Ag = C.Al = C.Af = C.As = C.Aa = C.agf = C.ags = C.agg = 0.0;
    C.be = y0.e; C.bg = y0.G; C.bl = y0.L; C.bf = y0.f; C.bs = y0.s; C.bgg = y0.g; C.ba = y0.a;
    F = pc_after(C, F);
  }
  if (steady) pc_scan_steady<GM>(F, P, Lq);
  else pc_scan<GM>(F);
  PState y = y0;
  if constexpr (GM == kGE) {
    const double e = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.be);
    if (lane > 0) y.e = e;
  } else if constexpr (GM == kGR) {
    const double G = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.bg);
    const double L = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.bl);
    if (lane > 0) {
      y.G = G;
      y.L = L;
    }
  } else if constexpr (GM == kGM) {
    const double f = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.bf);
    const double sl = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.bs);
    const double g = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.bgg);
    if (lane > 0) {
      y.f = f;
      y.s = sl;
      y.g = g;
    }
  } else {
    const double v = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, F.ba);
    if (lane > 0) y.a = v;
  }
  if (steady) pipe_replay<true, GM>(P, S, B, c, a, y, io);
  else pipe_replay<false, GM>(P, S, B, c, a, y, io);
}

template <unsigned GM>
__device__ __forceinline__ void pipe_emit_group(const PipeP& P, const PSeeds& S, const PBars& B,
                                                const Seg& c, int a, bool steady,
                                                const PState& y0, const PackIO& io) {
  PQLane Lq;
  if (steady) pq_lane<GM>(P, Lq);
  const PC F =
      steady ? pipe_fold<true, GM>(P, S, B, a, c.s1) : pipe_fold<false, GM>(P, S, B, a, c.s1);
  pipe_emit_scan<GM>(P, S, B, c, a, steady, Lq, F, y0, io);
}

union PipeFoldLds {
  PLds rec;
  ColSplitLds<ColTile<64>::E> win;
};

// One segment per workgroup: a recurrence segment (one wave per recurrence
// group) or a window segment of the pass's window part (one family group
// per wave).  With a window part, the blocks come in runs of 24: the 8
// recurrence segments u = 8m .. 8m+7 and then, in the same order, the two
// 256-bar window segments over each one's bars, so that the three blocks
// of segment u run on one XCD (workgroup b on XCD b mod 8) in consecutive
// slots and the later ones read the bars through its L2.
constexpr int kPipeRun = 8;
__global__ __launch_bounds__(kPW* kWave) void k_pipe_fold(PipeP P, PackP prm, PackIO io,
                                                          double* agg, Geometry gw, int nbW) {
  __shared__ __align__(16) PipeFoldLds sm;
  const int w = threadIdx.x / kWave;
  const int segs = P.rows * P.segsPerRow;
  int seg = (int)blockIdx.x;
  if (nbW > 0) {
    const int m = (int)blockIdx.x / (3 * kPipeRun), in = (int)blockIdx.x % (3 * kPipeRun);
    const int kind = in / kPipeRun, u = m * kPipeRun + in % kPipeRun;
    if (u >= segs) return;
    if (kind > 0) {
      const int row = u / P.segsPerRow, j = u - row * P.segsPerRow;
      const int wj = 2 * j + (kind - 1);
      if (wj >= gw.segsPerRow) return;
      Seg c;
      c.tree = false;
      c.epoch = 0;
      if (resolve(gw, ColTile<64>::SEG, row * gw.segsPerRow + wj, c))
        pack_windows_split<ColTile<64>, kPWinIL>(c, sm.win, prm, io, w);
      return;
    }
    seg = u;
  }
  if (seg >= segs) return;
  int row;
  const Seg c = pipe_seg(P, seg, row);
  PLds& L = sm.rec;
  const float c0 = io.c[c.rowOff];  // in flight with the staging loads
  pipe_stage_block<PTile>(L, io, c, P, w);
  const int f = __builtin_isnan(c0) ? first_valid(c, io.c) : 0;
  double* o = agg + (size_t)seg * 16;
  if (w == 0) pipe_fold_wave<kGE>(P, L, io, c, f, o);
  else if (w == 1) pipe_fold_wave<kGR>(P, L, io, c, f, o);
  else if (w == 2) pipe_fold_wave<kGM>(P, L, io, c, f, o);
  else pipe_fold_wave<kGA>(P, L, io, c, f, o);
}

// Passes with one recurrence group and no window part (the RSI, MACD and
// ATR family calls): four segments per workgroup, one per wave, so no wave
// of the block idles.  The other groups' fields of the composite are the
// identity.
template <unsigned GM>
__global__ __launch_bounds__(kPW* kWave) void k_pipe_fold1(PipeP P, PackIO io, double* agg) {
  __shared__ __align__(16) PLds sm[kPW];
  const int w = threadIdx.x / kWave;
  const int seg = blockIdx.x * kPW + w;
  if (seg >= P.rows * P.segsPerRow) return;
  int row;
  const Seg c = pipe_seg(P, seg, row);
  const float c0 = io.c[c.rowOff];  // in flight with the staging loads
  pipe_stage<PTile>(sm[w], io, c, P);
  const int f = __builtin_isnan(c0) ? first_valid(c, io.c) : 0;
  double* o = agg + (size_t)seg * 16;
  if (lane_id() == kWave - 1) {
    PC I;
    pc_id(I);
    if constexpr (GM != kGE) pc_store_group<kGE>(o, I);
    if constexpr (GM != kGR) pc_store_group<kGR>(o, I);
    if constexpr (GM != kGM) pc_store_group<kGM>(o, I);
    if constexpr (GM != kGA) pc_store_group<kGA>(o, I);
  }
  pipe_fold_wave<GM>(P, sm[w], io, c, f, o);
}

// One wave of the emit pass: group GM of the block's segment.
template <unsigned GM>
__device__ __forceinline__ void pipe_emit_wave(const PipeP& P, const PLds& L, const PackIO& io,
                                               const Seg& c, int row, int seg, int f,
                                               const double* agg, const double* gtot,
                                               const double* genter) {
  if (!pipe_wants<GM>(P)) return;
  using T = PTile;
  const int a = c.s0 + lane_id() * kPK;
  PBars B;
  pipe_bars<T>(L.c, L.h, L.l, io, c, P, a, B);
  const PSeeds S = pipe_seeds<T, GM>(L.c, L.h, L.l, io, c, P, f);
  // state entering the segment: its group's entering state (from k_pipe_rows,
  // or scanned here when genter is null), then the group's segments before it
  PState ge;
  if (genter) {
    const double* g = genter + ((size_t)row * pipe_groups(P) + c.seg / kWave) * 7;
    ge = PState{g[0], g[1], g[2], g[3], g[4], g[5], g[6]};
  } else {
    ge = pipe_group_enter<GM>(P, gtot, row, c.seg / kWave);
  }
  const PState y0 = pc_apply(pc_load(agg + (size_t)seg * 16), ge);
  pipe_emit_group<GM>(P, S, B, c, a, pipe_steady<GM>(c, P, S), y0, io);
}

// One workgroup per segment, one wave per recurrence group: the state
// entering the segment, then the replay of its bars.
__global__ __launch_bounds__(kPW* kWave) void k_pipe_emit(PipeP P, PackIO io, const double* agg,
                                                          const double* gtot,
                                                          const double* genter) {
  __shared__ __align__(16) PLds L;
  const int seg = blockIdx.x, w = threadIdx.x / kWave;
  if (seg >= P.rows * P.segsPerRow) return;
  int row;
  const Seg c = pipe_seg(P, seg, row);
  const float c0 = io.c[c.rowOff];  // in flight with the staging loads
  pipe_stage_block<PTile>(L, io, c, P, w);
  const int f = __builtin_isnan(c0) ? first_valid(c, io.c) : 0;
  if (w == 0) pipe_emit_wave<kGE>(P, L, io, c, row, seg, f, agg, gtot, genter);
  else if (w == 1) pipe_emit_wave<kGR>(P, L, io, c, row, seg, f, agg, gtot, genter);
  else if (w == 2) pipe_emit_wave<kGM>(P, L, io, c, row, seg, f, agg, gtot, genter);
  else pipe_emit_wave<kGA>(P, L, io, c, row, seg, f, agg, gtot, genter);
}

// The emit pass of k_pipe_fold1's passes: one segment per wave.
template <unsigned GM>
__global__ __launch_bounds__(kPW* kWave) void k_pipe_emit1(PipeP P, PackIO io, const double* agg,
                                                           const double* gtot,
                                                           const double* genter) {
  __shared__ __align__(16) PLds sm[kPW];
  const int w = threadIdx.x / kWave;
  const int seg = blockIdx.x * kPW + w;
  if (seg >= P.rows * P.segsPerRow) return;
  int row;
  const Seg c = pipe_seg(P, seg, row);
  const float c0 = io.c[c.rowOff];  // in flight with the staging loads
  pipe_stage<PTile>(sm[w], io, c, P);
  const int f = __builtin_isnan(c0) ? first_valid(c, io.c) : 0;
  pipe_emit_wave<GM>(P, sm[w], io, c, row, seg, f, agg, gtot, genter);
}

// ------------------------------------------------------- one-launch pass --
// Passes of one recurrence group and no window part (the EMA, RSI, MACD
// and ATR family calls) on few rows run as one launch.  A tile is kPW
// consecutive segments of a row (2048 bars) and one workgroup, a wave per
// segment: the waves fold their bars, the workgroup publishes the tile's
// composite, one wave obtains the state entering the tile from the
// composites of every tile before it in the row (up to kSoloM records per
// lane, composed in the lane, then across the wave), and the waves replay
// their segments from it.  Every wait is on a tile of a lower index in its
// row, whose workgroup starts earlier (workgroup u holds tile u), so the
// launch always completes; each wait is bounded in time all the same
// (kSoloWaitTicks, then the row's stall word and NaN states).  (Several
// groups per pass or a window part multiply the workgroups that wait at
// once beyond what stays resident, and those passes keep the two passes.)
//
// Records carry a tag (row, group and the launch epoch of that row and
// group).  The epoch word is read by each tile and bumped by the row's last
// tile once its look-back has seen every other tile's record, which proves
// that every tile of the row has read it; the next launch on the workspace
// (stream-ordered after this one) so tags its records anew, with no reset
// launch and no end-of-launch counter.  The workspace part holds only these
// records, so no other kernel rewrites an epoch word.
constexpr int kSoloMaxRows = 1024;
constexpr long long kSoloMaxSegs = 1ll << 14;
constexpr int kSoloM = 8;                          // records per lane of a look-back
constexpr int kSoloMaxTiles = kSoloM * kWave + 1;  // tiles per row
constexpr int kSR = 9;  // doubles per record: tag, then a group's composite fields
constexpr unsigned long long kSoloWaitTicks = 20000000ull;  // 0.2 s of the 100 MHz clock
constexpr int kSoloPart = 2;

struct SoloWs {
  unsigned* epoch;  // [kSoloMaxRows * 4]: row-major, one word per group
  unsigned* stall;  // [kSoloMaxRows * 4]: epoch | 2^31 once a wait of that launch gave up
  double* rec;      // [rows * tilesPerRow * 4 * kSR]
};
template <unsigned GM>
constexpr int pipe_gi() {
  return GM == kGE ? 0 : GM == kGR ? 1 : GM == kGM ? 2 : 3;
}

__device__ __forceinline__ unsigned long long solo_tag(int row, int gi, unsigned e) {
  return (1ull << 62) | ((unsigned long long)(row * 4 + gi) << 32) | e;
}

// Group GM's fields of lane j's composite in every lane (j uniform); the
// other fields are the identity.
template <unsigned GM>
__device__ __forceinline__ PC pc_lane_g(const PC& p, int j) {
  PC r;
  pc_id(r);
  if constexpr (GM == kGE) {
    r.Ae = bcast_d(p.Ae, j); r.be = bcast_d(p.be, j);
  } else if constexpr (GM == kGR) {
    r.Ag = bcast_d(p.Ag, j); r.bg = bcast_d(p.bg, j);
    r.Al = bcast_d(p.Al, j); r.bl = bcast_d(p.bl, j);
  } else if constexpr (GM == kGM) {
    r.Af = bcast_d(p.Af, j); r.bf = bcast_d(p.bf, j);
    r.As = bcast_d(p.As, j); r.bs = bcast_d(p.bs, j);
    r.agf = bcast_d(p.agf, j); r.ags = bcast_d(p.ags, j);
    r.agg = bcast_d(p.agg, j); r.bgg = bcast_d(p.bgg, j);
  } else {
    r.Aa = bcast_d(p.Aa, j); r.ba = bcast_d(p.ba, j);
  }
  return r;
}

// Publish group GM's fields of the (wave-uniform) composite C: lane 0
// stores the fields, drains its stores, then stores the tag.
template <unsigned GM>
__device__ __forceinline__ void solo_put(double* r, const PC& C, unsigned long long tag) {
  if (lane_id() != 0) return;
  if constexpr (GM == kGE) {
    st_rel_d(r + 1, C.Ae); st_rel_d(r + 2, C.be);
  } else if constexpr (GM == kGR) {
    st_rel_d(r + 1, C.Ag); st_rel_d(r + 2, C.bg); st_rel_d(r + 3, C.Al); st_rel_d(r + 4, C.bl);
  } else if constexpr (GM == kGM) {
    st_rel_d(r + 1, C.Af); st_rel_d(r + 2, C.bf); st_rel_d(r + 3, C.As); st_rel_d(r + 4, C.bs);
    st_rel_d(r + 5, C.agf); st_rel_d(r + 6, C.ags); st_rel_d(r + 7, C.agg);
    st_rel_d(r + 8, C.bgg);
  } else {
    st_rel_d(r + 1, C.Aa); st_rel_d(r + 2, C.ba);
  }
  vmem_order();
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(r), tag, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

template <unsigned GM>
__device__ __forceinline__ void solo_get(const double* r, PC& F) {
  if constexpr (GM == kGE) {
    F.Ae = ld_rel_d(r + 1); F.be = ld_rel_d(r + 2);
  } else if constexpr (GM == kGR) {
    F.Ag = ld_rel_d(r + 1); F.bg = ld_rel_d(r + 2); F.Al = ld_rel_d(r + 3); F.bl = ld_rel_d(r + 4);
  } else if constexpr (GM == kGM) {
    F.Af = ld_rel_d(r + 1); F.bf = ld_rel_d(r + 2); F.As = ld_rel_d(r + 3); F.bs = ld_rel_d(r + 4);
    F.agf = ld_rel_d(r + 5); F.ags = ld_rel_d(r + 6); F.agg = ld_rel_d(r + 7);
    F.bgg = ld_rel_d(r + 8);
  } else {
    F.Aa = ld_rel_d(r + 1); F.ba = ld_rel_d(r + 2);
  }
}

// The composite of records 0 .. cnt-1 (1 <= cnt <= kSoloM * 64) at `base`,
// `stride` doubles apart, uniform: lane l waits for records l*m .. l*m+m-1
// (m = ceil(cnt / 64)) to carry `tag`, composes them in order, and the
// wave scan composes the lanes.  A wait longer than kSoloWaitTicks, or one
// of a launch whose row already gave up, clears `ok` (the caller then
// carries NaN).
template <unsigned GM>
__device__ __forceinline__ PC solo_gather(const double* base, size_t stride, int cnt,
                                          unsigned long long tag, unsigned* stall, bool& ok) {
  const int lane = lane_id();
  const int m = (cnt + kWave - 1) / kWave;
  const int k0 = min(cnt, lane * m), n = min(cnt, k0 + m) - k0;
  // every load is issued unconditionally (lanes past the count reread
  // record 0), so the lane's loads are in flight together
  const double* rk[kSoloM];
#pragma unroll
  for (int k = 0; k < kSoloM; ++k) rk[k] = base + (size_t)(k < n ? k0 + k : 0) * stride;
  if (ok) {
    const unsigned stallTag = (unsigned)tag | 0x80000000u;
    const unsigned long long t0 = wall_clock64();
    for (int polls = 1;; ++polls) {
      unsigned long long fl[kSoloM];
#pragma unroll
      for (int k = 0; k < kSoloM; ++k) fl[k] = ld_flag(rk[k]);
      bool rdy = true;
#pragma unroll
      for (int k = 0; k < kSoloM; ++k) rdy = rdy & (k >= n || fl[k] == tag);
      if (__ballot(!rdy) == 0ull) break;
      if ((polls & 63) == 0 &&
          (wall_clock64() - t0 > kSoloWaitTicks ||
           __hip_atomic_load(stall, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == stallTag)) {
        ok = false;
        if (lane == 0)
          __hip_atomic_store(stall, stallTag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  vmem_order();
  // the fields in two halves of kSoloM / 2 records (the second only when a
  // lane holds more), each half's loads in flight together
  PC acc;
  pc_id(acc);
#pragma unroll
  for (int h = 0; h < kSoloM; h += kSoloM / 2) {
    if (h > 0 && m <= h) break;
    PC r[kSoloM / 2];
#pragma unroll
    for (int k = 0; k < kSoloM / 2; ++k) {
      pc_id(r[k]);
      solo_get<GM>(rk[h + k], r[k]);
    }
#pragma unroll
    for (int k = 0; k < kSoloM / 2; ++k)
      if (ok && h + k < n) acc = pc_after(acc, r[k]);
  }
  pc_scan<GM>(acc);
  return pc_lane_g<GM>(acc, kWave - 1);
}

struct SoloTileLds {
  PLds seg[kPW];
  double comp[kPW][16];  // the segments' composites
  double ent[kPW][7];    // the states entering the segments
};

// Group GM of tile t of a row: one wave per segment.
template <unsigned GM>
__device__ __forceinline__ void solo_tile(const PipeP& P, SoloTileLds& L, const PackIO& io,
                                          const SoloWs& W, int row, int t, int tilesPerRow) {
  using T = PTile;
  constexpr int gi = pipe_gi<GM>();
  const int w = threadIdx.x / kWave, lane = lane_id();
  const int sr = t * kPW + w;  // segment index in the row
  const bool on = sr < P.segsPerRow;
  int rowOf;
  const Seg c = pipe_seg(P, row * P.segsPerRow + min(sr, P.segsPerRow - 1), rowOf);
  unsigned* ep = W.epoch + row * 4 + gi;
  unsigned eRaw = 0;
  if (w == 0 && lane == 0) eRaw = __hip_atomic_load(ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const int a = c.s0 + lane * kPK;
  PBars B;
  PSeeds S;
  PQLane Lq;
  PC F;
  bool steady = false;
  if (on) {
    PLds& Ls = L.seg[w];
    const float c0 = io.c[c.rowOff];  // in flight with the staging loads
    stage<T>(Ls.c, io.c, c);
    if (GM == kGA) {
      stage<T>(Ls.h, io.h, c);
      stage<T>(Ls.l, io.l, c);
    }
    wave_sync();
    const int f = __builtin_isnan(c0) ? first_valid(c, io.c) : 0;
    pipe_bars<T>(Ls.c, Ls.h, Ls.l, io, c, P, a, B);
    S = pipe_seeds<T, GM>(Ls.c, Ls.h, Ls.l, io, c, P, f);
    steady = pipe_steady<GM>(c, P, S);
    if (steady) pq_lane<GM>(P, Lq);
    F = steady ? pipe_fold<true, GM>(P, S, B, a, c.s1) : pipe_fold<false, GM>(P, S, B, a, c.s1);
    PC Fs = F;
    if (steady) {
      pc_scan_steady<GM>(Fs, P, Lq);
      pc_set_a(Fs, P.q[6]);
    } else {
      pc_scan<GM>(Fs);
    }
    const PC C = pc_lane_g<GM>(Fs, kWave - 1);
    if (lane == 0) pc_store(L.comp[w], C);
  } else if (lane == 0) {
    PC I;
    pc_id(I);
    pc_store(L.comp[w], I);
  }
  __syncthreads();
  if (w == 0) {
    PC Ct = pc_load(L.comp[0]);
#pragma unroll
    for (int k = 1; k < kPW; ++k) Ct = pc_after(Ct, pc_load(L.comp[k]));
    const unsigned e = (unsigned)__builtin_amdgcn_readfirstlane((int)eRaw);
    const unsigned long long tag = solo_tag(row, gi, e);
    constexpr size_t kStride = 4 * kSR;
    const double* base = W.rec + ((size_t)row * tilesPerRow * 4 + gi) * kSR;
    if (t + 1 < tilesPerRow) solo_put<GM>(const_cast<double*>(base) + (size_t)t * kStride, Ct, tag);
    bool ok = true;
    PC Cp;
    pc_id(Cp);
    if (t > 0) Cp = solo_gather<GM>(base, kStride, t, tag, W.stall + row * 4 + gi, ok);
    // the row's last tile has seen every tile of the row publish, so all of
    // them have read the epoch: the next launch may use the next one
    if (t + 1 == tilesPerRow && lane == 0)
      __hip_atomic_store(ep, e + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    PState y = pc_apply(Cp, PState{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
    if (!ok) {
      const double qn = qnan_d();
      y = PState{qn, qn, qn, qn, qn, qn, qn};
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < kPW; ++k) {
        double* o = L.ent[k];
        o[0] = y.e; o[1] = y.G; o[2] = y.L; o[3] = y.f; o[4] = y.s; o[5] = y.g; o[6] = y.a;
        y = pc_apply(pc_load(L.comp[k]), y);
      }
    }
  }
  __syncthreads();
  if (on) {
    const double* g = L.ent[w];
    const PState y0{g[0], g[1], g[2], g[3], g[4], g[5], g[6]};
    pipe_emit_scan<GM>(P, S, B, c, a, steady, Lq, F, y0, io);
  }
}

// The one-launch pass of group GM: workgroup u holds tile u (rows in order).
template <unsigned GM>
__global__ __launch_bounds__(kPW* kWave) void k_pipe_solo(PipeP P, PackIO io, SoloWs W,
                                                          int tilesPerRow, int tiles) {
  __shared__ __align__(16) SoloTileLds sm;
  const int u = (int)blockIdx.x;
  if (u >= tiles) return;
  const int row = u / tilesPerRow, t = u - row * tilesPerRow;
  solo_tile<GM>(P, sm, io, W, row, t, tilesPerRow);
}

// Whether the two passes serve this pass: dense rows of at least kPMinLen
// bars, recurrences among EMA, RSI, MACD, ATR with periods of 2 or more (the
// one-bar contracts and DEMA, TEMA, ADX keep the look-back kernels), and
// window look-backs within the window blocks' 64-bar halo.
bool pipe_serves(int len, const PackP& p) {
  if (t_ragged || len < kPMinLen) return false;
  if (p.mask & kPUnsupported) return false;
  if (!(p.mask & (kCloseRecMask | kHlcRecMask))) return false;
  if (window_halo(p) > 64) return false;
  const bool wE = p.mask & (1u << kOEma), wR = p.mask & (1u << kORsi);
  const bool wM = p.mask & ((1u << kOMacd) | (1u << kOMacdSig) | (1u << kOMacdHist));
  const bool wA = p.mask & (1u << kOAtr);
  if ((wE && p.emaP < 2) || (wR && p.rsiP < 2) || (wA && p.atrP < 2)) return false;
  if (wM && (p.mFast < 2 || p.mSlow < 2 || p.mSig < 2)) return false;
  return true;
}

// q[0]: the multipliers of one lane's 8 steady bars, in pipe_fold's order;
// q[s + 1] = q[s] after q[s], in pc_step's order (IEEE double on both sides,
// so they equal what the general scan computes).
static double mulc_h(double a, double m) {
  const double r = a * m;
  return (r == 0.0 && a != 0.0 && m != 0.0) ? 2.2250738585072014e-308 : r;
}
static void pipe_steady_q(PipeP& P) {
  double Ae = 1.0, Ag = 1.0, Al = 1.0, Af = 1.0, As = 1.0, agf = 0.0, ags = 0.0, agg = 1.0,
         Aa = 1.0;
  for (int i = 0; i < kPK; ++i) {
    Ae = mulc_h(Ae, P.eb);
    Ag = mulc_h(Ag, P.br);
    Al = mulc_h(Al, P.br);
    Af = mulc_h(Af, P.fb);
    As = mulc_h(As, P.sb);
    agf = std::fma(P.gb, agf, P.gk * Af);
    ags = std::fma(P.gb, ags, -P.gk * As);
    agg = P.gb * agg;
    Aa = mulc_h(Aa, P.wb);
  }
  double* q = P.q[0];
  q[0] = Ae; q[1] = Ag; q[2] = Al; q[3] = Af; q[4] = As; q[5] = agf; q[6] = ags; q[7] = agg;
  q[8] = Aa;
  for (int s = 0; s < 6; ++s) {
    const double* a = P.q[s];
    double* r = P.q[s + 1];
    r[0] = mulc_h(a[0], a[0]);
    r[1] = mulc_h(a[1], a[1]);
    r[2] = mulc_h(a[2], a[2]);
    r[5] = std::fma(a[5], a[3], a[7] * a[5]);
    r[6] = std::fma(a[6], a[4], a[7] * a[6]);
    r[7] = a[7] * a[7];
    r[3] = mulc_h(a[3], a[3]);
    r[4] = mulc_h(a[4], a[4]);
    r[8] = mulc_h(a[8], a[8]);
  }
}

// The table k_pipe_qt writes, on the host: every lane starts from q[0]
// and takes the same DPP steps (row_shr 1, 2, 4, 8, then row_bcast 15 on
// rows 1 and 3; a lane without a source reads the identity) with the same
// products, fmas and clamps in the same order, so the doubles are the
// kernel's.
static void pipe_qt_host(const PipeP& P, double* out) {
  struct Q {
    double Ae, Ag, Al, Af, As, agf, ags, agg, Aa;
  };
  const double* q = P.q[0];
  Q v[kWave];
  for (int l = 0; l < kWave; ++l) v[l] = Q{q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7], q[8]};
  const Q id{1.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 1.0, 1.0};
  auto step = [&](auto src) {
    Q nv[kWave];
    for (int l = 0; l < kWave; ++l) {
      const int j = src(l);
      const Q o = j < 0 ? id : v[j];
      Q f = v[l];
      f.Ae = mulc_h(f.Ae, o.Ae);
      f.Ag = mulc_h(f.Ag, o.Ag);
      f.Al = mulc_h(f.Al, o.Al);
      const double agf = std::fma(f.agf, o.Af, f.agg * o.agf);
      const double ags = std::fma(f.ags, o.As, f.agg * o.ags);
      const double agg = f.agg * o.agg;
      f.agf = agf;
      f.ags = ags;
      f.agg = agg;
      f.Af = mulc_h(f.Af, o.Af);
      f.As = mulc_h(f.As, o.As);
      f.Aa = mulc_h(f.Aa, o.Aa);
      nv[l] = f;
    }
    for (int l = 0; l < kWave; ++l) v[l] = nv[l];
  };
  for (int d = 1; d <= 8; d *= 2) step([d](int l) { return (l % 16) >= d ? l - d : -1; });
  auto put = [&](int k) {
    for (int l = 0; l < kWave; ++l) {
      const double f[kQF] = {v[l].Ae, v[l].Ag, v[l].Al, v[l].Af, v[l].As,
                             v[l].agf, v[l].ags, v[l].agg, v[l].Aa};
      for (int i = 0; i < kQF; ++i) out[(k * kQF + i) * kWave + l] = f[i];
    }
  };
  put(0);
  step([](int l) {
    const int r = l / 16;
    return r == 1 ? 15 : r == 3 ? 47 : -1;
  });
  put(1);
}

// Device copies of the multiplier table per (device, q[0]), uploaded once
// outside capture; null where that cannot be done here (under capture,
// inside the allocation-free entry point, or past the cache's bound), and
// the caller then runs k_pipe_qt.  Entries live until the process ends:
// captured graphs may hold their addresses.
static ctStatus_t pipe_qt_cached(const PipeP& P, hipStream_t s, const double** out) {
  *out = nullptr;
  if (t_ws_fixed) return CT_STATUS_SUCCESS;
  int dev = 0;
  TCA_HIP(hipGetDevice(&dev));
  std::array<unsigned long long, kQF + 1> key{};
  key[0] = (unsigned long long)dev;
  std::memcpy(key.data() + 1, P.q[0], sizeof(double) * kQF);
  static std::mutex mu;
  static std::map<std::array<unsigned long long, kQF + 1>, double*> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(key);
  if (it != cache.end()) {
    *out = it->second;
    return CT_STATUS_SUCCESS;
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (s != nullptr) TCA_HIP(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone || cache.size() >= 256) return CT_STATUS_SUCCESS;
  std::vector<double> host(2 * kQF * kWave);
  pipe_qt_host(P, host.data());
  double* d = nullptr;
  const size_t bytes = host.size() * sizeof(double);
  TCA_HIP(hipMalloc(&d, bytes));
  hipError_t e = hipMemcpyAsync(d, host.data(), bytes, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "pipe multiplier table upload");
  }
  cache[key] = d;
  *out = d;
  return CT_STATUS_SUCCESS;
}

// The recurrence group of a pass of one group and no window part, else 0.
static unsigned solo_group(const PipeP& P, const PackP& prm) {
  if (prm.mask & kWinMask) return 0u;
  const unsigned g = (P.wE ? kGE : 0u) | (P.wR ? kGR : 0u) | (P.wM ? kGM : 0u) | (P.wA ? kGA : 0u);
  return (g == kGE || g == kGR || g == kGM || g == kGA) ? g : 0u;
}

// Whether the one-launch form serves a pass: one recurrence group, no window
// part, at most kSoloMaxRows rows, kSoloMaxSegs segments and kSoloMaxTiles
// tiles per row.
static bool solo_serves(const PipeP& P, const PackP& prm) {
  return solo_group(P, prm) != 0u && P.rows <= kSoloMaxRows &&
         (long long)P.rows * P.segsPerRow <= kSoloMaxSegs &&
         div_up(P.segsPerRow, kPW) <= kSoloMaxTiles;
}

// The one-launch form.  Its workspace part: the epoch and stall words of
// every row and group and the steady multiplier table (when not cached) at
// fixed offsets, then the tile records, so records of an earlier launch of
// another shape never overlay a word read as a tag.
static ctStatus_t launch_pipe_solo(PipeP& P, const PackP& prm, const PackIO& io,
                                   hipStream_t s) {
  const int tilesPerRow = div_up(P.segsPerRow, kPW);
  const int tiles = tilesPerRow * P.rows;
  constexpr size_t kHdr = 2 * (size_t)kSoloMaxRows * 4 * sizeof(unsigned);
  constexpr size_t kQt = 2 * kQF * kWave * sizeof(double);
  const size_t bytes = kHdr + kQt + (size_t)tiles * 4 * kSR * sizeof(double);
  Ws ws;
  ctStatus_t st = workspace_for(s, bytes, &ws, kSoloPart);
  if (st != CT_STATUS_SUCCESS) return st;
  char* base = reinterpret_cast<char*>(ws.area);
  SoloWs W;
  W.epoch = reinterpret_cast<unsigned*>(base);
  W.stall = W.epoch + kSoloMaxRows * 4;
  W.rec = reinterpret_cast<double*>(base + kHdr + kQt);
  const double* qt = nullptr;
  st = pipe_qt_cached(P, s, &qt);
  if (st != CT_STATUS_SUCCESS) return st;
  if (qt) {
    P.qt = const_cast<double*>(qt);
  } else {
    P.qt = reinterpret_cast<double*>(base + kHdr);
    KTime kt(s, "k_pipe_qt");
    hipLaunchKernelGGL(k_pipe_qt, dim3(1), dim3(kWave), 0, s, P);
    TCA_HIP(hipGetLastError());
  }
  const unsigned g = solo_group(P, prm);
  const dim3 grid((unsigned)tiles), blk(kPW * kWave);
  KTime kt(s, "k_pipe_solo");
  if (g == kGE) hipLaunchKernelGGL(k_pipe_solo<kGE>, grid, blk, 0, s, P, io, W, tilesPerRow, tiles);
  else if (g == kGR) hipLaunchKernelGGL(k_pipe_solo<kGR>, grid, blk, 0, s, P, io, W, tilesPerRow, tiles);
  else if (g == kGM) hipLaunchKernelGGL(k_pipe_solo<kGM>, grid, blk, 0, s, P, io, W, tilesPerRow, tiles);
  else hipLaunchKernelGGL(k_pipe_solo<kGA>, grid, blk, 0, s, P, io, W, tilesPerRow, tiles);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}


ctStatus_t launch_pipe(int len, int rows, hipStream_t s, const PackP& prm, const PackIO& io) {
  PipeP P{};
  P.n = len;
  P.rows = rows;
  P.segsPerRow = div_up(len, kPSeg);
  P.wE = (prm.mask & (1u << kOEma)) ? 1 : 0;
  P.wR = (prm.mask & (1u << kORsi)) ? 1 : 0;
  P.wM = (prm.mask & ((1u << kOMacd) | (1u << kOMacdSig) | (1u << kOMacdHist))) ? 1 : 0;
  P.wA = (prm.mask & (1u << kOAtr)) ? 1 : 0;
  P.pe = prm.emaP;
  P.pr = prm.rsiP;
  P.fast = prm.mFast;
  P.slow = prm.mSlow;
  P.sig = prm.mSig;
  P.pa = prm.atrP;
  auto coef = [](int p, double& b, double& k) {
    double bb = (double)(p - 1) / (double)(p + 1);
    const double kk = 1.0 - bb;
    if (bb < 0.5) bb = 1.0 - kk;
    b = bb;
    k = kk;
  };
  coef(P.pe, P.eb, P.ek);
  coef(P.fast, P.fb, P.fk);
  coef(P.slow, P.sb, P.sk);
  coef(P.sig, P.gb, P.gk);
  P.inv = 1.0 / (double)P.pr;
  P.pm1 = (double)(P.pr - 1);
  P.br = P.pm1 * P.inv;
  P.wb = (double)(P.pa - 1) / (double)P.pa;
  P.wa = 1.0 - P.wb;
  const long long segs = (long long)P.segsPerRow * rows;
  if (segs > (1ll << 30)) return fail(CT_STATUS_UNSUPPORTED, "pack: batch too large");
  pipe_steady_q(P);
  if (solo_serves(P, prm)) return launch_pipe_solo(P, prm, io, s);
  const long long groups = (long long)div_up(P.segsPerRow, kWave) * rows;
  const size_t bytes =
      ((size_t)segs * 16 + (size_t)groups * (16 + 7) + 2 * kQF * kWave) * sizeof(double);
  Ws ws;
  ctStatus_t st = workspace_for(s, bytes, &ws, 0);
  if (st != CT_STATUS_SUCCESS) return st;
  double* agg = ws.area;
  double* gtot = agg + (size_t)segs * 16;
  double* genter = gtot + (size_t)groups * 16;
  const double* qt = nullptr;
  st = pipe_qt_cached(P, s, &qt);
  if (st != CT_STATUS_SUCCESS) return st;
  if (qt) {
    P.qt = const_cast<double*>(qt);
  } else {
    P.qt = genter + (size_t)groups * 7;
    KTime kt(s, "k_pipe_qt");
    hipLaunchKernelGGL(k_pipe_qt, dim3(1), dim3(kWave), 0, s, P);
    TCA_HIP(hipGetLastError());
  }
  Geometry gw = dense_geometry(len, rows, ColTile<64>::SEG);
  const int nbW = (prm.mask & kWinMask) ? gw.totalSegs : 0;
  // one recurrence group and no window part: the one-segment-per-wave kernels
  const unsigned one = nbW ? 0u
                           : (P.wE ? kGE : 0u) | (P.wR ? kGR : 0u) | (P.wM ? kGM : 0u) |
                                 (P.wA ? kGA : 0u);
  const dim3 g1(div_up(segs, kPW)), blk(kPW * kWave);
  {
    KTime kt(s, "k_pipe_fold");
    if (one == kGE) hipLaunchKernelGGL(k_pipe_fold1<kGE>, g1, blk, 0, s, P, io, agg);
    else if (one == kGR) hipLaunchKernelGGL(k_pipe_fold1<kGR>, g1, blk, 0, s, P, io, agg);
    else if (one == kGM) hipLaunchKernelGGL(k_pipe_fold1<kGM>, g1, blk, 0, s, P, io, agg);
    else if (one == kGA) hipLaunchKernelGGL(k_pipe_fold1<kGA>, g1, blk, 0, s, P, io, agg);
    else
      hipLaunchKernelGGL(k_pipe_fold,
                         dim3(nbW ? 3 * kPipeRun * div_up(segs, kPipeRun) : segs), blk, 0, s,
                         P, prm, io, agg, gw, nbW);
  }
  TCA_HIP(hipGetLastError());
  {
    KTime kt(s, "k_pipe_group");
    hipLaunchKernelGGL(k_pipe_group, dim3(div_up(groups, kPW)), dim3(kPW * kWave), 0, s, P, agg,
                       gtot);
  }
  TCA_HIP(hipGetLastError());
  // Few segments in all: each emit wave scans its row's group composites
  // itself (at most segs * groups-per-row composites read, from L2) instead
  // of a k_pipe_rows launch.
  const bool rowsInEmit = segs * div_up(P.segsPerRow, kWave) <= kPipeRowsInEmit;
  if (!rowsInEmit) {
    KTime kt(s, "k_pipe_rows");
    hipLaunchKernelGGL(k_pipe_rows, dim3(div_up(rows, kPW)), dim3(kPW * kWave), 0, s, P, gtot,
                       genter);
  }
  TCA_HIP(hipGetLastError());
  {
    KTime kt(s, "k_pipe_emit");
    const double* ge = rowsInEmit ? nullptr : genter;
    if (one == kGE) hipLaunchKernelGGL(k_pipe_emit1<kGE>, g1, blk, 0, s, P, io, agg, gtot, ge);
    else if (one == kGR) hipLaunchKernelGGL(k_pipe_emit1<kGR>, g1, blk, 0, s, P, io, agg, gtot, ge);
    else if (one == kGM) hipLaunchKernelGGL(k_pipe_emit1<kGM>, g1, blk, 0, s, P, io, agg, gtot, ge);
    else if (one == kGA) hipLaunchKernelGGL(k_pipe_emit1<kGA>, g1, blk, 0, s, P, io, agg, gtot, ge);
    else hipLaunchKernelGGL(k_pipe_emit, dim3(segs), blk, 0, s, P, io, agg, gtot, ge);
  }
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// The fused pass: the window kernel and the two recurrence kernels read the
// same inputs and write disjoint outputs -- as one launch (tiles.one) when
// the window look-backs fit 64 bars and at least two parts are requested,
// else as up to three launches in stream order.
ctStatus_t pack_launch(int len, int rows, hipStream_t s, const PackP& p, const PackIO& io,
                       PackTiles t) {
  if (t.twoPass && pipe_serves(len, p)) return launch_pipe(len, rows, s, p, io);
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
  if (len <= 4 * kWave) return PackTiles{{4, 4}, {4, 4}, 4, 1};
  return PackTiles{{8, 1}, {8, 4}, 8, 1};
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
  if (window_halo(P) > kPackMaxWindow)
    return fail(CT_STATUS_UNSUPPORTED, "pack: window look-back above %d bars is not provided",
                kPackMaxWindow);
  if (!(P.mask & (1u << kORsi))) P.rsiP = 2;
  if (!(P.mask & (1u << kOAtr))) P.atrP = 2;
  if (!(P.mask & (1u << kOAdx))) P.adxP = 2;
  return CT_STATUS_SUCCESS;
}

}  // namespace

namespace tca {

ctStatus_t pipe_family(const ctxPackParams& params, const ctxPackOutputs& outputs,
                       const float* h, const float* l, const float* c, int len, int rows,
                       hipStream_t stream, bool* served) {
  *served = false;
  if (t_ragged || len < kPMinLen) return CT_STATUS_SUCCESS;
  PackP P;
  PackIO io;
  if (pack_validate(&params, &outputs, P, io, h, l, c) != CT_STATUS_SUCCESS) {
    clear_error();  // the family path reports its own errors
    return CT_STATUS_SUCCESS;
  }
  if (!pipe_serves(len, P)) return CT_STATUS_SUCCESS;
  // launch_pipe takes the one-launch form where it serves, the two passes
  // otherwise
  *served = true;
  return launch_pipe(len, rows, stream, P, io);
}

ctStatus_t win_family(const ctxPackParams& params, const ctxPackOutputs& outputs,
                      const float* h, const float* l, const float* c, int len, int rows,
                      hipStream_t stream, bool* served) {
  *served = false;
  if (t_ragged) return CT_STATUS_SUCCESS;
  PackP P;
  PackIO io;
  if (pack_validate(&params, &outputs, P, io, h, l, c) != CT_STATUS_SUCCESS) {
    clear_error();  // the family path reports its own errors
    return CT_STATUS_SUCCESS;
  }
  if (!(P.mask & kWinMask) || (P.mask & ~kWinMask)) return CT_STATUS_SUCCESS;
  *served = true;
  return launch_windows(len, rows, stream, P, io);
}

}  // namespace tca

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_pack_batch_d(const float* h, const float* l, const float* c,
                                         int seriesLen, int numSeries, const ctxPackParams* p,
                                         const ctxPackOutputs* o, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    return pack_batch_body(h, l, c, seriesLen, numSeries, p, o, s);
  });
}

}  // extern "C"

namespace tca {
// ctx_pack_batch_d without the error reset and exception guard, for the
// allocation-free entry point that sets up its own scratch first.
ctStatus_t pack_batch_body(const float* h, const float* l, const float* c, int seriesLen,
                           int numSeries, const ctxPackParams* p, const ctxPackOutputs* o,
                           hipStream_t s) {
  if (seriesLen <= 0 || numSeries <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "pack: seriesLen and numSeries must be positive");
  PackP P;
  PackIO io;
  ctStatus_t st = pack_validate(p, o, P, io, h, l, c);
  if (st != CT_STATUS_SUCCESS) return st;
  st = pack_dense(seriesLen, numSeries, s, P, io);
  if (st != CT_STATUS_SUCCESS) return st;
  return finish_batch(s);
}
}  // namespace tca

extern "C" {

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
  // Recorded after every replay; the next replay's stream waits on it, so
  // replays on different streams never share the plan's workspace at once.
  hipEvent_t done = nullptr;
  std::mutex mu;
  PackTiles tiles{{4, 4}, {4, 4}, 4, 1};
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
    ctStatus_t r = time_tiles(len, rows, st, Q, io, PackTiles{cf, cf, 0, 0}, 10, &us);
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
  if (p->done) (void)hipEventSynchronize(p->done);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->done) (void)hipEventDestroy(p->done);
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
      // the whole pass decides: the default (two passes where they serve,
      // else one launch), three launches with the tuned tiles, or one launch
      // of 4 or 8 bars per lane
      float du = 0.0f;
      if (r == CT_STATUS_SUCCESS) r = time_tiles(len, rows, p->stream, P, io, dflt, 10, &du);
      PackTiles best = dflt;
      float bu = du;
      for (int one : {0, 4, 8}) {
        if (r != CT_STATUS_SUCCESS) break;
        PackTiles c = t;
        c.one = one;
        c.twoPass = 0;
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
      const bool two = p->tiles.twoPass && !rg && pipe_serves(len, P);
      p->runK = two ? kPK : one ? p->tiles.one : t.K;
      p->runW = two ? kPW : one ? kFuseW : t.W;
    }
    e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: begin capture"));
    ctStatus_t r = pack_launch(len, rows, p->stream, P, io, p->tiles);
    e = hipStreamEndCapture(p->stream, &p->graph);
    if (r != CT_STATUS_SUCCESS) return bail(r);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: end capture"));
    e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: instantiate"));
    e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(p->done, p->stream);
    if (e != hipSuccess) return bail(hip_fail(e, "plan: event"));
    *out = p;
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ctx_plan_launch(ctxPlan plan, hipStream_t stream) {
  return guarded([&]() -> ctStatus_t {
    if (!plan || !plan->exec) return fail(CT_STATUS_INVALID_ARGUMENT, "plan: null plan");
    // The graph's look-back slots, epoch and carries live in the plan's
    // workspace: a replay starts after the previous one ended, whatever
    // stream either was launched on.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr) TCA_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(CT_STATUS_UNSUPPORTED,
                  "plan: a plan is a captured graph already; launch it outside stream capture");
    std::lock_guard<std::mutex> lk(plan->mu);
    TCA_HIP(hipStreamWaitEvent(stream, plan->done, 0));
    TCA_HIP(hipGraphLaunch(plan->exec, stream));
    TCA_HIP(hipEventRecord(plan->done, stream));
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
