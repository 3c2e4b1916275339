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

struct PackIO {
  const float* h;
  const float* l;
  const float* c;
  float* out[kONum];
};


// Per-bar affine step of a recurrence with warm-up [start, L) summed and the
// seed divided at L.  Returns (A, b) for the composite; `steady` gives the
// update after L.
struct Phase {
  int start, L;
};

__device__ __forceinline__ void warm_coef(const Phase& ph, int t, double u, double invP, double& A,
                                          double& b) {
  if (t == ph.start) {
    A = -0.0;  // structural reset: a new recurrence starts here
    b = (t == ph.L) ? u * invP : u;
  } else if (t < ph.L) {
    A = 1.0;
    b = u;
  } else {  // t == ph.L
    A = invP;
    b = u * invP;
  }
}

// Replay of the warm-up part, TA-Lib's arithmetic: sum, then (sum+u)/p.
__device__ __forceinline__ double warm_step(const Phase& ph, int t, double y, double u, int p) {
  if (t == ph.start) return (t == ph.L) ? u / p : u;
  if (t < ph.L) return y + u;
  return (y + u) / p;
}

__device__ __forceinline__ double warm_step_inv(const Phase& ph, int t, double y, double u,
                                                double inv) {
  if (t == ph.start) return (t == ph.L) ? u * inv : u;
  if (t < ph.L) return y + u;
  return (y + u) * inv;
}

template <class T>
__device__ void pack_store(const Seg& c, float* lo, const float (&v)[T::K], float* dst) {
  if (dst) store_lane<T>(lo, v, dst, c);
}

template <int R>
__device__ __forceinline__ void pack_carry(const Seg& c, int level, Aff<R> loc, double (&yin)[R]) {
  Aff<R> tot;
  wave_scan_excl(loc, tot);
  double ys[R];
#pragma unroll
  for (int r = 0; r < R; ++r) ys[r] = qnan_d();
  if (c.tree) {
    publish_tree(c, level, tot);
    if (c.seg > 0) lookback<R>(c, level, ys);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) yin[r] = aff_apply(loc, r, ys[r]);
}

// Recurrences.  PART 1 reads close only (EMA cascade, RSI, MACD); PART 2
// reads high/low/close (ATR, ADX).  Separate launches keep each kernel's
// registers low enough for several waves per SIMD.
template <class T, int PART>
__device__ __forceinline__ void pack_run(const Seg& c, float* li, float* lo, const PackP& P, const PackIO& io) {
  constexpr int K = T::K;
  float* lh = li;
  float* ll = li + T::PLEN;
  float* lc = li + 2 * T::PLEN;
  if (PART == 2) {
    stage<T>(lh, io.h, c);
    stage<T>(ll, io.l, c);
  }
  stage<T>(lc, io.c, c);
  wave_sync();
  float hv[K], lv[K], cv[K], ov[K];
  if (PART == 2) {
    lane_vals<T>(lh, hv);
    lane_vals<T>(ll, lv);
  }
  lane_vals<T>(lc, cv);
  const int a = c.s0 + lane_id() * K;
  const unsigned m = P.mask;

  // ------------------------------------------------------------ level 1 --
  // PART 1: close recurrences (EMA cascade, RSI, MACD); PART 2: high/low/close
  // recurrences (ATR, ADX).  Each shares one scan per level.
  constexpr bool P1 = PART == 1;
  constexpr int R1 = P1 ? 5 : 4;
  constexpr int R2 = P1 ? 2 : 1;
  // level-1 slot indices
  constexpr int iEma = 0, iG = 1, iL = 2, iFast = 3, iSlow = 4;
  constexpr int iAtr = 0, iPdm = 1, iMdm = 2, iTr = 3;
  const bool wantE = P1 && (m & ((1u << kOEma) | (1u << kODema) | (1u << kOTema)));
  const bool wantR = P1 && (m & (1u << kORsi));
  const bool wantM = P1 && (m & ((1u << kOMacd) | (1u << kOMacdSig) | (1u << kOMacdHist)));
  const bool wantA = !P1 && (m & (1u << kOAtr));
  const bool wantX = !P1 && (m & (1u << kOAdx));
  if (!(wantE || wantR || wantM || wantA || wantX)) return;
  const int f = P1 ? first_valid(c, io.c) : 0;

  const int pe = P.emaP;
  const Phase phE{f, sat_add(f, pe - 1)};
  const Phase phG{sat_add(f, 1), sat_add(f, P.rsiP)};
  const Phase phS{f, sat_add(f, P.mSlow - 1)};
  const Phase phF{sat_add(f, P.mSlow - P.mFast), sat_add(f, P.mSlow - 1)};
  const Phase phA{1, P.atrP};
  const Phase phX{1, P.adxP - 1};  // sums over bars 1..p-1, state is the sum
  double eb, ek, fb, fk2, sb, sk2, gb, gk;
  ema_coef(pe, eb, ek);
  ema_coef(P.mFast, fb, fk2);
  ema_coef(P.mSlow, sb, sk2);
  ema_coef(P.mSig, gb, gk);
  const double invE = 1.0 / pe, invR = 1.0 / (double)P.rsiP, pm1R = (double)(P.rsiP - 1);
  const double invF = 1.0 / P.mFast, invS = 1.0 / P.mSlow;
  const double wbA = (double)(P.atrP - 1) / (double)P.atrP, waA = 1.0 - wbA, invA = 1.0 / P.atrP;
  const double ppX = (double)P.adxP, amX = 1.0 - 1.0 / ppX;

  // per-bar inputs: close deltas (PART 1) or TR and +DM/-DM (PART 2)
  double u0[K], u1[K], u2[K];
  {
    double pc = (double)at<T>(lc, io.c, c, a - 1);
    double ph = P1 ? 0.0 : (double)at<T>(lh, io.h, c, a - 1);
    double pl = P1 ? 0.0 : (double)at<T>(ll, io.l, c, a - 1);
#pragma unroll
    for (int i = 0; i < K; ++i) {
      if (P1) {
        u0[i] = (double)cv[i] - pc;
        u1[i] = u2[i] = 0.0;
      } else {
        u0[i] = true_range(hv[i], lv[i], pc);
        dm_pair(ph, hv[i], pl, lv[i], u1[i], u2[i]);
        ph = hv[i];
        pl = lv[i];
      }
      pc = cv[i];
    }
  }

  Aff<R1> l1;
  aff_id(l1);
  auto push = [&](int r, double A, double b) {
    l1.b[r] = fma_carry(A, l1.b[r], b);
    l1.A[r] = is_reset(A) ? -0.0 : decay_mul(l1.A[r], A);
  };
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    double A, b;
    if (P1) {
      const double x = (double)cv[i];
      if (wantE && t >= phE.start) {
        if (t <= phE.L) warm_coef(phE, t, x, invE, A, b);
        else { A = eb; b = ek * x; }
        push(iEma, A, b);
      }
      if (wantR && t >= phG.start) {
        const double gd = u0[i] > 0.0 ? u0[i] : 0.0, ld = gd - u0[i];
        if (t <= phG.L) {
          warm_coef(phG, t, gd, invR, A, b);
          push(iG, A, b);
          warm_coef(phG, t, ld, invR, A, b);
          push(iL, A, b);
        } else {
          l1.b[iG] = (l1.b[iG] * pm1R + gd) * invR;
          l1.b[iL] = (l1.b[iL] * pm1R + ld) * invR;
          l1.A[iG] = decay_mul(l1.A[iG], pm1R * invR);
          l1.A[iL] = decay_mul(l1.A[iL], pm1R * invR);
        }
      }
      if (wantM) {
        if (t >= phS.start) {
          if (t <= phS.L) warm_coef(phS, t, x, invS, A, b);
          else { A = sb; b = sk2 * x; }
          push(iSlow, A, b);
        }
        if (t >= phF.start) {
          if (t <= phF.L) warm_coef(phF, t, x, invF, A, b);
          else { A = fb; b = fk2 * x; }
          push(iFast, A, b);
        }
      }
    } else {
      if (wantA && t >= phA.start) {
        if (t <= phA.L) warm_coef(phA, t, u0[i], invA, A, b);
        else { A = wbA; b = waA * u0[i]; }
        push(iAtr, A, b);
      }
      if (wantX && t >= phX.start) {
        const double u[3] = {u1[i], u2[i], u0[i]};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const int r = iPdm + q;
          if (t == phX.start) {
            l1.b[r] = u[q];
            l1.A[r] = -0.0;
          } else if (t <= phX.L) {
            l1.b[r] = l1.b[r] + u[q];
          } else {
            l1.b[r] = l1.b[r] - l1.b[r] / ppX + u[q];
            l1.A[r] = is_reset(l1.A[r]) ? -0.0 : decay_mul(l1.A[r], amX);
          }
        }
      }
    }
  }
  double y1[R1];
  pack_carry<R1>(c, 0, l1, y1);

  // replay level 1: e1 / macd line (PART 1) or DX (PART 2) per bar
  double v1[K], v2[K];
  {
    float oa[K], ob[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      v1[i] = v2[i] = qnan_();
      oa[i] = ob[i] = qnanf_();
    }
    if (P1) {
      double ye = y1[iEma], yg = y1[iG], yl = y1[iL], yf = y1[iFast], ys = y1[iSlow];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        const double x = (double)cv[i];
        if (wantE && t >= phE.start) {
          ye = (t <= phE.L) ? warm_step(phE, t, ye, x, pe) : fma(eb, ye, ek * x);
          if (t >= phE.L) {
            v1[i] = ye;
            oa[i] = (float)ye;
          }
        }
        if (pe == 1) oa[i] = cv[i];
        if (wantR && t >= phG.start) {
          const double gd = u0[i] > 0.0 ? u0[i] : 0.0;
          if (t <= phG.L) {
            yg = warm_step_inv(phG, t, yg, gd, invR);
            yl = warm_step_inv(phG, t, yl, gd - u0[i], invR);
          } else {
            yl *= pm1R;
            yg *= pm1R;
            yg += gd;
            yl += gd - u0[i];
            yl *= invR;
            yg *= invR;
          }
          if (t >= phG.L) {
            const double s = yg + yl;
            ob[i] = (float)(s > 0.0 ? 100.0 * (yg / s) : 50.0);
          }
        }
        if (wantM) {
          if (t >= phS.start) ys = (t <= phS.L) ? warm_step(phS, t, ys, x, P.mSlow) : fma(sb, ys, sk2 * x);
          if (t >= phF.start) yf = (t <= phF.L) ? warm_step(phF, t, yf, x, P.mFast) : fma(fb, yf, fk2 * x);
          if (t >= phS.L) v2[i] = yf - ys;
        }
      }
      if (m & (1u << kOEma)) pack_store<T>(c, lo, oa, io.out[kOEma]);
      if (wantR) pack_store<T>(c, lo, ob, io.out[kORsi]);
    } else {
      double ya = y1[iAtr], yp = y1[iPdm], yq = y1[iMdm], yt = y1[iTr];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (wantA && t >= phA.start) {
          ya = (t <= phA.L) ? warm_step(phA, t, ya, u0[i], P.atrP) : fma(wbA, ya, waA * u0[i]);
          if (t >= phA.L) oa[i] = (float)ya;
        }
        if (wantX && t >= phX.start) {
          if (t == phX.start) {
            yp = u1[i];
            yq = u2[i];
            yt = u0[i];
          } else if (t <= phX.L) {
            yp += u1[i];
            yq += u2[i];
            yt += u0[i];
          } else {
            yp = yp - yp / ppX + u1[i];
            yq = yq - yq / ppX + u2[i];
            yt = yt - yt / ppX + u0[i];
          }
          if (t > phX.L) {
            double pdi, mdi, d;
            dir_outputs(yp, yq, yt, pdi, mdi, d);
            v1[i] = d;
          }
        }
      }
      if (wantA) pack_store<T>(c, lo, oa, io.out[kOAtr]);
    }
  }

  // ------------------------------------------------------------ level 2 --
  // PART 1: DEMA/TEMA e2 (input e1) and the MACD signal (input macd line);
  // PART 2: ADX (input DX).
  const bool wantE2 = P1 && (m & ((1u << kODema) | (1u << kOTema))) && pe > 1;
  if (P1 && (m & (1u << kODema)) && pe == 1) pack_store<T>(c, lo, cv, io.out[kODema]);
  if (P1 && (m & (1u << kOTema)) && pe == 1) pack_store<T>(c, lo, cv, io.out[kOTema]);
  if (!(wantE2 || wantM || wantX)) return;
  const Phase phE2{phE.L, sat_add(phE.L, pe - 1)};
  const Phase phSg{phS.L, sat_add(phS.L, P.mSig - 1)};
  const Phase phAx{P.adxP, 2 * P.adxP - 1};
  const double invSg = 1.0 / P.mSig;
  constexpr int jE2 = 0, jSig = 1, jAdx = 0;
  Aff<R2> l2;
  aff_id(l2);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n) continue;
    double A, b;
    if (P1) {
      if (wantE2 && t >= phE2.start) {
        if (t <= phE2.L) warm_coef(phE2, t, v1[i], invE, A, b);
        else { A = eb; b = ek * v1[i]; }
        l2.b[jE2] = fma_carry(A, l2.b[jE2], b);
        l2.A[jE2] = is_reset(A) ? -0.0 : decay_mul(l2.A[jE2], A);
      }
      if (wantM && t >= phSg.start) {
        if (t <= phSg.L) warm_coef(phSg, t, v2[i], invSg, A, b);
        else if (P.mSig == 1) { A = -0.0; b = v2[i]; }
        else { A = gb; b = gk * v2[i]; }
        l2.b[jSig % R2] = fma_carry(A, l2.b[jSig % R2], b);
        l2.A[jSig % R2] = is_reset(A) ? -0.0 : decay_mul(l2.A[jSig % R2], A);
      }
    } else if (t >= phAx.start) {
      if (t <= phAx.L) {
        warm_coef(phAx, t, v1[i], 1.0 / ppX, A, b);
        l2.b[jAdx] = fma_carry(A, l2.b[jAdx], b);
      } else {
        A = (ppX - 1.0) / ppX;
        l2.b[jAdx] = ((l2.b[jAdx] * (ppX - 1.0)) + v1[i]) / ppX;
      }
      l2.A[jAdx] = is_reset(A) ? -0.0 : decay_mul(l2.A[jAdx], A);
    }
  }
  double y2[R2];
  pack_carry<R2>(c, 1, l2, y2);
  double e2[K];
  {
    float od[K], om[K], os[K], oh[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      e2[i] = qnan_();
      od[i] = om[i] = os[i] = oh[i] = qnanf_();
    }
    if (P1) {
      double ye = y2[jE2], yg = y2[jSig % R2];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (wantE2 && t >= phE2.start) {
          ye = (t <= phE2.L) ? warm_step(phE2, t, ye, v1[i], pe) : fma(eb, ye, ek * v1[i]);
          if (t >= phE2.L) {
            e2[i] = ye;
            od[i] = (float)(2.0 * v1[i] - ye);
          }
        }
        if (wantM && t >= phSg.start) {
          if (t <= phSg.L) yg = warm_step(phSg, t, yg, v2[i], P.mSig);
          else yg = (P.mSig == 1) ? v2[i] : fma(gb, yg, gk * v2[i]);
          if (t >= phSg.L) {
            om[i] = (float)v2[i];
            os[i] = (float)yg;
            oh[i] = (float)(v2[i] - yg);
          }
        }
      }
      if ((m & (1u << kODema)) && pe > 1) pack_store<T>(c, lo, od, io.out[kODema]);
      if (wantM) {
        pack_store<T>(c, lo, om, io.out[kOMacd]);
        pack_store<T>(c, lo, os, io.out[kOMacdSig]);
        pack_store<T>(c, lo, oh, io.out[kOMacdHist]);
      }
    } else {
      double yx = y2[jAdx];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = a + i;
        if (t >= phAx.start) {
          yx = (t <= phAx.L) ? warm_step(phAx, t, yx, v1[i], P.adxP)
                             : ((yx * (ppX - 1.0)) + v1[i]) / ppX;
          if (t >= phAx.L) od[i] = (float)yx;
        }
      }
      pack_store<T>(c, lo, od, io.out[kOAdx]);
    }
  }

  // ------------------------------------------------------------ level 3 --
  if (!P1 || !(m & (1u << kOTema)) || pe == 1) return;
  const Phase phE3{phE2.L, sat_add(phE2.L, pe - 1)};
  Aff<1> l3;
  aff_id(l3);
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const int t = a + i;
    if (t >= c.n || t < phE3.start) continue;
    double A, b;
    if (t <= phE3.L) warm_coef(phE3, t, e2[i], invE, A, b);
    else { A = eb; b = ek * e2[i]; }
    l3.b[0] = fma_carry(A, l3.b[0], b);
    l3.A[0] = is_reset(A) ? -0.0 : decay_mul(l3.A[0], A);
  }
  double y3[1];
  pack_carry<1>(c, 2, l3, y3);
  {
    double ye = y3[0];
#pragma unroll
    for (int i = 0; i < K; ++i) {
      const int t = a + i;
      ov[i] = qnanf_();
      if (t >= phE3.start) {
        ye = (t <= phE3.L) ? warm_step(phE3, t, ye, e2[i], pe) : fma(eb, ye, ek * e2[i]);
        if (t >= phE3.L) ov[i] = (float)(ye + (3.0 * v1[i] - 3.0 * e2[i]));
      }
    }
    pack_store<T>(c, lo, ov, io.out[kOTema]);
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

  if (m & (1u << kOSma)) {
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
  if (m & ((1u << kOBbU) | (1u << kOBbM) | (1u << kOBbL))) {
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
    pack_run<T, PART>(c, s_in[w], s_out[w], prm, io);
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

#ifndef TCA_PACK_W
#define TCA_PACK_W 4
#endif

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

template <int K>
static ctStatus_t launch_pack(Geometry g, int maxSegsPerRow, int len, int rows, hipStream_t s,
                              const PackP& prm, const PackIO& io) {
  if (g.totalSegs <= 0) return CT_STATUS_SUCCESS;
  constexpr int W = TCA_PACK_W;
  const int blocks = div_up(g.totalSegs, W);
  if (prm.mask & kWinMask) {
    ctStatus_t st = launch_windows(len, rows, s, prm, io);
    if (st != CT_STATUS_SUCCESS) return st;
  }
  const bool useLB = maxSegsPerRow > 1;
  // 3 levels x (segments + groups) slots of up to 5 (A, b) pairs
  Ws ws{nullptr, nullptr, 2 * 5 + 1};
  if (useLB && (prm.mask & ~kWinMask)) {
    const size_t bytes = (size_t)3 * ((size_t)g.totalSegs + g.totalSegs / kGroup) * (2 * 5 + 1) * 8;
    ctStatus_t st = workspace_for(s, bytes, &ws);
    if (st != CT_STATUS_SUCCESS) return st;
  }
  if (prm.mask & kCloseRecMask) {
    hipLaunchKernelGGL((k_pack<K, K, W, 1>), dim3(blocks), dim3(W * kWave), 0, s, g, ws,
                       useLB ? 1 : 0, prm, io);
    TCA_HIP(hipGetLastError());
  }
  if (prm.mask & kHlcRecMask) {
    hipLaunchKernelGGL((k_pack<K, K, W, 2>), dim3(blocks), dim3(W * kWave), 0, s, g, ws,
                       useLB ? 1 : 0, prm, io);
    TCA_HIP(hipGetLastError());
  }
  return CT_STATUS_SUCCESS;
}

// Dense or single-series pack: 4 bars per lane (one wave covers a 256-bar
// row), segments chained by the look-back tree for longer rows.
ctStatus_t pack_dense(int len, int rows, hipStream_t s, const PackP& p, const PackIO& io) {
  Geometry g = dense_geometry(len, rows, 4 * kWave);
  return launch_pack<4>(g, g.segsPerRow, len, rows, s, p, io);
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

}  // extern "C"
