// Fused long-series scan: EMA, KAMA, T3 and SMA of one close series per row
// from one launch sequence (no counterpart in the reference, whose families
// run as separate calls: EMA.cu, KAMA.cu, T3.cu, SMA.cu).  Results follow the
// per-family definitions (ta_EMA.c, ta_KAMA.c, ta_T3.c, the reference SMA).
//
// Long rows are scanned in two streaming passes instead of a decoupled
// look-back, so no workgroup ever waits on another:
//   1. k_scan_lead  -- first non-NaN bar of every row (the seeds start there);
//   2. k_scan_fold  -- every 1024-bar segment folds its bars from a zero state
//                      and writes its composite (EMA offset, T3 6-vector
//                      offset, KAMA affine pair);
//   3. k_scan_carry -- one wave per row scans the segment composites into the
//                      state entering every segment;
//   4. k_scan_emit  -- every segment re-reads its bars, rebuilds the states
//                      entering each lane from its carry and replays the bars
//                      with the families' own sequential arithmetic.
// The input is read twice and each output written once.
//
// Every recurrence here starts at a seed bar with everything before it
// contributing exactly 0, so a run of L bars maps the state s to
// M^L s + o: EMA has M = beta, T3's six cascaded EMAs form one linear
// recurrence whose matrix is lower-triangular Toeplitz (entry (i, j) =
// (1-k) k^(i-j)), so its powers are Toeplitz too (six numbers each) and are
// computed on the host.  KAMA's multiplier (1 - sc_t) varies per bar and is
// carried as an (A, b) pair.  Powers of nonzero factors that underflow are
// clamped to the smallest normal double, so a carried Inf survives them as
// it does in the sequential recurrence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "../../include/tacuda_amd.h"
#include "families.hpp"
#include "host.hpp"

namespace tca {
namespace {

constexpr int kLK = 16;             // bars per lane
constexpr int kLW = 4;              // waves (segments) per workgroup
constexpr int kLSeg = kLK * kWave;  // bars per segment
constexpr int kLMaxHalo = 1024;     // longest window look-back staged
constexpr unsigned kLEma = 1, kLKama = 2, kLT3 = 4, kLSma = 8;

struct Toe {
  double c[6];  // first column of a lower-triangular Toeplitz 6x6 matrix
};

struct ScanP {
  int n, rows, segsPerRow, cpl;  // cpl: segments per lane in the carry scan
  unsigned mask;
  int useCarry;                  // rows span several segments (carries exist)
  int pe, pk, pt, ps;            // EMA, KAMA, T3, SMA periods
  double eb, ek;                 // EMA beta and weight (ta_EMA.c)
  double kmax, kdiff;            // KAMA constants (ta_KAMA.c)
  double tk, tomk;               // T3 weight and 1 - weight (ta_T3.c)
  double c1, c2, c3, c4;         // T3 combination
  double ePow[6], eSeg, eRow[6]; // beta^(K 2^s), beta^SEG, beta^(SEG cpl 2^s)
  Toe tPow[6], tSeg, tRow[6];    // the same powers of the T3 matrix
};

// y += T v
__device__ __forceinline__ void toe_mv_add(const Toe& T, const double (&v)[6], double (&y)[6]) {
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
#pragma unroll
    for (int j = 0; j <= i; ++j) s = fma(T.c[i - j], v[j], s);
    y[i] = s;
  }
}

template <class T>
__device__ __forceinline__ float xat(const float* lds, const float* x, const Seg& c, int t) {
  return at<T>(lds, x, c, t);
}

// --------------------------------------------------------------- seeds --
// Computed by every lane of the wave holding the seed bar, in lockstep (the
// same bar at a time), in the oracle's order.
template <class T>
__device__ double ema_seed(const float* lds, const float* x, const Seg& c, int f, int p) {
  double s = 0.0;
  for (int t = f; t < f + p; ++t) s += (double)xat<T>(lds, x, c, t);
  return s / p;
}

// State (e1..e6) of ta_T3.c's cascade at its first output bar f + 6(p-1).
template <class T>
__device__ void t3_seed(const float* lds, const float* x, const Seg& c, int f, int p,
                        const ScanP& P, double (&e)[6]) {
  const double k = P.tk, omk = P.tomk;
  int t = f;
  double tmp = 0.0;
  for (int i = 0; i < p; ++i) tmp += (double)xat<T>(lds, x, c, t++);
  e[0] = tmp / p;
  for (int lev = 1; lev < 6; ++lev) {
    tmp = e[lev - 1];
    for (int i = p - 1; i > 0; --i) {
      double in = (double)xat<T>(lds, x, c, t++);
      e[0] = fma(omk, e[0], k * in);
      for (int q = 1; q < lev; ++q) e[q] = fma(omk, e[q], k * e[q - 1]);
      tmp += e[lev - 1];
    }
    e[lev] = tmp / p;
  }
}

// ------------------------------------------------------------ KAMA bar --
// Sliding |delta| window of one lane: sum over bars t-p+1..t of |x_j - x_{j-1}|.
struct KamaWin {
  double sum;
  int t;  // bar the sum is for; -1 = not started
};

template <class T>
__device__ __forceinline__ double kama_sc(const float* lds, const float* x, const Seg& c,
                                          const ScanP& P, KamaWin& w, int t, double xt) {
  const int p = P.pk;
  if (w.t < 0) {
    double s = 0.0;
    for (int j = t - p + 1; j <= t; ++j)
      s += fabs((double)xat<T>(lds, x, c, j) - (double)xat<T>(lds, x, c, j - 1));
    w.sum = s;
  } else {
    const double xo = (double)xat<T>(lds, x, c, t - p), xo1 = (double)xat<T>(lds, x, c, t - p - 1);
    const double xp = (double)xat<T>(lds, x, c, t - 1);
    w.sum = w.sum + fabs(xt - xp) - fabs(xo - xo1);
  }
  w.t = t;
  const double sum = w.sum;
  const double roc = xt - (double)xat<T>(lds, x, c, t - p);
  double er;
  if (sum <= 0.0 || sum <= roc) {
    er = 1.0;
  } else {
    er = fabs(roc / sum);
    if (er > 1.0) er = 1.0;
  }
  double sc = fma(er, P.kdiff, P.kmax);
  return sc * sc;
}

// -------------------------------------------------------------- lane fold --
// Offsets of the lane's K bars from a zero state (contributions before a
// seed are 0; the seed bar sets the state).
struct Fold {
  double be;          // EMA
  double bt[6];       // T3 cascade
  double Ak, bk;      // KAMA: state -> Ak*state + bk
};

struct Seeds {
  int f, Le, Lk, L6;
  double se;     // EMA seed
  double st[6];  // T3 state at L6
};

template <class T>
__device__ __forceinline__ Seeds make_seeds(const float* lds, const float* x, const Seg& c,
                                            const ScanP& P, int f) {
  Seeds S;
  S.f = f;
  S.Le = f + P.pe - 1;
  S.Lk = f + P.pk - 1;
  S.L6 = f + 6 * (P.pt - 1);
  S.se = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) S.st[i] = 0.0;
  auto here = [&](int L) { return L >= c.s0 && L < c.s1; };
  if ((P.mask & kLEma) && P.pe > 1 && here(S.Le)) S.se = ema_seed<T>(lds, x, c, f, P.pe);
  if ((P.mask & kLT3) && here(S.L6)) t3_seed<T>(lds, x, c, f, P.pt, P, S.st);
  return S;
}

template <class T>
__device__ __forceinline__ Fold fold_lane(const float* lds, const float* x, const Seg& c,
                                          const ScanP& P, const Seeds& S, const float (&xv)[kLK],
                                          int a) {
  Fold F;
  F.be = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) F.bt[i] = 0.0;
  F.Ak = 1.0;
  F.bk = 0.0;
  const bool wE = (P.mask & kLEma) && P.pe > 1, wT = P.mask & kLT3, wK = (P.mask & kLKama) && P.pk > 1;
  KamaWin kw{0.0, -1};
#pragma unroll
  for (int i = 0; i < kLK; ++i) {
    const int t = a + i;
    if (t >= c.s1) break;
    const double v = (double)xv[i];
    if (wE) {
      if (t == S.Le) F.be = S.se;
      else if (t > S.Le) F.be = fma(P.eb, F.be, P.ek * v);
    }
    if (wT) {
      if (t == S.L6) {
#pragma unroll
        for (int q = 0; q < 6; ++q) F.bt[q] = S.st[q];
      } else if (t > S.L6) {
        F.bt[0] = fma(P.tomk, F.bt[0], P.tk * v);
#pragma unroll
        for (int q = 1; q < 6; ++q) F.bt[q] = fma(P.tomk, F.bt[q], P.tk * F.bt[q - 1]);
      }
    }
    if (wK) {
      if (t == S.Lk) {
        F.Ak = 0.0;
        F.bk = v;
      } else if (t > S.Lk) {
        const double sc = kama_sc<T>(lds, x, c, P, kw, t, v);
        const double m = 1.0 - sc;
        F.Ak *= m;
        F.bk = fma(m, F.bk, sc * v);
      }
    }
  }
  return F;
}

// Inclusive lane scan of the folds (lane 0 earliest), of the requested
// recurrences only.
__device__ __forceinline__ void scan_lanes(const ScanP& P, Fold& F) {
  const int lane = lane_id();
  const bool wE = P.mask & kLEma, wT = P.mask & kLT3, wK = P.mask & kLKama;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int d = 1 << s;
    if (wE) {
      const double be = __shfl_up(F.be, d);
      if (lane >= d) F.be = fma(P.ePow[s], be, F.be);
    }
    if (wT) {
      double bt[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) bt[q] = __shfl_up(F.bt[q], d);
      if (lane >= d) toe_mv_add(P.tPow[s], bt, F.bt);
    }
    if (wK) {
      const double Ak = __shfl_up(F.Ak, d), bk = __shfl_up(F.bk, d);
      if (lane >= d) {
        F.bk = fma(F.Ak, bk, F.bk);
        F.Ak *= Ak;
      }
    }
  }
}

__device__ __forceinline__ Seg seg_of(const ScanP& P, int seg, int& row) {
  Seg c;
  row = seg / P.segsPerRow;
  c.seg = seg - row * P.segsPerRow;
  c.nseg = P.segsPerRow;
  c.gseg = seg;
  c.rowOff = (long long)row * P.n;
  c.n = P.n;
  c.s0 = c.seg * kLSeg;
  c.s1 = min(P.n, c.s0 + kLSeg);
  c.totalSegs = P.rows * P.segsPerRow;
  c.tree = false;
  c.epoch = 0;
  return c;
}

// ------------------------------------------------------- steady segments --
// A full segment whose bars all lie past every seed and warm-up and whose
// staged in-row bars are all finite: every bar takes the steady step, so the
// lane loops carry no per-bar conditions.  Results equal the general path's
// (the same operations in the same order).
template <class T>
__device__ __forceinline__ bool steady(const float* lds, const Seg& c, const ScanP& P,
                                       const Seeds& S) {
  if (c.s1 - c.s0 != kLSeg) return false;
  if (((P.mask & kLEma) && P.pe == 1) || ((P.mask & kLKama) && P.pk == 1)) return false;
  int last = -1;
  if (P.mask & kLEma) last = max(last, S.Le);
  if (P.mask & kLT3) last = max(last, S.L6);
  if (P.mask & kLKama) last = max(last, S.Lk);
  if (P.mask & kLSma) last = max(last, P.ps - 1);
  if (c.s0 <= last) return false;
  bool bad = false;
  for (int e = lane_id(); e < T::LEN; e += kWave) {
    const int t = c.s0 - T::H + e;
    if (t >= 0 && t < c.n) bad |= !__builtin_isfinite(lds_ld(lds, T::pos(e)));
  }
  return __ballot(bad) == 0;
}

// Fold of a steady lane; sc receives KAMA's smoothing constant of each bar.
template <class T>
__device__ __forceinline__ Fold fold_steady(const float* lds, const Seg& c, const ScanP& P,
                                            const float (&xv)[kLK], int a, double (&sc)[kLK]) {
  Fold F;
  F.be = 0.0;
#pragma unroll
  for (int i = 0; i < 6; ++i) F.bt[i] = 0.0;
  F.Ak = 1.0;
  F.bk = 0.0;
  if (P.mask & kLEma) {
#pragma unroll
    for (int i = 0; i < kLK; ++i) F.be = fma(P.eb, F.be, P.ek * (double)xv[i]);
  }
  if (P.mask & kLT3) {
#pragma unroll
    for (int i = 0; i < kLK; ++i) {
      F.bt[0] = fma(P.tomk, F.bt[0], P.tk * (double)xv[i]);
#pragma unroll
      for (int q = 1; q < 6; ++q) F.bt[q] = fma(P.tomk, F.bt[q], P.tk * F.bt[q - 1]);
    }
  }
  if (P.mask & kLKama) {
    const int p = P.pk;
    const int e0 = a - p - 1 - (c.s0 - T::H);  // staged index of x[a - p - 1]
    float xb[kLK + 1];                          // x[a - p - 1 + j]
#pragma unroll
    for (int j = 0; j <= kLK; ++j) xb[j] = lds_ld(lds, T::pos(e0 + j));
    // the window of bar a, in the oracle's order
    double sum = 0.0;
    double xprev = (double)xb[1];
    for (int e = e0 + 2; e <= e0 + 1 + p; ++e) {
      const double v = (double)lds_ld(lds, T::pos(e));
      sum += fabs(v - xprev);
      xprev = v;
    }
#pragma unroll
    for (int i = 0; i < kLK; ++i) {
      const double xt = (double)xv[i];
      if (i > 0) sum = sum + fabs(xt - xprev) - fabs((double)xb[i + 1] - (double)xb[i]);
      xprev = xt;
      const double roc = xt - (double)xb[i + 1];
      double er;
      if (sum <= 0.0 || sum <= roc) {
        er = 1.0;
      } else {
        er = fabs(roc / sum);
        if (er > 1.0) er = 1.0;
      }
      double s2 = fma(er, P.kdiff, P.kmax);
      s2 *= s2;
      sc[i] = s2;
      const double m = 1.0 - s2;
      F.Ak *= m;
      F.bk = fma(m, F.bk, s2 * xt);
    }
  }
  return F;
}

// ------------------------------------------------------------- kernels --
__global__ __launch_bounds__(64) void k_scan_lead(const float* x, int n, int rows, int* lead) {
  const int row = blockIdx.x;
  if (row >= rows) return;
  const float* r = x + (size_t)row * n;
  int f = n;
  for (int t0 = 0; t0 < n; t0 += kWave) {
    const int t = t0 + lane_id();
    const bool ok = t < n && !__builtin_isnan(r[t]);
    const unsigned long long m = __ballot(ok);
    if (m) {
      f = t0 + __builtin_ctzll(m);
      break;
    }
  }
  if (lane_id() == 0) lead[row] = f;
}

// FM: the requested families when fixed at compile time (a single-family
// call), so the others' code and registers drop out; 0 = P.mask at run time.
template <int H, unsigned FM>
__global__ __launch_bounds__(kLW* kWave) void k_scan_fold(const float* x, const int* lead, ScanP P,
                                                          double* agg) {
  using T = Tile<kLK, H>;
  if (FM) P.mask = FM;
  __shared__ __align__(16) float s_in[kLW][T::PLEN];
  const int w = threadIdx.x / kWave;
  const int seg = blockIdx.x * kLW + w;
  if (seg >= P.rows * P.segsPerRow) return;
  int row;
  const Seg c = seg_of(P, seg, row);
  float* lds = s_in[w];
  stage<T>(lds, x, c);
  wave_sync();
  float xv[kLK];
  lane_vals<T>(lds, xv);
  const int a = c.s0 + lane_id() * kLK;
  const Seeds S = make_seeds<T>(lds, x, c, P, lead[row]);
  Fold F;
  if (steady<T>(lds, c, P, S)) {
    double sc[kLK];
    F = fold_steady<T>(lds, c, P, xv, a, sc);
  } else {
    F = fold_lane<T>(lds, x, c, P, S, xv, a);
  }
  scan_lanes(P, F);
  if (lane_id() == kWave - 1) {
    double* o = agg + (size_t)seg * 9;
    o[0] = F.be;
#pragma unroll
    for (int q = 0; q < 6; ++q) o[1 + q] = F.bt[q];
    o[7] = F.Ak;
    o[8] = F.bk;
  }
}

// One wave per row: the state entering every segment of the row.
__global__ __launch_bounds__(64) void k_scan_carry(ScanP P, const double* agg, double* carry) {
  const int row = blockIdx.x;
  if (row >= P.rows) return;
  const int S = P.segsPerRow, cpl = P.cpl, lane = lane_id();
  const size_t base = (size_t)row * S;
  const int j0 = lane * cpl, j1 = min(S, j0 + cpl);
  Fold F;
  F.be = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) F.bt[q] = 0.0;
  F.Ak = 1.0;
  F.bk = 0.0;
  for (int j = j0; j < j1; ++j) {
    const double* g = agg + (base + j) * 9;
    F.be = fma(P.eSeg, F.be, g[0]);
    double v[6], y[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      v[q] = F.bt[q];
      y[q] = g[1 + q];
    }
    toe_mv_add(P.tSeg, v, y);
#pragma unroll
    for (int q = 0; q < 6; ++q) F.bt[q] = y[q];
    F.bk = fma(g[7], F.bk, g[8]);
    F.Ak *= g[7];
  }
  // inclusive scan over lanes with cpl-segment multipliers
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int d = 1 << s;
    const double be = __shfl_up(F.be, d);
    double bt[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) bt[q] = __shfl_up(F.bt[q], d);
    const double Ak = __shfl_up(F.Ak, d), bk = __shfl_up(F.bk, d);
    if (lane >= d) {
      F.be = fma(P.eRow[s], be, F.be);
      toe_mv_add(P.tRow[s], bt, F.bt);
      F.bk = fma(F.Ak, bk, F.bk);
      F.Ak *= Ak;
    }
  }
  // exclusive: the state entering this lane's first segment
  double ce = __shfl_up(F.be, 1), ck = __shfl_up(F.bk, 1), ct[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) ct[q] = __shfl_up(F.bt[q], 1);
  if (lane == 0) {
    ce = ck = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) ct[q] = 0.0;
  }
  for (int j = j0; j < j1; ++j) {
    double* o = carry + (base + j) * 8;
    const double* g = agg + (base + j) * 9;
    o[0] = ce;
#pragma unroll
    for (int q = 0; q < 6; ++q) o[1 + q] = ct[q];
    o[7] = ck;
    ce = fma(P.eSeg, ce, g[0]);
    double y[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) y[q] = g[1 + q];
    toe_mv_add(P.tSeg, ct, y);
#pragma unroll
    for (int q = 0; q < 6; ++q) ct[q] = y[q];
    ck = fma(g[7], ck, g[8]);
  }
}

// Eight outputs of a lane at bar a + 8h: float4 stores when whole and
// aligned, else per bar.
__device__ __forceinline__ void put8(float* dst, const Seg& c, int t0, const float (&v)[8]) {
  if (!dst) return;
  float* p = dst + c.rowOff + t0;
  if (t0 + 8 <= c.s1 && ((reinterpret_cast<size_t>(p) & 15) == 0)) {
    const v4f a = {v[0], v[1], v[2], v[3]}, b = {v[4], v[5], v[6], v[7]};
    __builtin_nontemporal_store(a, reinterpret_cast<v4f*>(p));
    __builtin_nontemporal_store(b, reinterpret_cast<v4f*>(p + 4));
    return;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (t0 + i < c.s1) p[i] = v[i];
}

// A steady lane's 16 outputs (the segment is full: always in the row).
__device__ __forceinline__ void put16(float* dst, const Seg& c, int a, const float (&v)[kLK]) {
  float* p = dst + c.rowOff + a;
  if ((reinterpret_cast<size_t>(p) & 15) == 0) {
#pragma unroll
    for (int q = 0; q < kLK / 4; ++q)
      *reinterpret_cast<float4*>(p + 4 * q) = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
    return;
  }
#pragma unroll
  for (int i = 0; i < kLK; ++i) p[i] = v[i];
}

struct ScanOut {
  float *ema, *kama, *t3, *sma;
};

// SMA window sums of a wave whose staged in-row bars are all finite: prefix
// sums of the staged range's 16-bar chunks (pc[k] = bars of chunks < k),
// so a lane's first window costs its two partial chunks plus a difference.
// Lane l holds the prefix of chunks 2l and 2l + 1 in registers; other lanes
// read them with a cross-lane shuffle, which keeps them out of LDS (and the
// emit launch's LDS at one staged tile per wave).
struct SmaPrefix {
  double excl;   // pc[2 * lane]
  double c0;     // bars of chunk 2 * lane: pc[2 * lane + 1] = excl + c0
  double total;  // pc of one past the last chunk
};

template <class T>
__device__ __forceinline__ bool sma_prefix(const float* lds, const Seg& c, SmaPrefix& pr,
                                           bool check) {
  constexpr int NCH = T::LEN / kLK;
  static_assert(NCH <= 2 * kWave, "two chunks per lane");
  const int lane = lane_id();
  bool bad = false;
  double cs[2] = {0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k = 2 * lane + q;
    if (k >= NCH) break;
#pragma unroll
    for (int i = 0; i < kLK; ++i) {
      const int e = k * kLK + i, t = c.s0 - T::H + e;
      if (t < 0 || t >= c.n) continue;
      const float v = lds_ld(lds, T::pos(e));
      bad |= !__builtin_isfinite(v);
      cs[q] += (double)v;
    }
  }
  if (check && __ballot(bad)) return false;
  double incl = cs[0] + cs[1];
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  pr.excl = incl - (cs[0] + cs[1]);
  pr.c0 = cs[0];
  pr.total = bcast_d(incl, kWave - 1);
  return true;
}

// pc[k] for this lane's k (0 <= k <= the chunk count).
__device__ __forceinline__ double sma_pc(const SmaPrefix& pr, int k) {
  const int src = k >> 1;
  const double ex = __shfl(pr.excl, src < kWave ? src : 0);
  const double c0 = __shfl(pr.c0, src < kWave ? src : 0);
  return src >= kWave ? pr.total : (k & 1) ? ex + c0 : ex;
}

// The chunk prefixes a window e_lo..e_hi needs, read by every lane of the
// wave together (the shuffles need all lanes).
struct SmaEnds {
  double lo, hi;
};
__device__ __forceinline__ SmaEnds sma_ends(const SmaPrefix& pr, int e_lo, int e_hi) {
  const int k_lo = (e_lo + kLK - 1) / kLK, k_hi = (e_hi + 1) / kLK;
  return SmaEnds{sma_pc(pr, k_lo), sma_pc(pr, k_hi)};
}

// Sum of staged bars e_lo..e_hi (inclusive) from the chunk prefix sums and
// the partial chunks at both ends.
template <class T>
__device__ __forceinline__ double sma_window(const float* lds, const SmaEnds& pc, int e_lo,
                                             int e_hi) {
  const int k_lo = (e_lo + kLK - 1) / kLK, k_hi = (e_hi + 1) / kLK;
  double s = 0.0;
  if (k_lo >= k_hi) {
    for (int e = e_lo; e <= e_hi; ++e) s += (double)lds_ld(lds, T::pos(e));
    return s;
  }
  for (int e = e_lo; e < k_lo * kLK; ++e) s += (double)lds_ld(lds, T::pos(e));
  s += pc.hi - pc.lo;
  for (int e = k_hi * kLK; e <= e_hi; ++e) s += (double)lds_ld(lds, T::pos(e));
  return s;
}

template <int H, unsigned FM>
__global__ __launch_bounds__(kLW* kWave) void k_scan_emit(const float* x, const int* lead, ScanP P,
                                                          const double* carry, ScanOut out) {
  using T = Tile<kLK, H>;
  if (FM) P.mask = FM;
  __shared__ __align__(16) float s_in[kLW][T::PLEN];
  const int w = threadIdx.x / kWave;
  const int seg = blockIdx.x * kLW + w;
  if (seg >= P.rows * P.segsPerRow) return;
  int row;
  const Seg c = seg_of(P, seg, row);
  float* lds = s_in[w];
  stage<T>(lds, x, c);
  wave_sync();
  float xv[kLK];
  lane_vals<T>(lds, xv);
  const int lane = lane_id();
  const int a = c.s0 + lane * kLK;
  const Seeds S = make_seeds<T>(lds, x, c, P, lead[row]);
  const bool fast = steady<T>(lds, c, P, S);
  double sc[kLK];
  Fold F = fast ? fold_steady<T>(lds, c, P, xv, a, sc) : fold_lane<T>(lds, x, c, P, S, xv, a);
  // lane 0 absorbs the state entering the segment; the inclusive scan then
  // gives the state after every lane's bars
  double ce = 0.0, ck = 0.0, ct[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (P.useCarry) {
    const double* g = carry + (size_t)seg * 8;
    ce = g[0];
    ck = g[7];
#pragma unroll
    for (int q = 0; q < 6; ++q) ct[q] = g[1 + q];
  }
  if (lane == 0) {
    F.be = fma(P.ePow[0], ce, F.be);
    toe_mv_add(P.tPow[0], ct, F.bt);
    F.bk = fma(F.Ak, ck, F.bk);
  }
  scan_lanes(P, F);
  double ye = 0.0, yk = 0.0, yt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (P.mask & kLEma) ye = __shfl_up(F.be, 1);
  if (P.mask & kLKama) yk = __shfl_up(F.bk, 1);
  if (P.mask & kLT3) {
#pragma unroll
    for (int q = 0; q < 6; ++q) yt[q] = __shfl_up(F.bt[q], 1);
  }
  if (lane == 0) {
    ye = ce;
    yk = ck;
#pragma unroll
    for (int q = 0; q < 6; ++q) yt[q] = ct[q];
  }
  const bool wE = P.mask & kLEma, wT = P.mask & kLT3, wK = P.mask & kLKama, wS = P.mask & kLSma;
  if (fast) {
    // steady replay: every bar takes the step.  One family at a time, so
    // each lane writes its 16 outputs of a family as four back-to-back
    // float4 stores (whole 64-byte runs); KAMA first, so its smoothing
    // constants die early.
    float o[kLK];
    if (wK) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        const double v = (double)xv[j];
        yk = fma(v - yk, sc[j], yk);
        o[j] = (float)yk;
      }
      put16(out.kama, c, a, o);
    }
    if (wS) {
      // slides from the chunk prefix sums (all staged bars are finite here)
      SmaPrefix pr;
      sma_prefix<T>(lds, c, pr, false);
      const int e = a - (c.s0 - T::H);
      double ssum = sma_window<T>(lds, sma_ends(pr, e - P.ps + 1, e), e - P.ps + 1, e);
      const int eo = e - P.ps;  // staged index of x[a - ps]
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        if (j > 0) ssum = ssum + ((double)xv[j] - (double)lds_ld(lds, T::pos(eo + j)));
        o[j] = (float)(ssum / (double)P.ps);
      }
      put16(out.sma, c, a, o);
    }
    if (wE) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        ye = fma(P.eb, ye, P.ek * (double)xv[j]);
        o[j] = (float)ye;
      }
      put16(out.ema, c, a, o);
    }
    if (wT) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        yt[0] = fma(P.tomk, yt[0], P.tk * (double)xv[j]);
#pragma unroll
        for (int q = 1; q < 6; ++q) yt[q] = fma(P.tomk, yt[q], P.tk * yt[q - 1]);
        o[j] = (float)fma(P.c4, yt[2], fma(P.c3, yt[3], fma(P.c1, yt[5], P.c2 * yt[4])));
      }
      put16(out.t3, c, a, o);
    }
    return;
  }
  // replay with the families' arithmetic, one family at a time (each
  // family's state and outputs only are live in its loop)
  const float qn = __builtin_nanf("");
  if (wE) {
#pragma unroll
    for (int h = 0; h < kLK / 8; ++h) {
      float o[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int t = a + 8 * h + i;
        const float xf = xv[8 * h + i];
        o[i] = qn;
        if (t >= c.s1) continue;
        if (P.pe == 1) {
          o[i] = xf;
        } else if (t >= S.Le) {
          ye = (t == S.Le) ? S.se : fma(P.eb, ye, P.ek * (double)xf);
          o[i] = (float)ye;
        }
      }
      if (a + 8 * h < c.s1) put8(out.ema, c, a + 8 * h, o);
    }
  }
  if (wT) {
#pragma unroll
    for (int h = 0; h < kLK / 8; ++h) {
      float o[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int t = a + 8 * h + i;
        o[i] = qn;
        if (t >= c.s1 || t < S.L6) continue;
        if (t == S.L6) {
#pragma unroll
          for (int q = 0; q < 6; ++q) yt[q] = S.st[q];
        } else {
          yt[0] = fma(P.tomk, yt[0], P.tk * (double)xv[8 * h + i]);
#pragma unroll
          for (int q = 1; q < 6; ++q) yt[q] = fma(P.tomk, yt[q], P.tk * yt[q - 1]);
        }
        o[i] = (float)fma(P.c4, yt[2], fma(P.c3, yt[3], fma(P.c1, yt[5], P.c2 * yt[4])));
      }
      if (a + 8 * h < c.s1) put8(out.t3, c, a + 8 * h, o);
    }
  }
  if (wK) {
    KamaWin kw{0.0, -1};
#pragma unroll
    for (int h = 0; h < kLK / 8; ++h) {
      float o[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int t = a + 8 * h + i;
        const float xf = xv[8 * h + i];
        const double v = (double)xf;
        o[i] = qn;
        if (t >= c.s1) continue;
        if (P.pk == 1) {
          o[i] = xf;
        } else if (t == S.Lk) {
          yk = v;
        } else if (t > S.Lk) {
          const double sc = kama_sc<T>(lds, x, c, P, kw, t, v);
          yk = fma(v - yk, sc, yk);
          o[i] = (float)yk;
        }
      }
      if (a + 8 * h < c.s1) put8(out.kama, c, a + 8 * h, o);
    }
  }
  if (wS) {
    // SMA window of the lane: finite-part sum and counts of NaN, +Inf, -Inf
    // (the counts stay 0 on the all-finite path)
    SmaPrefix spr;
    const bool sfast = sma_prefix<T>(lds, c, spr, true);
    // the prefixes of the lane's first window (its first bar at or past the
    // SMA warm-up), read before the lanes' bar loops diverge
    SmaEnds sends{0.0, 0.0};
    if (sfast) {
      const int tf = max(a, P.ps - 1);
      const int ef = tf - (c.s0 - T::H);
      sends = sma_ends(spr, max(ef - P.ps + 1, 0), max(ef, 0));
    }
    double ssum = 0.0;
    int snan = 0, spos = 0, sneg = 0, st = -1;
    auto sadd = [&](float v, int sgn) {
      if (__builtin_isnan(v)) snan += sgn;
      else if (__builtin_isinf(v)) (v > 0.0f ? spos : sneg) += sgn;
      else ssum += sgn > 0 ? (double)v : -(double)v;
    };
#pragma unroll
    for (int h = 0; h < kLK / 8; ++h) {
      float o[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int t = a + 8 * h + i;
        const float xf = xv[8 * h + i];
        o[i] = qn;
        if (t >= c.s1 || t < P.ps - 1) continue;
        const int e = t - (c.s0 - T::H);
        if (sfast) {
          ssum = st < 0 ? sma_window<T>(lds, sends, e - P.ps + 1, e)
                        : ssum + ((double)xf - (double)lds_ld(lds, T::pos(e - P.ps)));
        } else if (st < 0) {
          for (int j = t - P.ps + 1; j <= t; ++j) sadd(xat<T>(lds, x, c, j), 1);
        } else {
          sadd(xf, 1);
          sadd(xat<T>(lds, x, c, t - P.ps), -1);
        }
        st = t;
        float r;
        if (snan > 0 || (spos > 0 && sneg > 0)) r = qn;
        else if (spos > 0) r = __builtin_inff();
        else if (sneg > 0) r = -__builtin_inff();
        else r = (float)(ssum / (double)P.ps);
        o[i] = r;
      }
      if (a + 8 * h < c.s1) put8(out.sma, c, a + 8 * h, o);
    }
  }
}

// ------------------------------------------------------------------ host --
double hpow(double b, long long e) {
  double r = 1.0, q = b;
  while (e) {
    if (e & 1) {
      r *= q;
      if (r == 0.0 && b != 0.0) r = DBL_MIN;
    }
    q *= q;
    if (q == 0.0 && b != 0.0) q = DBL_MIN;
    e >>= 1;
  }
  return r;
}

Toe toe_mul(const Toe& a, const Toe& b) {
  Toe r;
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
    for (int j = 0; j <= i; ++j) s = std::fma(a.c[j], b.c[i - j], s);
    r.c[i] = s == 0.0 ? DBL_MIN : s;  // every entry of a power is positive
  }
  return r;
}

Toe toe_pow(const Toe& m, long long e) {
  Toe r{{1.0, 0.0, 0.0, 0.0, 0.0, 0.0}}, q = m;
  bool first = true;
  while (e) {
    if (e & 1) {
      r = first ? q : toe_mul(r, q);
      first = false;
    }
    e >>= 1;
    if (e) q = toe_mul(q, q);
  }
  return r;
}

template <int H, unsigned FM>
ctStatus_t launch_scan(const float* x, int* lead, double* agg, double* carry, const ScanP& P,
                       const ScanOut& o, hipStream_t s) {
  const int segs = P.rows * P.segsPerRow;
  const int blocks = div_up(segs, kLW);
  {
    KTime kt(s, "k_scan_lead");
    hipLaunchKernelGGL(k_scan_lead, dim3(P.rows), dim3(kWave), 0, s, x, P.n, P.rows, lead);
  }
  TCA_HIP(hipGetLastError());
  if (P.useCarry) {
    {
      KTime kt(s, "k_scan_fold");
      // the fold's only look-back is KAMA's window (pk + 1 bars), not SMA's:
      // a 32-bar halo stages 1056 bars per segment instead of 1280
      if (H > 32 && P.pk + 1 <= 32)
        hipLaunchKernelGGL((k_scan_fold<32, FM>), dim3(blocks), dim3(kLW * kWave), 0, s, x, lead,
                           P, agg);
      else
        hipLaunchKernelGGL((k_scan_fold<H, FM>), dim3(blocks), dim3(kLW * kWave), 0, s, x, lead,
                           P, agg);
    }
    TCA_HIP(hipGetLastError());
    {
      KTime kt(s, "k_scan_carry");
      hipLaunchKernelGGL(k_scan_carry, dim3(P.rows), dim3(kWave), 0, s, P, agg, carry);
    }
    TCA_HIP(hipGetLastError());
  }
  {
    KTime kt(s, "k_scan_emit");
    hipLaunchKernelGGL((k_scan_emit<H, FM>), dim3(blocks), dim3(kLW * kWave), 0, s, x, lead, P,
                       carry, o);
  }
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

}  // namespace

ctStatus_t scan_run(const float* close, int seriesLen, int numSeries, const ctxScanParams* params,
                    const ctxScanOutputs* outputs, hipStream_t stream) {
  if (!close || !params || !outputs)
    return fail(CT_STATUS_INVALID_ARGUMENT, "scan: null pointer");
  if (seriesLen <= 0 || numSeries <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "scan: seriesLen and numSeries must be positive");
  ScanP P{};
  P.mask = (outputs->ema ? kLEma : 0u) | (outputs->kama ? kLKama : 0u) |
           (outputs->t3 ? kLT3 : 0u) | (outputs->sma ? kLSma : 0u);
  if (!P.mask) return fail(CT_STATUS_INVALID_ARGUMENT, "scan: no output requested");
  P.pe = (P.mask & kLEma) ? params->ema_period : 2;
  P.pk = (P.mask & kLKama) ? params->kama_period : 2;
  P.pt = (P.mask & kLT3) ? params->t3_period : 2;
  P.ps = (P.mask & kLSma) ? params->sma_period : 1;
  const int fast = (P.mask & kLKama) ? params->kama_fast : 2;
  const int slow = (P.mask & kLKama) ? params->kama_slow : 30;
  if (P.pe < 1 || P.pk < 1 || P.ps < 1 || fast < 1 || slow < 1)
    return fail(CT_STATUS_INVALID_ARGUMENT, "scan: invalid period");
  if (P.pt < 2) return fail(CT_STATUS_INVALID_ARGUMENT, "scan: T3 needs period >= 2");
  const int halo = std::max(P.ps - 1, P.pk + 1);
  if (halo > kLMaxHalo)
    return fail(CT_STATUS_UNSUPPORTED, "scan: window look-back above %d bars is not provided",
                kLMaxHalo);
  P.n = seriesLen;
  P.rows = numSeries;
  P.segsPerRow = div_up(seriesLen, kLSeg);
  if ((long long)P.segsPerRow * numSeries > (1ll << 30))
    return fail(CT_STATUS_UNSUPPORTED, "scan: batch too large");
  P.cpl = div_up(P.segsPerRow, kWave);
  P.useCarry = P.segsPerRow > 1 && (P.mask & (kLEma | kLKama | kLT3)) ? 1 : 0;
  {
    double b = (double)(P.pe - 1) / (double)(P.pe + 1);
    const double kk = 1.0 - b;
    if (b < 0.5) b = 1.0 - kk;
    P.eb = b;
    P.ek = kk;
  }
  P.kmax = 2.0 / (slow + 1.0);
  P.kdiff = 2.0 / (fast + 1.0) - P.kmax;
  P.tk = 2.0 / (P.pt + 1.0);
  P.tomk = 1.0 - P.tk;
  {
    const double v = (double)params->t3_vfactor, tmp = v * v;
    P.c1 = -(tmp * v);
    P.c2 = 3.0 * (tmp - P.c1);
    P.c3 = -6.0 * tmp - 3.0 * (v - P.c1);
    P.c4 = std::fma(3.0, tmp, std::fma(3.0, v, 1.0) - P.c1);
  }
  Toe M;
  for (int i = 0; i < 6; ++i) M.c[i] = P.tomk * std::pow(P.tk, (double)i);
  for (int s = 0; s < 6; ++s) {
    const long long L = (long long)kLK << s, R = (long long)kLSeg * P.cpl << s;
    P.ePow[s] = hpow(P.eb, L);
    P.eRow[s] = hpow(P.eb, R);
    P.tPow[s] = toe_pow(M, L);
    P.tRow[s] = toe_pow(M, R);
  }
  P.eSeg = hpow(P.eb, kLSeg);
  P.tSeg = toe_pow(M, kLSeg);
  const size_t segs = (size_t)P.segsPerRow * numSeries;
  const size_t leadB = ((size_t)numSeries * sizeof(int) + 255) & ~(size_t)255;
  const size_t bytes = leadB + segs * (9 + 8) * sizeof(double);
  Ws ws;
  ctStatus_t st = workspace_for(stream, bytes, &ws);
  if (st != CT_STATUS_SUCCESS) return st;
  int* lead = reinterpret_cast<int*>(ws.area);
  double* agg = reinterpret_cast<double*>(reinterpret_cast<char*>(ws.area) + leadB);
  double* carry = agg + segs * 9;
  const ScanOut o{outputs->ema, outputs->kama, outputs->t3, outputs->sma};
  // single-family calls run kernels compiled for that family alone
  if (P.mask == kLEma) return launch_scan<32, kLEma>(close, lead, agg, carry, P, o, stream);
  if (P.mask == kLT3) return launch_scan<32, kLT3>(close, lead, agg, carry, P, o, stream);
  if (P.mask == kLKama) {
    if (halo <= 32) return launch_scan<32, kLKama>(close, lead, agg, carry, P, o, stream);
    if (halo <= 256) return launch_scan<256, kLKama>(close, lead, agg, carry, P, o, stream);
    return launch_scan<kLMaxHalo, kLKama>(close, lead, agg, carry, P, o, stream);
  }
  if (P.mask == kLSma) {
    if (halo <= 32) return launch_scan<32, kLSma>(close, lead, agg, carry, P, o, stream);
    if (halo <= 256) return launch_scan<256, kLSma>(close, lead, agg, carry, P, o, stream);
    return launch_scan<kLMaxHalo, kLSma>(close, lead, agg, carry, P, o, stream);
  }
  if (halo <= 32) return launch_scan<32, 0>(close, lead, agg, carry, P, o, stream);
  if (halo <= 256) return launch_scan<256, 0>(close, lead, agg, carry, P, o, stream);
  return launch_scan<kLMaxHalo, 0>(close, lead, agg, carry, P, o, stream);
}

bool scan_serves(int len, int rows, int windowLookback) {
  const long long segsPerRow = (len + kLSeg - 1) / kLSeg;
  return rows >= kScanMinRows && segsPerRow >= 2 && segsPerRow * rows >= kScanMinSegments &&
         windowLookback <= kLMaxHalo;
}

}  // namespace tca

using namespace tca;

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_scan_batch_d(const float* close, int seriesLen, int numSeries,
                                         const ctxScanParams* params,
                                         const ctxScanOutputs* outputs, hipStream_t stream) {
  return guarded([&]() -> ctStatus_t {
    ctStatus_t st = scan_run(close, seriesLen, numSeries, params, outputs, stream);
    if (st != CT_STATUS_SUCCESS) return st;
    return finish_batch(stream);
  });
}

}  // extern "C"
