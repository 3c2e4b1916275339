// Fused long-series scan: EMA, KAMA, T3 and SMA of one close series per row
// from one launch sequence (no counterpart in the reference, whose families
// run as separate calls: EMA.cu, KAMA.cu, T3.cu, SMA.cu).  Results follow the
// per-family definitions (ta_EMA.c, ta_KAMA.c, ta_T3.c, the reference SMA).
//
// Two launches:
//   1. k_scan_lead -- first non-NaN bar of every row (the seeds start there);
//                     block 0 also opens the launch (epoch tag, stall flag)
//                     and writes the power tables of the EMA and T3
//                     transition maps;
//   2. k_scan_one  -- one wave per 1024-bar segment reads its bars once and
//                     folds them from a zero state; the workgroup's kLW
//                     segments (a tile of one row) combine their composites
//                     in LDS, publish the tile's composite, obtain the state
//                     entering the tile by a decoupled look-back along the
//                     row and publish the state leaving it; every wave then
//                     replays its bars from its segment's entering state
//                     with the families' own sequential arithmetic.
// Workgroups take tiles column-major (the j-th tiles of a set of rows before
// any (j+1)-th), so a workgroup only ever waits on tiles of workgroups
// started before it (scan_tile_of), and each wait is bounded (kSpinLimit).
//
// Every recurrence here starts at a seed bar with everything before it
// contributing exactly 0, so a run of L bars maps the state s to
// M^L s + o: EMA has M = beta; T3's six cascaded EMAs form one linear
// recurrence, carried in the scaled state h_q = e_q / k^(q+1) whose matrix
// is lower-triangular Toeplitz with every entry 1 - k, so its powers are
// Toeplitz too (six numbers each) and are computed on the host.  KAMA's multiplier (1 - sc_t) varies per bar and is
// carried as an (A, b) pair.  Powers of nonzero factors that underflow are
// clamped to the smallest normal double, so a carried Inf survives them as
// it does in the sequential recurrence.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "../../include/tacuda_amd.h"
#include "divide.hpp"
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
  int n, rows, segsPerRow, tilesPerRow;  // tile: kLW segments sharing a record
  unsigned mask;
  int useCarry;                  // rows span several segments (carries exist)
  int pe, pk, pt, ps;            // EMA, KAMA, T3, SMA periods
  double eb, ek;                 // EMA beta and weight (ta_EMA.c)
  double kmax, kdiff;            // KAMA constants (ta_KAMA.c)
  double tk, tomk;               // T3 weight and 1 - weight (ta_T3.c)
  double c1, c2, c3, c4;         // T3 combination of e3..e6, scaled to h3..h6
  double tsc[6];                 // 1 / tk^(q+1): e_q -> h_q
  double rps;                    // RN(1 / SMA period)
  double ePow[6], eSeg, eTile;   // beta^(K 2^s), beta^SEG, beta^(SEG kLW)
  Toe tPow[6], tSeg, tTile;      // the same powers of the T3 matrix
};

// Power tables written by k_scan_lead, kPowStride doubles an entry: entries
// 0..63 hold beta^(K l) at [0] and the first column of M^(K l) at [1..6]
// (entry 2^s is the multiplier of step s of the lane scan, entry l carries
// a segment's entering state over lanes 0..l-1); entries 64..127 hold the
// same for the tile map (kLW segments), beta^(SEG kLW l) and M^(SEG kLW l).
constexpr int kPowStride = 8;
constexpr int kPowEntries = 2 * 64;

// Launch header of the scan's workspace: the tag of this launch's record
// flags and the stall flag (both set by k_scan_lead).
struct ScanHdr {
  unsigned epoch;
  unsigned stalled;  // set when a look-back wait gave up (see scan_lookback)
  unsigned pad[62];
};

// Upper bound on the polls of one look-back wait (about a second of
// s_sleep): the segment waited on is always running, so a wait this long
// means a broken launch; the wave then carries NaN on and flags `stalled`
// instead of hanging the device, and every other waiting wave gives up at
// its next check of the flag.
constexpr int kSpinLimit = 1 << 24;

// One record per tile (kLW segments of a row, one workgroup): [0] flag (low
// word: launch epoch; high word: kAgg once [1..9] holds the tile's
// composite -- EMA offset, T3 6-vector offset, KAMA (A, b) -- then kIncl
// once [10..17] holds the state leaving the tile -- EMA, T3 6-vector,
// KAMA).  A row's first tile publishes kIncl only, so a look-back always
// ends at a kIncl record.
constexpr int kRec = 18;
constexpr unsigned kAgg = 1u, kIncl = 2u;

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

// a / b for finite a and a normal b > 0: the hardware reciprocal refined
// by two Newton steps, then the same correction; it equals the IEEE
// quotient but for rare last-bit cases.
__device__ __forceinline__ double div_nr(double a, double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  const double q = a * y;
  return fma(fma(-q, b, a), y, q);
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
  // to the scaled state h_q = e_q / tk^(q+1)
#pragma unroll
  for (int q = 0; q < 6; ++q) e[q] *= P.tsc[q];
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
        F.bt[0] = fma(P.tomk, F.bt[0], v);
#pragma unroll
        for (int q = 1; q < 6; ++q) F.bt[q] = fma(P.tomk, F.bt[q], F.bt[q - 1]);
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

__device__ __forceinline__ Toe toe_at(const double* p) {
  Toe t;
#pragma unroll
  for (int i = 0; i < 6; ++i) t.c[i] = p[i];
  return t;
}

// Inclusive lane scan of the folds (lane 0 earliest), of the requested
// recurrences only; pw is the per-lane power table.
__device__ __forceinline__ void scan_lanes(const ScanP& P, const double* __restrict__ pw, Fold& F) {
  const int lane = lane_id();
  const bool wE = P.mask & kLEma, wT = P.mask & kLT3, wK = P.mask & kLKama;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int d = 1 << s;
    if (wE) {
      const double be = __shfl_up(F.be, d);
      if (lane >= d) F.be = fma(pw[d * kPowStride], be, F.be);
    }
    if (wT) {
      double bt[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) bt[q] = __shfl_up(F.bt[q], d);
      if (lane >= d) toe_mv_add(toe_at(pw + d * kPowStride + 1), bt, F.bt);
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
__device__ __forceinline__ bool steady(const Seg& c, const ScanP& P, const Seeds& S,
                                       bool allFinite) {
  if (!allFinite || c.s1 - c.s0 != kLSeg) return false;
  if (((P.mask & kLEma) && P.pe == 1) || ((P.mask & kLKama) && P.pk == 1)) return false;
  int last = -1;
  if (P.mask & kLEma) last = max(last, S.Le);
  if (P.mask & kLT3) last = max(last, S.L6);
  if (P.mask & kLKama) last = max(last, S.Lk);
  if (P.mask & kLSma) last = max(last, P.ps - 1);
  return c.s0 > last;
}

// Stages bars [s0-H, s0+SEG) of the row into LDS, as stage() does, and
// returns whether every staged bar inside the row is finite (wave-uniform),
// tested on the loaded registers.
template <class T>
__device__ __forceinline__ bool stage_finite(float* lds, const float* src, const Seg& c) {
  const int lane = lane_id();
  const long long g0 = c.rowOff + c.s0 - T::H;
  bool bad = false;
  if (((g0 & 3) == 0) && (c.s0 - T::H >= 0) && (c.s0 + T::SEG <= c.n)) {
    constexpr int NV = (T::LEN / 4 + kWave - 1) / kWave;
    v4f v[NV];
    const v4f* p = reinterpret_cast<const v4f*>(src + g0);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int q = lane + i * kWave;
      if (q < T::LEN / 4) v[i] = p[q];
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int q = lane + i * kWave;
      if (q < T::LEN / 4) {
        *reinterpret_cast<v4f*>(lds + T::pos(4 * q)) = v[i];
        bad |= !(__builtin_isfinite(v[i].x) && __builtin_isfinite(v[i].y) &&
                 __builtin_isfinite(v[i].z) && __builtin_isfinite(v[i].w));
      }
    }
    return __ballot(bad) == 0;
  }
  const float* row = src + c.rowOff;
  constexpr int NS = (T::LEN + kWave - 1) / kWave;
  float v[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const int e = lane + i * kWave;
    const int t = c.s0 - T::H + e;
    v[i] = qnanf_();
    if (e < T::LEN && t >= 0 && t < c.n) {
      v[i] = row[t];
      bad |= !__builtin_isfinite(v[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const int e = lane + i * kWave;
    if (e < T::LEN) lds[T::pos(e)] = v[i];
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
    // the offset scaled by 1 / weight: one fma a bar
    double se = 0.0;
#pragma unroll
    for (int i = 0; i < kLK; ++i) se = fma(P.eb, se, (double)xv[i]);
    F.be = P.ek * se;
  }
  if (P.mask & kLT3) {
#pragma unroll
    for (int i = 0; i < kLK; ++i) {
      F.bt[0] = fma(P.tomk, F.bt[0], (double)xv[i]);
#pragma unroll
      for (int q = 1; q < 6; ++q) F.bt[q] = fma(P.tomk, F.bt[q], F.bt[q - 1]);
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
        er = fabs(div_nr(roc, sum));
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
__device__ Toe toe_mul_d(const Toe& a, const Toe& b) {
  Toe r;
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j <= i; ++j) s = fma(a.c[j], b.c[i - j], s);
    r.c[i] = s == 0.0 ? DBL_MIN : s;  // every entry of a power is positive
  }
  return r;
}

// First non-NaN bar of every row.  Block 0 also opens the launch: it takes
// the stream workspace's epoch as this launch's record tag (and advances
// it, so no later launch on the workspace reuses the tag), clears the stall
// flag and writes the power tables.
__global__ __launch_bounds__(64) void k_scan_lead(const float* x, ScanP P, int* lead, WsHeader* hdr,
                                                  ScanHdr* sh, double* pw) {
  const int row = blockIdx.x;
  if (row >= P.rows) return;
  const int n = P.n;
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
  if (row != 0) return;
  const int l = lane_id();
  if (l == 0) {
    const unsigned e =
        __hip_atomic_load(&hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&hdr->epoch, e + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&sh->epoch, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&sh->stalled, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  // beta^(K l) and M^(K l) by binary powering over the bits of l
  double pe = 1.0;
  Toe pt{{1.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
  bool first = true;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    if ((l >> s) & 1) {
      const double q = pe * P.ePow[s];
      pe = (q == 0.0 && P.eb != 0.0) ? DBL_MIN : q;
      pt = first ? P.tPow[s] : toe_mul_d(pt, P.tPow[s]);
      first = false;
    }
  }
  double* o = pw + l * kPowStride;
  o[0] = pe;
#pragma unroll
  for (int q = 0; q < 6; ++q) o[1 + q] = pt.c[q];
  o[7] = 0.0;
  // powers of the tile map for the look-back
  double sq = P.eTile, se = 1.0;
  Toe tq = P.tTile, ts{{1.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
  bool f1 = true;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    if ((l >> s) & 1) {
      const double q = se * sq;
      se = (q == 0.0 && P.eb != 0.0) ? DBL_MIN : q;
      ts = f1 ? tq : toe_mul_d(ts, tq);
      f1 = false;
    }
    const double q2 = sq * sq;
    sq = (q2 == 0.0 && P.eb != 0.0) ? DBL_MIN : q2;
    if (s < 5) tq = toe_mul_d(tq, tq);
  }
  double* so = pw + (kWave + l) * kPowStride;
  so[0] = se;
#pragma unroll
  for (int q = 0; q < 6; ++q) so[1 + q] = ts.c[q];
  so[7] = 0.0;
}

// Lane 0 writes n values of a record at offset off, then the flag.
__device__ __forceinline__ void rec_publish(double* rp, const double* v, int n, int off,
                                            unsigned epoch, unsigned status) {
  if (lane_id() == 0) {
    for (int i = 0; i < n; ++i) st_rel_d(rp + off + i, v[i]);
    vmem_order();
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(rp),
                       ((unsigned long long)status << 32) | epoch, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Wave sum of v over lanes (result in every lane).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) v += __shfl_xor(v, d);
  return v;
}

// State entering tile j of a row (j >= 1) from the row's records.  The
// immediate predecessor is polled first; if it has only its composite, lane
// l reads the record of tile j-1-l, so one poll covers 64 predecessors:
// once the nearest published leaving state lies among them and every newer
// tile has published its composite, the wave folds them in one pass -- the
// EMA and T3 terms weighted by powers of the tile map (sp: entry l holds
// beta^(SEG kLW l) and the first column of M^(SEG kLW l)), KAMA's by the
// exclusive product of the newer composites' multipliers -- and sums over
// lanes.  The tiles waited on belong to workgroups started before this one
// (scan_tile_of), and the oldest unfinished tile of a row finds its
// predecessor's leaving state, so the published frontier keeps moving.
__device__ __forceinline__ void scan_lookback(const double* rowRec, int j, unsigned epoch,
                                              const double* __restrict__ sp, ScanHdr* sh,
                                              double& ye, double (&yt)[6], double& yk) {
  const int lane = lane_id();
  int polls = 0;
  auto give_up = [&]() {
    ++polls;
    if (polls > kSpinLimit ||
        ((polls & 1023) == 0 &&
         __hip_atomic_load(&sh->stalled, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
      const double qn = qnan_d();
      ye = yk = qn;
#pragma unroll
      for (int r = 0; r < 6; ++r) yt[r] = qn;
      if (lane == 0)
        __hip_atomic_store(&sh->stalled, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return true;
    }
    __builtin_amdgcn_s_sleep(2);
    return false;
  };
  // the immediate predecessor first, one record per poll: usually it has
  // published at least its composite by the time this wave looks
  {
    const double* r1 = rowRec + (size_t)(j - 1) * kRec;
    unsigned long long f;
    for (;;) {
      unsigned lo = 0, hi = 0;
      if (lane == 0) {
        const unsigned long long g = ld_flag(r1);
        lo = (unsigned)g;
        hi = (unsigned)(g >> 32);
      }
      f = ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)hi) << 32) |
          (unsigned)__builtin_amdgcn_readfirstlane((int)lo);
      if ((unsigned)f == epoch) break;
      if (give_up()) return;
    }
    if ((unsigned)(f >> 32) == kIncl) {
      vmem_order();
      ye = ld_rel_d(r1 + 10);
#pragma unroll
      for (int r = 0; r < 6; ++r) yt[r] = ld_rel_d(r1 + 11 + r);
      yk = ld_rel_d(r1 + 17);
      return;
    }
  }
  const int i = j - 1 - lane;
  const double* rp = rowRec + (size_t)(i < 0 ? 0 : i) * kRec;
  int lim;
  for (;;) {
    unsigned st = 0;
    if (i >= 0) {
      const unsigned long long f = ld_flag(rp);
      if ((unsigned)f == epoch) st = (unsigned)(f >> 32);
    }
    const unsigned long long inc = __ballot(st == kIncl), rdy = __ballot(st != 0);
    lim = inc ? __builtin_ctzll(inc) : kWave;
    const unsigned long long need = (1ull << (lim & 63)) - 1ull;
    if (lim < kWave && (rdy & need) == need) break;
    if (give_up()) return;
  }
  vmem_order();
  // lane l < lim: composite of tile j-1-l; lane lim: leaving state of tile
  // j-1-lim
  double e = 0.0, k = 0.0, m = 1.0, t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (lane < lim) {
    e = ld_rel_d(rp + 1);
#pragma unroll
    for (int r = 0; r < 6; ++r) t[r] = ld_rel_d(rp + 2 + r);
    m = ld_rel_d(rp + 8);
    k = ld_rel_d(rp + 9);
  } else if (lane == lim) {
    e = ld_rel_d(rp + 10);
#pragma unroll
    for (int r = 0; r < 6; ++r) t[r] = ld_rel_d(rp + 11 + r);
    k = ld_rel_d(rp + 17);
  }
  // KAMA: weight of lane l = product of the multipliers of lanes 0..l-1
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = __shfl_up(m, d);
    if (lane >= d) m *= o;
  }
  double wx = __shfl_up(m, 1);
  if (lane == 0) wx = 1.0;
  const bool on = lane <= lim;
  const double* pl = sp + lane * kPowStride;
  ye = wave_sum(on ? pl[0] * e : 0.0);
  yk = wave_sum(on ? wx * k : 0.0);
#pragma unroll
  for (int r = 0; r < 6; ++r) yt[r] = 0.0;
  if (on) toe_mv_add(toe_at(pl + 1), t, yt);
#pragma unroll
  for (int r = 0; r < 6; ++r) yt[r] = wave_sum(yt[r]);
}

// Eight outputs of a lane at bar a + 8h (the general path: segments with a
// seed, a non-finite bar or the row end), per bar.
__device__ __forceinline__ void put8(float* dst, const Seg& c, int t0, const float (&v)[8]) {
  if (!dst) return;
  float* p = dst + c.rowOff + t0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (t0 + i < c.s1) p[i] = v[i];
}

// A steady wave's outputs (a full segment), stored coalesced: the lanes'
// 16 bars go through the staging area (no longer read) so consecutive
// lanes store consecutive float4s; lane-contiguous stores reach a fraction
// of that bandwidth.
template <class T>
__device__ __forceinline__ void put16(float* lds, float* dst, const Seg& c, const float (&v)[kLK]) {
  wave_sync();
  float* p = lds + T::pos(lane_id() * kLK);
#pragma unroll
  for (int q = 0; q < kLK / 4; ++q)
    *reinterpret_cast<float4*>(p + 4 * q) = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  wave_sync();
  float* row = dst + c.rowOff + c.s0;
  if ((reinterpret_cast<size_t>(row) & 15) == 0) {
#pragma unroll
    for (int i = 0; i < kLK / 4; ++i) {
      const int q = lane_id() + i * kWave;
      __builtin_nontemporal_store(*reinterpret_cast<const v4f*>(lds + T::pos(4 * q)),
                                  reinterpret_cast<v4f*>(row) + q);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < kLK; ++i) {
    const int e = lane_id() + i * kWave;
    __builtin_nontemporal_store(lds_ld(lds, T::pos(e)), row + e);
  }
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

// Tile of workgroup b: kLW consecutive segments of one row, one per wave,
// sharing one look-back record.  Workgroups are dealt to the XCDs
// round-robin (workgroup b runs on XCD b mod kXcd) and start in order within
// each XCD, so with at least kXcd rows every XCD owns the rows r = XCD
// (mod kXcd) and its workgroups take those rows' tiles column-major (the
// j-th tiles of its rows before any (j+1)-th): the tiles a workgroup waits
// on belong to workgroups started before it on its own XCD, and the halo it
// reads was loaded through that XCD's L2.  With fewer rows the tiles are
// taken column-major over all rows.
constexpr int kXcd = 8;
__device__ __forceinline__ bool scan_tile_of(const ScanP& P, int b, int& row, int& tile) {
  int rowsX, rowStep, rowBase, u;
  if (P.rows >= kXcd) {
    const int xcd = b % kXcd;
    rowsX = (P.rows - xcd + kXcd - 1) / kXcd;
    rowStep = kXcd;
    rowBase = xcd;
    u = b / kXcd;
  } else {
    rowsX = P.rows;
    rowStep = 1;
    rowBase = 0;
    u = b;
  }
  if ((long long)u >= (long long)rowsX * P.tilesPerRow) return false;
  tile = u / rowsX;
  row = rowBase + (u - tile * rowsX) * rowStep;
  return true;
}

// Workgroups of a launch (the grid scan_tile_of expects).
int scan_blocks(const ScanP& P) {
  if (P.rows >= kXcd) return div_up(P.rows, kXcd) * P.tilesPerRow * kXcd;
  return P.rows * P.tilesPerRow;
}

// y -> the state after a segment with composite g (EMA offset, T3 offset
// 6-vector, KAMA multiplier and offset), every segment but a row's last
// being full.
__device__ __forceinline__ void seg_apply(const ScanP& P, const double* g, double& ye,
                                          double (&yt)[6], double& yk) {
  ye = fma(P.eSeg, ye, g[0]);
  double n6[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) n6[q] = g[1 + q];
  toe_mv_add(P.tSeg, yt, n6);
#pragma unroll
  for (int q = 0; q < 6; ++q) yt[q] = n6[q];
  yk = fma(g[7], yk, g[8]);
}

template <int H, unsigned FM>
__global__ __launch_bounds__(kLW* kWave) void k_scan_one(const float* x, const int* lead, ScanP P,
                                                         ScanHdr* sh, const double* __restrict__ pw,
                                                         double* rec, ScanOut out) {
  using T = Tile<kLK, H>;
  if (FM) P.mask = FM;
  __shared__ __align__(16) float s_in[kLW][T::PLEN];
  __shared__ double s_tot[kLW][9];  // each wave's segment composite
  __shared__ double s_ent[kLW][8];  // the state entering each wave's segment
  int row, tile;
  if (!scan_tile_of(P, blockIdx.x, row, tile)) return;  // uniform over the workgroup
  const int w = threadIdx.x / kWave;
  const int sj = tile * kLW + w;
  const bool active = sj < P.segsPerRow;  // a row's last tile may be partial
  int rowChk;
  const Seg c = seg_of(P, row * P.segsPerRow + (active ? sj : P.segsPerRow - 1), rowChk);
  float* lds = s_in[w];
  const int lane = lane_id();
  const int a = c.s0 + lane * kLK;
  float xv[kLK];
  double sc[kLK];
  Fold F;
  F.be = 0.0;
#pragma unroll
  for (int q = 0; q < 6; ++q) F.bt[q] = 0.0;
  F.Ak = 1.0;
  F.bk = 0.0;
  Seeds S{};
  bool fast = false;
  if (active) {
    const bool fin = stage_finite<T>(lds, x, c);
    wave_sync();
    lane_vals<T>(lds, xv);
    S = make_seeds<T>(lds, x, c, P, lead[row]);
    fast = steady<T>(c, P, S, fin);
    F = fast ? fold_steady<T>(lds, c, P, xv, a, sc) : fold_lane<T>(lds, x, c, P, S, xv, a);
    scan_lanes(P, pw, F);
  }
  // the state entering the segment: zero for a row's first segment (all
  // contributions before the seeds are exactly 0), else carried from the
  // tile's entering state, which the look-back provides
  double ce = 0.0, ck = 0.0, ct[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (P.useCarry) {
    if (active && lane == kWave - 1) {
      s_tot[w][0] = F.be;
#pragma unroll
      for (int q = 0; q < 6; ++q) s_tot[w][1 + q] = F.bt[q];
      s_tot[w][7] = F.Ak;
      s_tot[w][8] = F.bk;
    }
    __syncthreads();
    if (w == 0) {
      const unsigned epoch =
          (unsigned)__builtin_amdgcn_readfirstlane((int)__hip_atomic_load(
              &sh->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      double* rowRec = rec + (size_t)row * P.tilesPerRow * kRec;
      double* mine = rowRec + (size_t)tile * kRec;
      const bool last = tile == P.tilesPerRow - 1;  // every earlier tile is full
      if (tile > 0) {
        if (!last) {
          double g[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
          double gt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          for (int v = 0; v < kLW; ++v) {
            seg_apply(P, s_tot[v], g[0], gt, g[8]);
            g[7] *= s_tot[v][7];
          }
#pragma unroll
          for (int q = 0; q < 6; ++q) g[1 + q] = gt[q];
          rec_publish(mine, g, 9, 1, epoch, kAgg);
        }
        scan_lookback(rowRec, tile, epoch, pw + kWave * kPowStride, sh, ce, ct, ck);
      }
      // entering states of the tile's segments, then the tile's leaving state
      for (int v = 0; v < kLW; ++v) {
        if (lane == 0) {
          s_ent[v][0] = ce;
#pragma unroll
          for (int q = 0; q < 6; ++q) s_ent[v][1 + q] = ct[q];
          s_ent[v][7] = ck;
        }
        if (tile * kLW + v < P.segsPerRow) seg_apply(P, s_tot[v], ce, ct, ck);
      }
      if (!last) {
        double leave[8];
        leave[0] = ce;
#pragma unroll
        for (int q = 0; q < 6; ++q) leave[1 + q] = ct[q];
        leave[7] = ck;
        rec_publish(mine, leave, 8, 10, epoch, kIncl);
      }
    }
    __syncthreads();
    if (!active) return;
    ce = s_ent[w][0];
#pragma unroll
    for (int q = 0; q < 6; ++q) ct[q] = s_ent[w][1 + q];
    ck = s_ent[w][7];
  } else if (!active) {
    return;
  }
  // state entering each lane: the exclusive lane prefix applied after the
  // segment's entering state carried over the earlier lanes' bars
  double ye = 0.0, yk = 0.0, yt[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  {
    const double* lp = pw + lane * kPowStride;
    if (P.mask & kLEma) {
      double ex = __shfl_up(F.be, 1);
      if (lane == 0) ex = 0.0;
      ye = fma(lp[0], ce, ex);
    }
    if (P.mask & kLKama) {
      double exA = __shfl_up(F.Ak, 1), exb = __shfl_up(F.bk, 1);
      if (lane == 0) {
        exA = 1.0;
        exb = 0.0;
      }
      yk = fma(exA, ck, exb);
    }
    if (P.mask & kLT3) {
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        yt[q] = __shfl_up(F.bt[q], 1);
        if (lane == 0) yt[q] = 0.0;
      }
      toe_mv_add(toe_at(lp + 1), ct, yt);
    }
  }
  // the lane's bars again from LDS (not held in registers across the
  // look-back)
  lane_vals<T>(lds, xv);
  const bool wE = P.mask & kLEma, wT = P.mask & kLT3, wK = P.mask & kLKama, wS = P.mask & kLSma;
  if (fast) {
    // steady replay: every bar takes the step, one family at a time.  SMA
    // first: it is the only family reading the staged bars, after which
    // the staging area carries each family's outputs to coalesced stores.
    float o[kLK];
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
        o[j] = (float)div_by(ssum, DivBy{(double)P.ps, P.rps});
      }
      put16<T>(lds, out.sma, c, o);
    }
    if (wK) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        const double v = (double)xv[j];
        yk = fma(v - yk, sc[j], yk);
        o[j] = (float)yk;
      }
      put16<T>(lds, out.kama, c, o);
    }
    if (wE) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        ye = fma(P.eb, ye, P.ek * (double)xv[j]);
        o[j] = (float)ye;
      }
      put16<T>(lds, out.ema, c, o);
    }
    if (wT) {
#pragma unroll
      for (int j = 0; j < kLK; ++j) {
        yt[0] = fma(P.tomk, yt[0], (double)xv[j]);
#pragma unroll
        for (int q = 1; q < 6; ++q) yt[q] = fma(P.tomk, yt[q], yt[q - 1]);
        o[j] = (float)fma(P.c4, yt[2], fma(P.c3, yt[3], fma(P.c1, yt[5], P.c2 * yt[4])));
      }
      put16<T>(lds, out.t3, c, o);
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
          yt[0] = fma(P.tomk, yt[0], (double)xv[8 * h + i]);
#pragma unroll
          for (int q = 1; q < 6; ++q) yt[q] = fma(P.tomk, yt[q], yt[q - 1]);
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
        else r = (float)div_by(ssum, DivBy{(double)P.ps, P.rps});
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
ctStatus_t launch_scan(const float* x, const Ws& ws, int* lead, ScanHdr* sh, double* pw,
                       double* rec, const ScanP& P, const ScanOut& o, hipStream_t s) {
  const int blocks = scan_blocks(P);
  {
    KTime kt(s, "k_scan_lead");
    hipLaunchKernelGGL(k_scan_lead, dim3(P.rows), dim3(kWave), 0, s, x, P, lead, ws.hdr, sh, pw);
  }
  TCA_HIP(hipGetLastError());
  {
    KTime kt(s, "k_scan_one");
    hipLaunchKernelGGL((k_scan_one<H, FM>), dim3(blocks), dim3(kLW * kWave), 0, s, x, lead, P,
                       sh, pw, rec, o);
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
  P.tilesPerRow = div_up(P.segsPerRow, kLW);
  if ((long long)P.segsPerRow * numSeries > (1ll << 30))
    return fail(CT_STATUS_UNSUPPORTED, "scan: batch too large");
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
  // T3 runs on the scaled cascade h_q = e_q / tk^(q+1), q = 0..5, whose step
  // is h_0 <- (1-tk) h_0 + x, h_q <- (1-tk) h_q + h_{q-1} (one fma a stage);
  // its matrix is lower-triangular Toeplitz with every entry 1 - tk
  double kq[7] = {1.0};
  for (int q = 1; q < 7; ++q) kq[q] = kq[q - 1] * P.tk;
  for (int q = 0; q < 6; ++q) P.tsc[q] = 1.0 / kq[q + 1];
  P.c1 *= kq[6];
  P.c2 *= kq[5];
  P.c3 *= kq[4];
  P.c4 *= kq[3];
  Toe M;
  for (int i = 0; i < 6; ++i) M.c[i] = P.tomk;
  for (int s = 0; s < 6; ++s) {
    const long long L = (long long)kLK << s;
    P.ePow[s] = hpow(P.eb, L);
    P.tPow[s] = toe_pow(M, L);
  }
  P.eSeg = hpow(P.eb, kLSeg);
  P.tSeg = toe_pow(M, kLSeg);
  P.eTile = hpow(P.eb, (long long)kLSeg * kLW);
  P.tTile = toe_pow(M, (long long)kLSeg * kLW);
  P.rps = 1.0 / (double)P.ps;
  // workspace: lead bars, launch header, lane powers, segment records
  const size_t tiles = (size_t)P.tilesPerRow * numSeries;
  const size_t leadB = ((size_t)numSeries * sizeof(int) + 255) & ~(size_t)255;
  const size_t powB = (((size_t)kPowEntries * kPowStride * sizeof(double)) + 255) & ~(size_t)255;
  const size_t bytes = leadB + sizeof(ScanHdr) + powB + (P.useCarry ? tiles * kRec * sizeof(double) : 0);
  Ws ws;
  ctStatus_t st = workspace_for(stream, bytes, &ws);
  if (st != CT_STATUS_SUCCESS) return st;
  char* base = reinterpret_cast<char*>(ws.area);
  int* lead = reinterpret_cast<int*>(base);
  ScanHdr* sh = reinterpret_cast<ScanHdr*>(base + leadB);
  double* pw = reinterpret_cast<double*>(base + leadB + sizeof(ScanHdr));
  double* rec = reinterpret_cast<double*>(base + leadB + sizeof(ScanHdr) + powB);
  const ScanOut o{outputs->ema, outputs->kama, outputs->t3, outputs->sma};
  // single-family calls run kernels compiled for that family alone
  if (P.mask == kLEma) return launch_scan<32, kLEma>(close, ws, lead, sh, pw, rec, P, o, stream);
  if (P.mask == kLT3) return launch_scan<32, kLT3>(close, ws, lead, sh, pw, rec, P, o, stream);
  if (P.mask == kLKama) {
    if (halo <= 32) return launch_scan<32, kLKama>(close, ws, lead, sh, pw, rec, P, o, stream);
    if (halo <= 256) return launch_scan<256, kLKama>(close, ws, lead, sh, pw, rec, P, o, stream);
    return launch_scan<kLMaxHalo, kLKama>(close, ws, lead, sh, pw, rec, P, o, stream);
  }
  if (P.mask == kLSma) {
    if (halo <= 32) return launch_scan<32, kLSma>(close, ws, lead, sh, pw, rec, P, o, stream);
    if (halo <= 256) return launch_scan<256, kLSma>(close, ws, lead, sh, pw, rec, P, o, stream);
    return launch_scan<kLMaxHalo, kLSma>(close, ws, lead, sh, pw, rec, P, o, stream);
  }
  if (halo <= 32) return launch_scan<32, 0>(close, ws, lead, sh, pw, rec, P, o, stream);
  if (halo <= 256) return launch_scan<256, 0>(close, ws, lead, sh, pw, rec, P, o, stream);
  return launch_scan<kLMaxHalo, 0>(close, ws, lead, sh, pw, rec, P, o, stream);
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
