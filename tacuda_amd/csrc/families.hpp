// Indicator families on the segment engine (engine.hpp).
//
// Every family is a functor with `run<T>(Seg&, lds_in, lds_out, params, io)`
// executed by one wave for one segment.  T is a Tile<K, H>: K bars per lane,
// H halo bars staged before the segment.  Semantics follow the reference C
// ABI (warm-up NaN, leading-NaN skip for EMA/DEMA/TEMA/T3/MACD/RSI/KAMA,
// window NaN semantics) and TA-Lib 0.8.3's arithmetic order (file:line
// cited per family).
#pragma once
#include <climits>

#include "engine.hpp"

namespace tca {

// ------------------------------------------------------------------ tiles --
template <int K_, int H_>
struct Tile {
  static constexpr int K = K_;
  static constexpr int H = H_;
  static constexpr int SEG = kWave * K;
  // one pad quad per K-block when K/4 is even keeps ds_read_b128 of
  // lane-contiguous bars free of bank conflicts
  static constexpr int PAD = ((K / 4) & 1) ? 0 : 4;
  static constexpr int LEN = SEG + H;
  static constexpr int PLEN = LEN + PAD * (LEN / K);
  static constexpr int POUT = SEG + PAD * (SEG / K);
  __device__ static __forceinline__ int pos(int e) { return e + PAD * (e / K); }
  static_assert(K % 4 == 0 && H % K == 0, "tile shape");
};

typedef float v4f __attribute__((ext_vector_type(4)));

struct IO {
  const float* in[4];
  float* out[4];
};

__device__ __forceinline__ void wave_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// Stage bars [s0-H, s0+SEG) of one input into LDS.  Every load of the wave
// is issued before the first LDS store; rows whose range is 16-byte aligned
// and fully inside the row move as float4.
template <class T>
__device__ __forceinline__ void stage(float* lds, const float* src, const Seg& c) {
  const int lane = lane_id();
  const float* row = src + c.rowOff;
  const long long g0 = c.rowOff + c.s0 - T::H;
  const bool vec = ((g0 & 3) == 0) && (c.s0 - T::H >= 0) && (c.s0 + T::SEG <= c.n);
  if (vec) {
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
        const int e = 4 * q;  // 4 | K: the quad stays inside one K-block
        *reinterpret_cast<v4f*>(lds + T::pos(e)) = v[i];
      }
    }
    return;
  }
  constexpr int NS = (T::LEN + kWave - 1) / kWave;
  float v[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const int e = lane + i * kWave;
    const int t = c.s0 - T::H + e;
    v[i] = qnanf_();
    if (e < T::LEN && t >= 0 && t < c.n) v[i] = row[t];
  }
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const int e = lane + i * kWave;
    if (e < T::LEN) lds[T::pos(e)] = v[i];
  }
}

// Load through an LDS-qualified pointer: a generic pointer that may alias
// global memory would otherwise be read with flat loads.
typedef __attribute__((address_space(3))) const float lds_cfloat;
__device__ __forceinline__ float lds_ld(const float* p, int i) {
  return ((lds_cfloat*)p)[i];
}

// Row-relative bar t (t < s0 + SEG): LDS when staged, global before that.
template <class T>
__device__ __forceinline__ float at(const float* lds, const float* src, const Seg& c, int t) {
  const int e = t - (c.s0 - T::H);
  if (e >= 0) return lds_ld(lds, T::pos(e));
  return t >= 0 ? src[c.rowOff + t] : qnanf_();
}

template <class T>
__device__ __forceinline__ void lane_vals(const float* lds, float (&v)[T::K]) {
  const float* p = lds + T::pos(T::H + lane_id() * T::K);
#pragma unroll
  for (int q = 0; q < T::K / 4; ++q) {
    const float4 f = *reinterpret_cast<const float4*>(p + 4 * q);
    v[4 * q + 0] = f.x;
    v[4 * q + 1] = f.y;
    v[4 * q + 2] = f.z;
    v[4 * q + 3] = f.w;
  }
}

// Writes the lane's K outputs.  The wave transposes its segment through
// ldo so that consecutive lanes store consecutive addresses: nontemporal
// float4 runs when the segment is full and 16-byte aligned, dwords
// otherwise.  Stores
// of K contiguous bars per lane reach a fraction of that bandwidth on
// MI355X (a quarter when nontemporal).
template <class T>
__device__ __forceinline__ void store_lane(float* ldo, const float (&v)[T::K], float* dst,
                                           const Seg& c) {
  wave_sync();
  float* p = ldo + T::pos(lane_id() * T::K);
#pragma unroll
  for (int q = 0; q < T::K / 4; ++q)
    *reinterpret_cast<float4*>(p + 4 * q) = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
  wave_sync();
  const long long g0 = c.rowOff + c.s0;
  if (((g0 & 3) == 0) && (c.s0 + T::SEG <= c.n)) {
    v4f* row = reinterpret_cast<v4f*>(dst + g0);
#pragma unroll
    for (int i = 0; i < T::K / 4; ++i) {
      const int q = lane_id() + i * kWave;
      __builtin_nontemporal_store(*reinterpret_cast<const v4f*>(ldo + T::pos(4 * q)), row + q);
    }
    return;
  }
  // rows that start or end inside a cache line share it with the next
  // row's wave: plain stores let L2 merge the two halves
  float* row = dst + c.rowOff;
  for (int e = lane_id(); e < T::SEG; e += kWave) {
    const int t = c.s0 + e;
    if (t < c.s1) row[t] = ldo[T::pos(e)];
  }
}

// Index of the first non-NaN sample of the row if it lies before s1,
// else the row length (nothing valid in or before this segment).
__device__ __forceinline__ int first_valid(const Seg& c, const float* x) {
  const float* row = x + c.rowOff;
  if (!isnan(row[0])) return 0;
  for (int base = 0; base < c.s1; base += kWave) {
    const int t = base + lane_id();
    const bool ok = t < c.s1 && !isnan(row[t]);
    const unsigned long long m = __ballot(ok);
    if (m) return base + __ffsll((long long)m) - 1;
  }
  return c.n;
}

__device__ __forceinline__ int sat_add(int a, int b) {
  const long long s = (long long)a + b;
  return s > INT_MAX ? INT_MAX : (int)s;
}

// EMA coefficients, beta first (ta_EMA.c:88-93).
__device__ __forceinline__ void ema_coef(int p, double& beta, double& k) {
  double b = (double)(p - 1) / (double)(p + 1);
  const double kk = 1.0 - b;
  if (b < 0.5) b = 1.0 - kk;
  beta = b;
  k = kk;
}

// True range, compare order of ta_ATR.c:135-150.
__device__ __forceinline__ double true_range(double h, double l, double pc) {
  double g = h - l;
  const double v2 = fabs(pc - h);
  if (v2 > g) g = v2;
  const double v3 = fabs(pc - l);
  if (v3 > g) g = v3;
  return g;
}

__device__ __forceinline__ bool is_zero14(double v) { return (-1e-14 < v) && (v < 1e-14); }

// ============================================================ EMA cascade ==
// EMA (1 level), DEMA (2), TEMA (3), T3 (6): each level is an SMA-seeded
// EMA of the level below; level j's seed bar is f + j*(p-1) + (p-1) - ... =
// f + (j+1)*(p-1).  ta_EMA.c:37-147, ta_DEMA.c:37-206, ta_TEMA.c:37-227,
// ta_T3.c:37-226.
template <class T, int NLV>
__device__ __forceinline__ void ema_cascade(const Seg& c, const float* li, const float* x, const float (&xv)[T::K],
                            int f, int p, double beta, double k, double (&e)[NLV][T::K],
                            int (&L)[NLV], int levelBase) {
  const int lane = lane_id();
  const int a = c.s0 + lane * T::K;
#pragma unroll
  for (int j = 0; j < NLV; ++j) L[j] = sat_add(f, (j + 1) * (p - 1));
  // prologue: exact TA-Lib sequence up to the last seed bar that lies here
  bool need = false;
#pragma unroll
  for (int j = 0; j < NLV; ++j) need |= (L[j] >= c.s0 && L[j] < c.s1);
  double seed[NLV];
#pragma unroll
  for (int j = 0; j < NLV; ++j) seed[j] = qnan_();
  if (need) {
    double st[NLV];
    if (lane == 0) {
      double s = 0.0;
      for (int t = f; t <= L[0]; ++t) s += (double)at<T>(li, x, c, t);
      double cur[NLV];
#pragma unroll
      for (int j = 0; j < NLV; ++j) cur[j] = 0.0;
      cur[0] = s / p;
      st[0] = cur[0];
#pragma unroll
      for (int lv = 1; lv < NLV; ++lv) {
        double tmp = 0.0;
        tmp += cur[lv - 1];
        for (int t = L[lv - 1] + 1; t <= L[lv]; ++t) {
          cur[0] = fma(beta, cur[0], k * (double)at<T>(li, x, c, t));
#pragma unroll
          for (int j = 1; j < NLV; ++j)
            if (j < lv) cur[j] = fma(beta, cur[j], k * cur[j - 1]);
          tmp += cur[lv - 1];
        }
        cur[lv] = tmp / p;
        st[lv] = cur[lv];
      }
    }
#pragma unroll
    for (int j = 0; j < NLV; ++j) seed[j] = bcast_d(lane == 0 ? st[j] : 0.0, 0);
  }
#pragma unroll
  for (int lv = 0; lv < NLV; ++lv) {
    Aff<1> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L[lv] && t < c.n) {
        const double u = lv == 0 ? (double)xv[i] : e[lv > 0 ? lv - 1 : 0][i];
        loc.b[0] = fma(beta, loc.b[0], k * u);
        loc.A[0] = decay_mul(loc.A[0], beta);
      }
    }
    double yl[1] = {seed[lv]}, yin[1];
    level_carry<1>(c, levelBase + lv, L[lv], yl, loc, yin);
    double y = yin[0];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      const double u = lv == 0 ? (double)xv[i] : e[lv > 0 ? lv - 1 : 0][i];
      if (t < L[lv]) {
        e[lv][i] = qnan_();
      } else {
        y = (t == L[lv]) ? seed[lv] : fma(beta, y, k * u);
        e[lv][i] = y;
      }
    }
  }
}

struct FEma {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 1;
  static constexpr int kMaxR = 1;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    if (pr.p == 1) {
      store_lane<T>(lo, xv, io.out[0], c);
      return;
    }
    const int f = first_valid(c, x);
    double beta, k;
    ema_coef(pr.p, beta, k);
    double e[1][T::K];
    int L[1];
    ema_cascade<T, 1>(c, li, x, xv, f, pr.p, beta, k, e, L, 0);
#pragma unroll
    for (int i = 0; i < T::K; ++i) ov[i] = (float)e[0][i];
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

struct FDema {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 2;
  static constexpr int kMaxR = 1;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    if (pr.p == 1) {
      store_lane<T>(lo, xv, io.out[0], c);
      return;
    }
    const int f = first_valid(c, x);
    double beta, k;
    ema_coef(pr.p, beta, k);
    double e[2][T::K];
    int L[2];
    ema_cascade<T, 2>(c, li, x, xv, f, pr.p, beta, k, e, L, 0);
    const int a = c.s0 + lane_id() * T::K;
#pragma unroll
    for (int i = 0; i < T::K; ++i)
      ov[i] = (a + i < L[1]) ? qnanf_() : (float)(2.0 * e[0][i] - e[1][i]);
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

struct FTema {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 3;
  static constexpr int kMaxR = 1;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    if (pr.p == 1) {
      store_lane<T>(lo, xv, io.out[0], c);
      return;
    }
    const int f = first_valid(c, x);
    double beta, k;
    ema_coef(pr.p, beta, k);
    double e[3][T::K];
    int L[3];
    ema_cascade<T, 3>(c, li, x, xv, f, pr.p, beta, k, e, L, 0);
    const int a = c.s0 + lane_id() * T::K;
#pragma unroll
    for (int i = 0; i < T::K; ++i)
      ov[i] = (a + i < L[2]) ? qnanf_() : (float)(e[2][i] + (3.0 * e[0][i] - 3.0 * e[1][i]));
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

struct FT3 {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 6;
  static constexpr int kMaxR = 1;
  struct P {
    int p;
    double v;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    const int f = first_valid(c, x);
    const double k = 2.0 / (pr.p + 1.0);
    const double omk = 1.0 - k;
    double e[6][T::K];
    int L[6];
    ema_cascade<T, 6>(c, li, x, xv, f, pr.p, omk, k, e, L, 0);
    // ta_T3.c:200-205
    const double v = pr.v, tmp = v * v;
    const double c1 = -(tmp * v);
    const double c2 = 3.0 * (tmp - c1);
    const double c3 = -6.0 * tmp - 3.0 * (v - c1);
    const double c4 = fma(3.0, tmp, fma(3.0, v, 1.0) - c1);
    const int a = c.s0 + lane_id() * T::K;
#pragma unroll
    for (int i = 0; i < T::K; ++i)
      ov[i] = (a + i < L[5]) ? qnanf_()
                             : (float)fma(c4, e[2][i], fma(c3, e[3][i], fma(c1, e[5][i], c2 * e[4][i])));
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// =================================================================== MACD ==
// ta_MACD.c:37-288 on x[f..]: level 1 = (fast, slow) EMAs seeded at
// L1 = f+slow-1 (fast over the window ending there), level 2 = signal EMA of
// the MACD line seeded at L2 = L1+signal-1.  Periods are pre-swapped.
struct FMacd {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 2;
  static constexpr int kMaxR = 2;
  struct P {
    int fast, slow, sig;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K];
    lane_vals<T>(li, xv);
    const int lane = lane_id();
    const int a = c.s0 + lane * T::K;
    const int f = first_valid(c, x);
    const int fast = pr.fast, slow = pr.slow, sig = pr.sig;
    const int L1 = sat_add(f, slow - 1), L2 = sat_add(L1, sig - 1);
    double fb, fk, sb, sk, gb, gk;
    ema_coef(fast, fb, fk);
    ema_coef(slow, sb, sk);
    ema_coef(sig, gb, gk);
    double s1[2] = {qnan_(), qnan_()}, s2[1] = {qnan_()};
    if ((L1 >= c.s0 && L1 < c.s1) || (L2 >= c.s0 && L2 < c.s1)) {
      double pf = 0, ps = 0, pg = 0;
      if (lane == 0) {
        double t0 = 0.0;
        for (int t = f; t <= L1; ++t) t0 += (double)at<T>(li, x, c, t);
        ps = t0 / slow;
        double t1 = 0.0;
        for (int t = L1 - fast + 1; t <= L1; ++t) t1 += (double)at<T>(li, x, c, t);
        pf = t1 / fast;
        s1[0] = pf;
        s1[1] = ps;
        double mv = pf - ps;
        pg = 0.0;
        pg += mv;
        for (int t = L1 + 1; t <= L2; ++t) {
          const double v = (double)at<T>(li, x, c, t);
          pf = fma(fb, pf, fk * v);
          ps = fma(sb, ps, sk * v);
          mv = pf - ps;
          pg += mv;
        }
        pg = pg / sig;
      }
      s1[0] = bcast_d(lane == 0 ? s1[0] : 0.0, 0);
      s1[1] = bcast_d(lane == 0 ? s1[1] : 0.0, 0);
      s2[0] = bcast_d(pg, 0);
    }
    // level 1: fast & slow
    Aff<2> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L1 && t < c.n) {
        const double v = (double)xv[i];
        loc.b[0] = fma(fb, loc.b[0], fk * v);
        loc.A[0] = decay_mul(loc.A[0], fb);
        loc.b[1] = fma(sb, loc.b[1], sk * v);
        loc.A[1] = decay_mul(loc.A[1], sb);
      }
    }
    double yin[2];
    level_carry<2>(c, 0, L1, s1, loc, yin);
    double mv[T::K];
    double pf = yin[0], ps = yin[1];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L1) {
        mv[i] = qnan_();
      } else {
        if (t == L1) {
          pf = s1[0];
          ps = s1[1];
        } else {
          const double v = (double)xv[i];
          pf = fma(fb, pf, fk * v);
          ps = fma(sb, ps, sk * v);
        }
        mv[i] = pf - ps;
      }
    }
    // level 2: signal
    Aff<1> l2;
    aff_id(l2);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L2 && t < c.n) {
        l2.b[0] = fma(gb, l2.b[0], gk * mv[i]);
        l2.A[0] = decay_mul(l2.A[0], gb);
      }
    }
    double yin2[1];
    level_carry<1>(c, 1, L2, s2, l2, yin2);
    double g = yin2[0];
    float om[T::K], os[T::K], oh[T::K];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L2) {
        om[i] = os[i] = oh[i] = qnanf_();
      } else {
        if (t == L2)
          g = s2[0];
        else
          g = (sig == 1) ? mv[i] : fma(gb, g, gk * mv[i]);
        om[i] = (float)mv[i];
        os[i] = (float)g;
        oh[i] = (float)(mv[i] - g);
      }
    }
    store_lane<T>(lo, om, io.out[0], c);
    store_lane<T>(lo, os, io.out[1], c);
    store_lane<T>(lo, oh, io.out[2], c);
  }
};

// ==================================================================== RSI ==
// ta_RSI.c:36-200 on x[f..]: seed at L = f+p, update (avg*(p-1)+d)*inv.
struct FRsi {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 1;
  static constexpr int kMaxR = 2;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    const int lane = lane_id();
    const int a = c.s0 + lane * T::K;
    const int p = pr.p;
    const int f = first_valid(c, x);
    const int L = sat_add(f, p);
    const double inv = 1.0 / (double)p;
    const double pm1 = (double)(p - 1);
    double seed[2] = {qnan_(), qnan_()};
    if (L >= c.s0 && L < c.s1) {
      double g = 0.0, l = 0.0;
      if (lane == 0) {
        double prev = (double)at<T>(li, x, c, f);
        for (int t = f + 1; t <= L; ++t) {
          const double v = (double)at<T>(li, x, c, t);
          const double d = v - prev;
          prev = v;
          const double gd = d > 0.0 ? d : 0.0;
          g += gd;
          l += gd - d;
        }
        l *= inv;
        g *= inv;
      }
      seed[0] = bcast_d(g, 0);
      seed[1] = bcast_d(l, 0);
    }
    double prevx = (double)at<T>(li, x, c, a - 1);
    double dd[T::K];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const double v = (double)xv[i];
      dd[i] = v - prevx;
      prevx = v;
    }
    Aff<2> loc;
    aff_id(loc);
    const double am = pm1 * inv;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L && t < c.n) {
        const double d = dd[i];
        const double gd = d > 0.0 ? d : 0.0;
        loc.b[0] = (loc.b[0] * pm1 + gd) * inv;
        loc.b[1] = (loc.b[1] * pm1 + (gd - d)) * inv;
        loc.A[0] = decay_mul(loc.A[0], am);
        loc.A[1] = decay_mul(loc.A[1], am);
      }
    }
    double yin[2];
    level_carry<2>(c, 0, L, seed, loc, yin);
    double g = yin[0], l = yin[1];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L) {
        ov[i] = qnanf_();
        continue;
      }
      if (t == L) {
        g = seed[0];
        l = seed[1];
      } else {
        const double d = dd[i];
        l *= pm1;
        g *= pm1;
        const double gd = d > 0.0 ? d : 0.0;
        g += gd;
        l += gd - d;
        l *= inv;
        g *= inv;
      }
      const double s = g + l;
      ov[i] = (float)(s > 0.0 ? 100.0 * (g / s) : 50.0);
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// ============================================================ ATR / NATR ==
// ta_ATR.c:37-222: seed mean TR[1..p] at L = p; fma(wBeta, prev, wAlpha*TR).
template <bool NORM>
struct FAtrT {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 1;
  static constexpr int kMaxR = 1;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    float* lc = li + 2 * T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    stage<T>(lc, io.in[2], c);
    wave_sync();
    float hv[T::K], lv[T::K], cv[T::K], ov[T::K];
    lane_vals<T>(lh, hv);
    lane_vals<T>(ll, lv);
    lane_vals<T>(lc, cv);
    const int lane = lane_id();
    const int a = c.s0 + lane * T::K;
    const int p = pr.p;
    if (p == 1) {
      // period 1 is TRANGE for ATR and NATR alike (unnormalized; NaN at bar
      // 0; recovers after a NaN bar): reference docs/BATCH_API.md:155-159
      double pc1 = (double)at<T>(lc, io.in[2], c, a - 1);
#pragma unroll
      for (int i = 0; i < T::K; ++i) {
        ov[i] = (a + i < 1) ? qnanf_() : (float)true_range(hv[i], lv[i], pc1);
        pc1 = cv[i];
      }
      store_lane<T>(lo, ov, io.out[0], c);
      return;
    }
    const int L = p;
    const double wb = (double)(p - 1) / (double)p;
    const double wa = 1.0 - wb;
    double seed[1] = {qnan_()};
    if (L >= c.s0 && L < c.s1) {
      double tot = 0.0;
      if (lane == 0) {
        for (int t = 1; t <= p; ++t)
          tot += true_range(at<T>(lh, io.in[0], c, t), at<T>(ll, io.in[1], c, t),
                            at<T>(lc, io.in[2], c, t - 1));
        tot = tot / p;
      }
      seed[0] = bcast_d(tot, 0);
    }
    double tr[T::K];
    double pc = (double)at<T>(lc, io.in[2], c, a - 1);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      tr[i] = true_range(hv[i], lv[i], pc);
      pc = (double)cv[i];
    }
    Aff<1> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L && t < c.n) {
        loc.b[0] = fma(wb, loc.b[0], wa * tr[i]);
        loc.A[0] = decay_mul(loc.A[0], wb);
      }
    }
    double yin[1];
    level_carry<1>(c, 0, L, seed, loc, yin);
    double y = yin[0];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L) {
        ov[i] = qnanf_();
        continue;
      }
      y = (t == L) ? seed[0] : fma(wb, y, wa * tr[i]);
      if (NORM) {
        const double cl = (double)cv[i];
        ov[i] = is_zero14(cl) ? 0.0f : (float)((y / cl) * 100.0);
      } else {
        ov[i] = (float)y;
      }
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};
using FAtr = FAtrT<false>;
using FNatr = FAtrT<true>;

// ============================================================ directional ==
// +DM/-DM/+DI/-DI/DX/ADX: Wilder sums seeded over bars 1..p-1 at L1 = p-1,
// update prev - prev/p + x; DI = 100*DM/TR; DX from the DIs; ADX = mean of
// the first p DX values at L2 = 2p-1, then (adx*(p-1)+dx)/p.
enum DirKind { kPDM = 0, kMDM = 1, kPDI = 2, kMDI = 3, kDX = 4, kADX = 5 };

__device__ __forceinline__ void dm_pair(double hp, double h, double lp, double l, double& pd,
                                        double& md) {
  const double up = h - hp;
  const double dn = lp - l;
  pd = (up > 0.0 && up > dn) ? up : 0.0;
  md = (dn > 0.0 && dn > up) ? dn : 0.0;
}

__device__ __forceinline__ void dir_outputs(double sp, double sm, double st, double& pdi, double& mdi,
                                            double& dx) {
  pdi = is_zero14(st) ? 0.0 : 100.0 * (sp / st);
  mdi = is_zero14(st) ? 0.0 : 100.0 * (sm / st);
  const double s = pdi + mdi;
  dx = is_zero14(s) ? 0.0 : 100.0 * (fabs(mdi - pdi) / s);
}

template <int KIND>
struct FDir {
  static constexpr bool kScan = true;
  static constexpr int kLevels = (KIND == kADX) ? 2 : 1;
  static constexpr int kMaxR = ((KIND == kPDM || KIND == kMDM) ? 1 : 3);
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    constexpr bool needTR = KIND >= kPDI;
    float* lh = li;
    float* ll = li + T::PLEN;
    float* lc = li + 2 * T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    if (needTR) stage<T>(lc, io.in[2], c);
    wave_sync();
    float hv[T::K], lv[T::K], cv[T::K], ov[T::K];
    lane_vals<T>(lh, hv);
    lane_vals<T>(ll, lv);
    if (needTR) lane_vals<T>(lc, cv);
    const int lane = lane_id();
    const int a = c.s0 + lane * T::K;
    const int p = pr.p;
    double pdm[T::K], mdm[T::K], tr[T::K];
    {
      double hp = (double)at<T>(lh, io.in[0], c, a - 1);
      double lp = (double)at<T>(ll, io.in[1], c, a - 1);
      double cp = needTR ? (double)at<T>(lc, io.in[2], c, a - 1) : 0.0;
#pragma unroll
      for (int i = 0; i < T::K; ++i) {
        dm_pair(hp, hv[i], lp, lv[i], pdm[i], mdm[i]);
        tr[i] = needTR ? true_range(hv[i], lv[i], cp) : 0.0;
        hp = hv[i];
        lp = lv[i];
        if (needTR) cp = cv[i];
      }
    }
    if (p == 1) {
      // period-one contracts: raw values from bar 1
#pragma unroll
      for (int i = 0; i < T::K; ++i) {
        const int t = a + i;
        if (t < 1) {
          ov[i] = qnanf_();
          continue;
        }
        double pdi, mdi, dx;
        dir_outputs(pdm[i], mdm[i], tr[i], pdi, mdi, dx);
        const double r = KIND == kPDM ? pdm[i] : KIND == kMDM ? mdm[i] : KIND == kPDI ? pdi
                       : KIND == kMDI ? mdi : dx;
        ov[i] = (float)r;
      }
      store_lane<T>(lo, ov, io.out[0], c);
      return;
    }
    const double pp = (double)p;
    const int L1 = p - 1;
    const int L2 = 2 * p - 1;
    const double am = 1.0 - 1.0 / pp;
    // prologue: sums over bars 1..p-1 and, for ADX, the sequential run to 2p-1
    double s1[3] = {qnan_(), qnan_(), qnan_()}, s2[1] = {qnan_()};
    const bool in1 = L1 >= c.s0 && L1 < c.s1;
    const bool in2 = KIND == kADX && L2 >= c.s0 && L2 < c.s1;
    if (in1 || in2) {
      double sp = 0, sm = 0, st = 0, adx = 0;
      if (lane == 0) {
        for (int t = 1; t <= L1; ++t) {
          double pd, md;
          dm_pair(at<T>(lh, io.in[0], c, t - 1), at<T>(lh, io.in[0], c, t),
                  at<T>(ll, io.in[1], c, t - 1), at<T>(ll, io.in[1], c, t), pd, md);
          sp += pd;
          sm += md;
          if (needTR)
            st += true_range(at<T>(lh, io.in[0], c, t), at<T>(ll, io.in[1], c, t),
                             at<T>(lc, io.in[2], c, t - 1));
        }
        s1[0] = KIND == kMDM ? sm : sp;
        s1[1] = sm;
        s1[2] = st;
        if (KIND == kADX) {
          double sum = 0.0;
          for (int t = p; t <= L2; ++t) {
            double pd, md;
            dm_pair(at<T>(lh, io.in[0], c, t - 1), at<T>(lh, io.in[0], c, t),
                    at<T>(ll, io.in[1], c, t - 1), at<T>(ll, io.in[1], c, t), pd, md);
            sp = sp - sp / pp + pd;
            sm = sm - sm / pp + md;
            st = st - st / pp +
                 true_range(at<T>(lh, io.in[0], c, t), at<T>(ll, io.in[1], c, t),
                            at<T>(lc, io.in[2], c, t - 1));
            double pdi, mdi, dx;
            dir_outputs(sp, sm, st, pdi, mdi, dx);
            sum += dx;
          }
          adx = sum / pp;
        }
      }
#pragma unroll
      for (int r = 0; r < 3; ++r) s1[r] = bcast_d(lane == 0 ? s1[r] : 0.0, 0);
      s2[0] = bcast_d(adx, 0);
    }
    constexpr int R = (KIND == kPDM || KIND == kMDM) ? 1 : 3;
    Aff<R> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L1 && t < c.n) {
        if (R == 1) {
          const double u = KIND == kMDM ? mdm[i] : pdm[i];
          loc.b[0] = loc.b[0] - loc.b[0] / pp + u;
          loc.A[0] = decay_mul(loc.A[0], am);
        } else {
          const double u[3] = {pdm[i], mdm[i], tr[i]};
#pragma unroll
          for (int r = 0; r < R; ++r) {
            loc.b[r] = loc.b[r] - loc.b[r] / pp + u[r];
            loc.A[r] = decay_mul(loc.A[r], am);
          }
        }
      }
    }
    double seedR[R], yin[R];
#pragma unroll
    for (int r = 0; r < R; ++r) seedR[r] = s1[R == 1 ? 0 : r];
    level_carry<R>(c, 0, L1, seedR, loc, yin);
    double st[R];
#pragma unroll
    for (int r = 0; r < R; ++r) st[r] = yin[r];
    double dxv[T::K];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      dxv[i] = qnan_();
      if (t < L1) {
        ov[i] = qnanf_();
        continue;
      }
      if (t == L1) {
#pragma unroll
        for (int r = 0; r < R; ++r) st[r] = seedR[r];
      } else {
        if (R == 1) {
          const double u = KIND == kMDM ? mdm[i] : pdm[i];
          st[0] = st[0] - st[0] / pp + u;
        } else {
          const double u[3] = {pdm[i], mdm[i], tr[i]};
#pragma unroll
          for (int r = 0; r < R; ++r) st[r] = st[r] - st[r] / pp + u[r];
        }
      }
      if (R == 1) {
        ov[i] = (float)st[0];
      } else {
        double pdi, mdi, dx;
        dir_outputs(st[0], st[1 % R], st[2 % R], pdi, mdi, dx);
        dxv[i] = dx;
        const double r = KIND == kPDI ? pdi : KIND == kMDI ? mdi : dx;
        ov[i] = (t == L1) ? qnanf_() : (float)r;
      }
    }
    if (KIND != kADX) {
      store_lane<T>(lo, ov, io.out[0], c);
      return;
    }
    Aff<1> l2;
    aff_id(l2);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L2 && t < c.n) {
        l2.b[0] = ((l2.b[0] * (pp - 1.0)) + dxv[i]) / pp;
        l2.A[0] = decay_mul(l2.A[0], (pp - 1.0) / pp);
      }
    }
    double yin2[1];
    level_carry<1>(c, 1, L2, s2, l2, yin2);
    double adx = yin2[0];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L2) {
        ov[i] = qnanf_();
        continue;
      }
      adx = (t == L2) ? s2[0] : ((adx * (pp - 1.0)) + dxv[i]) / pp;
      ov[i] = (float)adx;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// =================================================================== KAMA ==
// ta_KAMA.c:37-297 on x[f..]: first KAMA at L = f+p from prev = x[L-1];
// sc = (ER*constDiff + constMax)^2, y = fma(x - y, sc, y).
struct FKama {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 1;
  static constexpr int kMaxR = 1;
  struct P {
    int p, fast, slow;
  };
  template <class T>
  __device__ static double sc_at(const float* li, const float* x, const Seg& c, int t, int p,
                                 double sum, double cdiff, double cmax) {
    const double roc = (double)at<T>(li, x, c, t) - (double)at<T>(li, x, c, t - p);
    double er;
    if (sum <= 0.0 || sum <= roc) {
      er = 1.0;
    } else {
      er = fabs(roc / sum);
      if (er > 1.0) er = 1.0;
    }
    double sc = fma(er, cdiff, cmax);
    return sc * sc;
  }
  template <class T>
  __device__ static double wsum(const float* li, const float* x, const Seg& c, int t, int p) {
    double s = 0.0;
    for (int j = t - p + 1; j <= t; ++j)
      s += fabs((double)at<T>(li, x, c, j) - (double)at<T>(li, x, c, j - 1));
    return s;
  }
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    if (pr.p == 1) {
      store_lane<T>(lo, xv, io.out[0], c);
      return;
    }
    const int lane = lane_id();
    const int a = c.s0 + lane * T::K;
    const int p = pr.p;
    const int f = first_valid(c, x);
    const int L = sat_add(f, p);
    const double cmax = 2.0 / (pr.slow + 1.0);
    const double cdiff = 2.0 / (pr.fast + 1.0) - cmax;
    double seed[1] = {qnan_()};
    if (L >= c.s0 && L < c.s1) {
      double y = 0.0;
      if (lane == 0) {
        const double prev = (double)at<T>(li, x, c, L - 1);
        const double sc = sc_at<T>(li, x, c, L, p, wsum<T>(li, x, c, L, p), cdiff, cmax);
        y = fma((double)at<T>(li, x, c, L) - prev, sc, prev);
      }
      seed[0] = bcast_d(y, 0);
    }
    // smoothing constants of the lane's active bars (window sums slide)
    double sc[T::K];
    double sum = 0.0;
    bool have = false;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      sc[i] = 0.0;
      if (t > L && t < c.n) {
        if (!have) {
          sum = wsum<T>(li, x, c, t, p);
          have = true;
        } else {
          const double oldd =
              fabs((double)at<T>(li, x, c, t - p) - (double)at<T>(li, x, c, t - p - 1));
          const double newd = fabs((double)xv[i] - (double)at<T>(li, x, c, t - 1));
          sum -= oldd;
          sum += newd;
        }
        sc[i] = sc_at<T>(li, x, c, t, p, sum, cdiff, cmax);
      }
    }
    Aff<1> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t > L && t < c.n) {
        loc.b[0] = fma((double)xv[i] - loc.b[0], sc[i], loc.b[0]);
        loc.A[0] = decay_mul(loc.A[0], (1.0 - sc[i]));
      }
    }
    double yin[1];
    level_carry<1>(c, 0, L, seed, loc, yin);
    double y = yin[0];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < L) {
        ov[i] = qnanf_();
        continue;
      }
      y = (t == L) ? seed[0] : fma((double)xv[i] - y, sc[i], y);
      ov[i] = (float)y;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// ================================================================ windows ==
// Direct evaluation of the window ending at each bar (window NaN semantics).

// SMA / SUM: exact double window sums slide inside a lane; a non-finite
// value entering or leaving forces a direct re-sum (ta_SMA.c:36-117).
template <bool MEAN>
struct FSumT {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    const int a = c.s0 + lane_id() * T::K;
    const int p = pr.p;
    double s = 0.0;
    bool fresh = true;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < p - 1 || t >= c.n) {
        ov[i] = qnanf_();
        fresh = true;
        continue;
      }
      const float xn = xv[i];
      const float xo = (t - p >= 0) ? at<T>(li, x, c, t - p) : 0.0f;
      if (fresh || !isfinite(xn) || !isfinite(xo) || !isfinite(s)) {
        s = 0.0;
        for (int j = t - p + 1; j <= t; ++j) s += (double)at<T>(li, x, c, j);
        fresh = false;
      } else {
        s += (double)xn;
        s -= (double)xo;
      }
      ov[i] = MEAN ? (float)(s / (double)p) : (float)s;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};
using FSma = FSumT<true>;
using FSum = FSumT<false>;

// WMA: weights 1..p, divider p(p+1)/2 (ta_WMA.c:36-244).
struct FWma {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float xv[T::K], ov[T::K];
    lane_vals<T>(li, xv);
    const int a = c.s0 + lane_id() * T::K;
    const int p = pr.p;
    if (p == 1) {
      store_lane<T>(lo, xv, io.out[0], c);
      return;
    }
    const double divider = (double)p * (p + 1) / 2.0;
    double sw = 0.0, sp = 0.0;
    bool fresh = true;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < p - 1 || t >= c.n) {
        ov[i] = qnanf_();
        fresh = true;
        continue;
      }
      const float xn = xv[i];
      const float xo = (t - p >= 0) ? at<T>(li, x, c, t - p) : 0.0f;
      if (fresh || !isfinite(xn) || !isfinite(xo) || !isfinite(sw) || !isfinite(sp)) {
        sw = 0.0;
        sp = 0.0;
        int w = 1;
        for (int j = t - p + 1; j <= t; ++j, ++w) {
          const double v = (double)at<T>(li, x, c, j);
          sp += v;
          sw += v * w;
        }
        fresh = false;
      } else {
        // slide: drop one copy of every element, add the new one with weight p
        sw -= sp;
        sw += (double)xn * p;
        sp += (double)xn;
        sp -= (double)xo;
      }
      ov[i] = (float)(sw / divider);
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// Population variance of the window shifted by its oldest sample, floored
// like ta_VAR (exactly 0 on a constant window).
template <class T>
__device__ __forceinline__ double window_var(const float* li, const float* x, const Seg& c, int t,
                                             int p, double* meanOut) {
  const double shift = (double)at<T>(li, x, c, t - p + 1);
  const double inv = 1.0 / (double)p;
  double s1 = 0.0, s2 = 0.0, sm = 0.0;
  for (int j = t - p + 1; j <= t; ++j) {
    const double v = (double)at<T>(li, x, c, j);
    const double d = v - shift;
    sm += v;
    s1 += d;
    s2 += d * d;
  }
  const double mean = s1 * inv;
  const double e2 = s2 * inv;
  double var = e2 - mean * mean;
  if (var < 1e-12 * e2) var = 0.0;
  if (!(var >= 0.0) && !isnan(var)) var = 0.0;
  if (meanOut) *meanOut = sm / p;
  return var;
}

// window_var of the lane's bars a+i for i in [i0, i0+IG): the IG windows are
// evaluated side by side (independent accumulator chains, each window in
// the same order as window_var).  Every bar of the windows is staged.
template <class T, int IG>
__device__ __forceinline__ void window_var_group(const float* li, const Seg& c, int a, int i0, int p,
                                                 double (&var)[IG], double (&mean)[IG]) {
  const int e0 = a + i0 - p + 1 - (c.s0 - T::H);  // staged index of the first window's oldest bar
  const double inv = 1.0 / (double)p;
  double shift[IG], s1[IG], s2[IG], sm[IG];
#pragma unroll
  for (int u = 0; u < IG; ++u) {
    shift[u] = (double)lds_ld(li, T::pos(e0 + u));
    s1[u] = s2[u] = sm[u] = 0.0;
  }
  for (int k = 0; k < p; ++k) {
#pragma unroll
    for (int u = 0; u < IG; ++u) {
      const double v = (double)lds_ld(li, T::pos(e0 + u + k));
      const double d = v - shift[u];
      sm[u] += v;
      s1[u] += d;
      s2[u] += d * d;
    }
  }
#pragma unroll
  for (int u = 0; u < IG; ++u) {
    const double m = s1[u] * inv;
    const double e2 = s2[u] * inv;
    double v = e2 - m * m;
    if (v < 1e-12 * e2) v = 0.0;
    if (!(v >= 0.0) && !isnan(v)) v = 0.0;
    var[u] = v;
    mean[u] = sm[u] / p;
  }
}

// Whether every window the segment outputs lies inside the staged range:
// bars before the row (staged as NaN) only enter windows whose output is
// NaN anyway, and bars past the row are never output.
template <class T>
__device__ __forceinline__ bool windows_staged(int p) {
  return p - 1 <= T::H;
}

template <int MODE>  // 0 = VAR, 1 = STDDEV
struct FVarT {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float ov[T::K];
    const int a = c.s0 + lane_id() * T::K;
    constexpr int IG = 4;
    static_assert(T::K % IG == 0, "bar groups");
    const bool grouped = windows_staged<T>(pr.p);
#pragma unroll
    for (int i0 = 0; i0 < T::K; i0 += IG) {
      double gv[IG] = {}, gm[IG] = {};
      if (grouped) window_var_group<T, IG>(li, c, a, i0, pr.p, gv, gm);
#pragma unroll
      for (int u = 0; u < IG; ++u) {
        const int i = i0 + u, t = a + i;
        if (t < pr.p - 1 || t >= c.n) {
          ov[i] = qnanf_();
          continue;
        }
        const double v = grouped ? gv[u] : window_var<T>(li, x, c, t, pr.p, nullptr);
        ov[i] = MODE == 0 ? (float)v : (float)sqrt(v);
      }
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};
using FVar = FVarT<0>;
using FStddev = FVarT<1>;

// BBANDS with an SMA middle band (ta_BBANDS.c:37-397).
struct FBbands {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
    double up, dn;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float ou[T::K], om[T::K], ol[T::K];
    const int a = c.s0 + lane_id() * T::K;
    constexpr int IG = 4;
    static_assert(T::K % IG == 0, "bar groups");
    const bool grouped = windows_staged<T>(pr.p);
#pragma unroll
    for (int i0 = 0; i0 < T::K; i0 += IG) {
      double gv[IG] = {}, gm[IG] = {};
      if (grouped) window_var_group<T, IG>(li, c, a, i0, pr.p, gv, gm);
#pragma unroll
    for (int u = 0; u < IG; ++u) {
      const int i = i0 + u, t = a + i;
      if (t < pr.p - 1 || t >= c.n) {
        ou[i] = om[i] = ol[i] = qnanf_();
        continue;
      }
      double m = gm[u];
      const double var = grouped ? gv[u] : window_var<T>(li, x, c, t, pr.p, &m);
      const double sd = (var != 0.0) ? sqrt(var) : 0.0;
      om[i] = (float)m;
      if (pr.up == pr.dn) {
        const double d = sd * pr.up;
        ou[i] = (float)(m + d);
        ol[i] = (float)(m - d);
      } else {
        ou[i] = (float)fma(sd, pr.up, m);
        ol[i] = (float)(m - sd * pr.dn);
      }
    }
    }
    store_lane<T>(lo, ou, io.out[0], c);
    store_lane<T>(lo, om, io.out[1], c);
    store_lane<T>(lo, ol, io.out[2], c);
  }
};

// MIN / MAX: NaN-ignoring window extremum; a window that starts inside the
// leading NaN prefix is NaN (reference MAX.cu:12-30, LeadingNaN.cuh:22-28).
template <bool ISMAX>
struct FExtT {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    const float* x = io.in[0];
    stage<T>(li, x, c);
    wave_sync();
    float ov[T::K];
    const int a = c.s0 + lane_id() * T::K;
    const int p = pr.p;
    // leading NaN prefix length, bounded by this segment's reach
    const int f = first_valid(c, x);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < p - 1 || t >= c.n) {
        ov[i] = qnanf_();
        continue;
      }
      const int s = t - p + 1;
      float m = at<T>(li, x, c, s);
      for (int j = s + 1; j <= t; ++j) {
        const float v = at<T>(li, x, c, j);
        m = ISMAX ? fmaxf(m, v) : fminf(m, v);
      }
      if (!isnan(m) && s < f) m = qnanf_();
      ov[i] = m;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};
using FMax = FExtT<true>;
using FMin = FExtT<false>;

template <class T>
__device__ __forceinline__ void window_hl(const float* lh, const float* ll, const IO& io,
                                          const Seg& c, int t, int p, float& hh, float& lw) {
  float hi = at<T>(lh, io.in[0], c, t), lo = at<T>(ll, io.in[1], c, t);
  for (int j = 1; j < p; ++j) {
    const float a = at<T>(lh, io.in[0], c, t - j), b = at<T>(ll, io.in[1], c, t - j);
    if (a > hi) hi = a;
    if (b < lo) lo = b;
  }
  hh = hi;
  lw = lo;
}

// WILLR (reference WILLR.cu:12-30): float extrema, double ratio, clamp.
struct FWillr {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    wave_sync();
    float ov[T::K];
    const int a = c.s0 + lane_id() * T::K;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < pr.p - 1 || t >= c.n) {
        ov[i] = qnanf_();
        continue;
      }
      float hh, lw;
      window_hl<T>(lh, ll, io, c, t, pr.p, hh, lw);
      const double den = (double)hh - (double)lw;
      const double cl = (double)io.in[2][c.rowOff + t];
      const double v = den == 0.0 ? 0.0 : (((double)hh - cl) / den) * -100.0;
      ov[i] = isnan(v) ? qnanf_() : (float)fmin(0.0, fmax(-100.0, v));
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// CCI (reference CCI.cu:11-37): typical price ((H+L)+C)/3 in double.
struct FCci {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ static double tp(const float* lh, const float* ll, const float* lc, const IO& io,
                              const Seg& c, int t) {
    return ((double)at<T>(lh, io.in[0], c, t) + (double)at<T>(ll, io.in[1], c, t) +
            (double)at<T>(lc, io.in[2], c, t)) /
           3.0;
  }
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    float* lc = li + 2 * T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    stage<T>(lc, io.in[2], c);
    wave_sync();
    float ov[T::K];
    const int a = c.s0 + lane_id() * T::K;
    const int p = pr.p;
    if (windows_staged<T>(p)) {
      // every window is staged: the typical price of each staged bar once,
      // in doubles over the high/low area (2 PLEN floats hold LEN doubles)
      constexpr int NE = (T::LEN + kWave - 1) / kWave;
      double tv[NE];
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        const int e = lane_id() + q * kWave;
        tv[q] = 0.0;
        if (e < T::LEN)
          tv[q] = ((double)lds_ld(lh, T::pos(e)) + (double)lds_ld(ll, T::pos(e)) +
                   (double)lds_ld(lc, T::pos(e))) /
                  3.0;
      }
      wave_sync();
      double* tpd = reinterpret_cast<double*>(lh);
#pragma unroll
      for (int q = 0; q < NE; ++q) {
        const int e = lane_id() + q * kWave;
        if (e < T::LEN) tpd[e] = tv[q];
      }
      wave_sync();
      const int e0 = a - (c.s0 - T::H);  // staged index of bar a
#pragma unroll
      for (int i = 0; i < T::K; ++i) {
        const int t = a + i, e = e0 + i;
        if (t < p - 1 || t >= c.n) {
          ov[i] = qnanf_();
          continue;
        }
        double s = 0.0;
        for (int j = 0; j < p; ++j) s += tpd[e - j];
        const double sma = s / p;
        double dev = 0.0;
        for (int j = 0; j < p; ++j) dev += fabs(tpd[e - j] - sma);
        const double md = dev / p;
        const double num = tpd[e] - sma;
        ov[i] = (num != 0.0 && md != 0.0) ? (float)(num / (0.015 * md)) : 0.0f;
      }
      store_lane<T>(lo, ov, io.out[0], c);
      return;
    }
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < p - 1 || t >= c.n) {
        ov[i] = qnanf_();
        continue;
      }
      double s = 0.0;
      for (int j = 0; j < p; ++j) s += tp<T>(lh, ll, lc, io, c, t - j);
      const double sma = s / p;
      double dev = 0.0;
      for (int j = 0; j < p; ++j) dev += fabs(tp<T>(lh, ll, lc, io, c, t - j) - sma);
      const double md = dev / p;
      const double num = tp<T>(lh, ll, lc, io, c, t) - sma;
      ov[i] = (num != 0.0 && md != 0.0) ? (float)(num / (0.015 * md)) : 0.0f;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// AROON (reference Aroon.cu:12-50): window of p+1 bars, newest tie wins,
// (100/p)*(p - bars since); MODE 0 = up/down pair, 1 = oscillator.
template <int MODE>
struct FAroonT {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int p;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    wave_sync();
    float od[T::K], ou[T::K];
    const int a = c.s0 + lane_id() * T::K;
    const int p = pr.p;
    const double factor = 100.0 / (double)p;
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < p || t >= c.n) {
        od[i] = ou[i] = qnanf_();
        continue;
      }
      int sh = 0, sl = 0;
      float mx = at<T>(lh, io.in[0], c, t), mn = at<T>(ll, io.in[1], c, t);
      for (int j = 1; j <= p; ++j) {
        const float hv = at<T>(lh, io.in[0], c, t - j), lv = at<T>(ll, io.in[1], c, t - j);
        if (hv > mx) {
          mx = hv;
          sh = j;
        }
        if (lv < mn) {
          mn = lv;
          sl = j;
        }
      }
      if (MODE == 0) {
        ou[i] = (float)(factor * (double)(p - sh));
        od[i] = (float)(factor * (double)(p - sl));
      } else {
        od[i] = (float)(factor * (double)(sl - sh));
      }
    }
    store_lane<T>(lo, od, io.out[0], c);
    if (MODE == 0) store_lane<T>(lo, ou, io.out[1], c);
  }
};
using FAroon = FAroonT<0>;
using FAroonOsc = FAroonT<1>;

// TRANGE (ta_TRANGE.c:36-123): NaN at bar 0.
struct FTrange {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int unused;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P&, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    float* lc = li + 2 * T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    stage<T>(lc, io.in[2], c);
    wave_sync();
    float hv[T::K], lv[T::K], cv[T::K], ov[T::K];
    lane_vals<T>(lh, hv);
    lane_vals<T>(ll, lv);
    lane_vals<T>(lc, cv);
    const int a = c.s0 + lane_id() * T::K;
    double pc = (double)at<T>(lc, io.in[2], c, a - 1);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      ov[i] = (a + i < 1) ? qnanf_() : (float)true_range(hv[i], lv[i], pc);
      pc = cv[i];
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};

// OBV and AD: running sums in double of a per-bar term, rounded to float
// per bar (reference OBV.cu:17-34 and AD.cu:22-71: the terms in double, a
// double inclusive scan, a float cast).  OBV's term at bar 0 is the volume,
// after that sign(close - previous close) x volume (0 when unchanged or
// either close is NaN); AD's is ((close - low) - (high - close)) / (high -
// low) x volume for finite bars with a positive range, 0 otherwise.  Bar 0
// is the scan's seed; every later bar adds its term, so the segment
// transforms are translations (A = 1) chained by the look-back like any
// other recurrence.  For volumes and prices of ordinary range every partial
// OBV sum is exact, so any association gives the sequential sums bit for
// bit; AD's quotients are rounded, so its sums may differ from a sequential
// loop in the last double bits (the reference's scan order differs too).
template <bool AD>
struct FCumT {
  static constexpr bool kScan = true;
  static constexpr int kLevels = 1;
  static constexpr int kMaxR = 1;
  struct P {
    int unused;
  };
  template <class T>
  __device__ __forceinline__ static double term(const float* li, const IO& io, const Seg& c,
                                                int t) {
    if constexpr (AD) {
      const double h = at<T>(li, io.in[0], c, t), l = at<T>(li + T::PLEN, io.in[1], c, t);
      const double cl = at<T>(li + 2 * T::PLEN, io.in[2], c, t);
      const double v = at<T>(li + 3 * T::PLEN, io.in[3], c, t);
      if (!(__builtin_isfinite(h) && __builtin_isfinite(l) && __builtin_isfinite(cl) &&
            __builtin_isfinite(v)))
        return 0.0;
      const double range = h - l;
      return range > 0.0 ? (((cl - l) - (h - cl)) / range) * v : 0.0;
    } else {
      const double v = at<T>(li + T::PLEN, io.in[1], c, t);
      if (t == 0) return v;
      const double p = at<T>(li, io.in[0], c, t), q = at<T>(li, io.in[0], c, t - 1);
      const double sg = p > q ? 1.0 : (p < q ? -1.0 : 0.0);
      return sg * v;
    }
  }
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P&,
                                             const IO& io) {
    constexpr int NIN = AD ? 4 : 2;
#pragma unroll
    for (int j = 0; j < NIN; ++j) stage<T>(li + j * T::PLEN, io.in[j], c);
    wave_sync();
    const int a = c.s0 + lane_id() * T::K;
    double tv[T::K];
    Aff<1> loc;
    aff_id(loc);
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      tv[i] = t < c.n ? term<T>(li, io, c, t) : 0.0;
      if (t > 0 && t < c.n) loc.b[0] = loc.b[0] + tv[i];
    }
    double seed[1] = {c.s0 == 0 ? term<T>(li, io, c, 0) : 0.0}, yin[1];
    level_carry<1>(c, 0, 0, seed, loc, yin);
    double y = yin[0];
    float ov[T::K];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      y = t == 0 ? tv[i] : y + tv[i];
      ov[i] = (float)y;
    }
    store_lane<T>(lo, ov, io.out[0], c);
  }
};
using FObv = FCumT<false>;
using FAd = FCumT<true>;

// STOCH (ta_STOCH.c:36-322): fastK over kP, SMA(dP) twice; both outputs
// start at kP-1 + 2(dP-1).  For dP <= kStochMaxD the fastK values of the
// lane's bars and the 2(dP-1) before them are kept in registers; longer
// smoothing re-evaluates them per output (any dP > 0, as the reference's
// StochasticFused.cu:184-220 accepts).
constexpr int kStochMaxD = 9;
struct FStoch {
  static constexpr bool kScan = false;
  static constexpr int kLevels = 0;
  static constexpr int kMaxR = 0;
  struct P {
    int kp, dp;
  };
  template <class T>
  __device__ __forceinline__ static void run(const Seg& c, float* li, float* lo, const P& pr, const IO& io) {
    float* lh = li;
    float* ll = li + T::PLEN;
    float* lc = li + 2 * T::PLEN;
    stage<T>(lh, io.in[0], c);
    stage<T>(ll, io.in[1], c);
    stage<T>(lc, io.in[2], c);
    wave_sync();
    constexpr int EX = 2 * (kStochMaxD - 1);
    const int a = c.s0 + lane_id() * T::K;
    const int kp = pr.kp, dp = pr.dp;
    const int lb = kp - 1 + 2 * (dp - 1);
    const int ex = 2 * (dp - 1);
    auto fast_k = [&](int t) -> double {
      float hh, lw;
      window_hl<T>(lh, ll, io, c, t, kp, hh, lw);
      const double range = (double)hh - (double)lw;
      const double cl = (double)at<T>(lc, io.in[2], c, t);
      return (fabs(range) <= 1e-14 * (fabs((double)hh) + fabs((double)lw)))
                 ? 0.0
                 : (cl - (double)lw) / range * 100.0;
    };
    if (dp > kStochMaxD) {
      // long smoothing: every output re-evaluates the fast %K windows it
      // averages, each sum in ascending bar order (the oracle's doubles)
      float ok[T::K], od[T::K];
      for (int i = 0; i < T::K; ++i) {
        const int t = a + i;
        ok[i] = od[i] = qnanf_();
        if (t < lb || t >= c.n) continue;
        double sd = 0.0, skt = 0.0;
        for (int q = dp - 1; q >= 0; --q) {
          double sum = 0.0;
          for (int r = dp - 1; r >= 0; --r) sum += fast_k(t - q - r);
          const double sk = sum / dp;
          sd += sk;
          skt = sk;
        }
        ok[i] = (float)skt;
        od[i] = (float)(sd / dp);
      }
      store_lane<T>(lo, ok, io.out[0], c);
      store_lane<T>(lo, od, io.out[1], c);
      return;
    }
    double fk[T::K + EX];
#pragma unroll
    for (int j = 0; j < T::K + EX; ++j) {
      const int t = a - EX + j;
      fk[j] = qnan_();
      if (j >= EX - ex && t >= kp - 1 && t < c.n) fk[j] = fast_k(t);
    }
    double sk[T::K + EX];
#pragma unroll
    for (int j = 0; j < T::K + EX; ++j) {
      sk[j] = qnan_();
      if (j >= EX - (dp - 1)) {
        // ascending window order; indices stay compile-time after unrolling
        double s = 0.0;
#pragma unroll
        for (int q = kStochMaxD - 1; q >= 0; --q)
          if (q < dp && j - q >= 0) s += fk[j - q >= 0 ? j - q : 0];
        sk[j] = s / dp;
      }
    }
    float ok[T::K], od[T::K];
#pragma unroll
    for (int i = 0; i < T::K; ++i) {
      const int t = a + i;
      if (t < lb || t >= c.n) {
        ok[i] = od[i] = qnanf_();
        continue;
      }
      const int j = EX + i;
      double s = 0.0;
#pragma unroll
      for (int q = kStochMaxD - 1; q >= 0; --q)
        if (q < dp) s += sk[j - q];
      ok[i] = (float)sk[j];
      od[i] = (float)(s / dp);
    }
    store_lane<T>(lo, ok, io.out[0], c);
    store_lane<T>(lo, od, io.out[1], c);
  }
};

}  // namespace tca
