// Differentiable indicators (no counterpart in the reference): forward and
// adjoint (reverse-mode) kernels behind torch.autograd.Function.
//
// One workgroup of kAW waves per row.  Thread i owns the contiguous chunk
// [i*C, i*C + C) of the row (C = ceil(n / (64 kAW))).  A recurrence
// y_t = beta*y_{t-1} + u_t is folded per chunk, composed across the lanes of
// a wave by an offset scan with the known multiplier beta^C, across the
// waves through LDS with beta^(64 C), and replayed.  The adjoint of such a
// recurrence is the same recurrence run backwards,
// lambda_t = g_t + beta*lambda_{t+1}: the chunk folds from its end and the
// scans run from the row's last thread down.  Forward passes save only the
// state entering each chunk (64 kAW doubles per row and recurrence);
// adjoint passes recompute each chunk's forward values from it (into LDS
// for EMA and MACD, block by block for RSI) before walking the chunk
// backwards.
//
// Operators and their seeds follow the TA-Lib definitions of the families
// (SMA-mean seed of the first p samples), with the smoothing weight alpha as
// a continuous parameter:
//   EMA_alpha   y_s = mean(x_0..x_{p-1}) at s = p-1, then y_t = (1-a)y + a x_t
//   RSI_alpha   G, L Wilder averages of gains / losses: seed = (sum of the first
//               p) * (1/p) at s = p, then G_t = (1-a)G + a*gain_t; out
//               100*G/(G+L) (50 when G+L is 0)
//   MACD        ta_MACD.c (fixed periods; linear in the prices)
//   BBANDS      SMA middle, population deviation, mid +/- k*sd
// Inputs must be finite (a NaN gives NaN gradients).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "../../include/tacuda_amd.h"
#include "engine.hpp"
#include "host.hpp"

namespace tca {
namespace {

constexpr int kAW = 8;            // waves per row
constexpr int kAT = kAW * kWave;  // threads per row (one workgroup)
constexpr int kMaxChunk = 64;     // C <= 64: rows up to 32768 bars (LDS columns <= 130 KB)

struct Chunk {
  int tid, lane, wave, C, a, e;  // chunk [a, e) of a row of n bars
};
__device__ __forceinline__ Chunk chunk_of(int n) {
  Chunk k;
  k.tid = (int)threadIdx.x;
  k.lane = lane_id();
  k.wave = k.tid / kWave;
  k.C = (n + kAT - 1) / kAT;
  k.a = min(n, k.tid * k.C);
  k.e = min(n, k.a + k.C);
  return k;
}

// Per-thread float array in LDS, padded so neighbouring threads hit
// distinct banks.
struct LdsCol {
  float* base;
  int stride;
  __device__ float& at(int tid, int i) { return base[tid * stride + i]; }
};

// The row's cross-wave exchange: one slot per wave.
struct Xw {
  double v[kAW][2];
};

// Inclusive sum scan over the wave (lane 0 earliest).
template <int R>
__device__ __forceinline__ void wave_iscan_add(double (&v)[R]) {
#define TCA_ADD_STEP(CTRL, RM)                                         \
  _Pragma("unroll") for (int r = 0; r < R; ++r) v[r] += dpp_d<CTRL, RM, 0xf>(0.0, v[r]);
  TCA_ADD_STEP(kRowShr1, 0xf)
  TCA_ADD_STEP(kRowShr2, 0xf)
  TCA_ADD_STEP(kRowShr4, 0xf)
  TCA_ADD_STEP(kRowShr8, 0xf)
  TCA_ADD_STEP(kRowBcast15, 0xa)
  TCA_ADD_STEP(kRowBcast31, 0xc)
#undef TCA_ADD_STEP
}

// Sum over the row's workgroup; every thread receives it.  Every thread of
// the workgroup calls it (it holds barriers).
template <int R>
__device__ __forceinline__ void rsum(double (&v)[R], Xw& x) {
  wave_total<R>(v);
  const int w = (int)threadIdx.x / kWave;
  if (lane_id() == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) x.v[w][r] = v[r];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double s = 0.0;
    for (int k = 0; k < kAW; ++k) s += x.v[k][r];
    v[r] = s;
  }
  __syncthreads();
}

// Forward offset scan: `b` holds each chunk's fold from a zero state;
// returns in `b` the state entering the thread's chunk.  Bars before a
// recurrence's seed contribute exactly 0, so every chunk's multiplier is
// beta^C (powers that underflow are clamped as pow_n does).
template <int R>
__device__ __forceinline__ void carry_fwd(double (&b)[R], const double (&beta)[R], int C,
                                          Xw& x) {
  const int lane = lane_id(), w = (int)threadIdx.x / kWave;
  double A[R];
#pragma unroll
  for (int r = 0; r < R; ++r) A[r] = pow_n(beta[r], C);
  bscan_incl<R, true>(b, A, beta);
  double ex[R], Al[R], Pw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (lane == kWave - 1) x.v[w][r] = b[r];
    ex[r] = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, b[r]);  // lanes before this one
    Al[r] = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, A[r]);  // beta^(C lane)
    Pw[r] = bcast_d(A[r], kWave - 1);               // beta^(64 C)
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double e = 0.0;  // state entering this wave
    for (int k = 0; k < w; ++k) e = fma(Pw[r], e, x.v[k][r]);
    b[r] = fma(Al[r], e, ex[r]);
  }
  __syncthreads();
}

// Reverse offset scan: `b` is each chunk's adjoint folded from its end (the
// adjoint at its first bar given zero beyond it); returns in `b` the
// adjoint just past the thread's chunk.
template <int R>
__device__ __forceinline__ void carry_bwd(double (&b)[R], const double (&beta)[R], int C,
                                          Xw& x) {
  const int lane = lane_id(), w = (int)threadIdx.x / kWave;
  double A[R];
#pragma unroll
  for (int r = 0; r < R; ++r) A[r] = pow_n(beta[r], C);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double ob = __shfl_down(b[r], d);
      const double oA = __shfl_down(A[r], d);
      if (lane + d < kWave) {
        b[r] = fma(A[r], ob, b[r]);
        A[r] = pow_clamp(A[r] * oA, beta[r]);
      }
    }
  }
  // b: adjoint at the chunk's first bar given zero past the wave; A:
  // beta^(C (64 - lane))
  double ex[R], An[R], Pw[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (lane == 0) x.v[w][r] = b[r];
    const double nb = __shfl_down(b[r], 1), nA = __shfl_down(A[r], 1);
    ex[r] = lane + 1 < kWave ? nb : 0.0;
    An[r] = lane + 1 < kWave ? nA : 1.0;
    Pw[r] = bcast_d(A[r], 0);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double e = 0.0;  // adjoint just past this wave
    for (int k = kAW - 1; k > w; --k) e = fma(Pw[r], e, x.v[k][r]);
    b[r] = fma(An[r], e, ex[r]);
  }
  __syncthreads();
}

// Threads walk their chunks in blocks of kB bars: every load of a block is
// issued before the block's arithmetic (one memory round trip per kB bars,
// not per bar), as two float4 accesses when the block lies inside [lo, hi)
// and is 16-byte aligned.  Bars outside [lo, hi) read as 0 and are not
// written.
constexpr int kB = 8;

__device__ __forceinline__ bool whole8(const float* r, int t, int lo, int hi) {
  return t >= lo && t + kB <= hi && ((reinterpret_cast<size_t>(r + t) & 15) == 0);
}

__device__ __forceinline__ void ld8(const float* r, int t, int lo, int hi, float (&v)[kB]) {
  if (whole8(r, t, lo, hi)) {
    const float4 a = *reinterpret_cast<const float4*>(r + t);
    const float4 b = *reinterpret_cast<const float4*>(r + t + 4);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
    v[4] = b.x;
    v[5] = b.y;
    v[6] = b.z;
    v[7] = b.w;
    return;
  }
#pragma unroll
  for (int i = 0; i < kB; ++i) {
    const int u = t + i;
    v[i] = (u >= lo && u < hi) ? r[u] : 0.0f;
  }
}

__device__ __forceinline__ void st8(float* r, int t, int lo, int hi, const float (&v)[kB]) {
  if (whole8(r, t, lo, hi)) {
    *reinterpret_cast<float4*>(r + t) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(r + t + 4) = make_float4(v[4], v[5], v[6], v[7]);
    return;
  }
#pragma unroll
  for (int i = 0; i < kB; ++i) {
    const int u = t + i;
    if (u >= lo && u < hi) r[u] = v[i];
  }
}

// Sum of bars [lo, hi) of a thread's chunk, in bar order.
__device__ __forceinline__ double chunk_sum(const float* r, int lo, int hi) {
  double part = 0.0;
  for (int t0 = lo; t0 < hi; t0 += kB) {
    float v[kB];
    ld8(r, t0, lo, hi, v);
#pragma unroll
    for (int i = 0; i < kB; ++i)
      if (t0 + i < hi) part += (double)v[i];
  }
  return part;
}

// =================================================================== EMA ==
__global__ __launch_bounds__(kAT) void k_ema_fwd(const float* x, float* y, double* ck, int rows,
                                                 int n, int p, double alphaArg,
                                                 const double* alphaPtr) {
  __shared__ Xw xw;
  const double alpha = alphaPtr ? *alphaPtr : alphaArg;
  const int row = blockIdx.x;
  if (row >= rows) return;
  const Chunk k = chunk_of(n);
  const float* xr = x + (size_t)row * n;
  float* yr = y + (size_t)row * n;
  const double beta = 1.0 - alpha;
  const int s = p - 1;
  double sv[1] = {chunk_sum(xr, k.a, min(k.e, s + 1))};
  rsum<1>(sv, xw);
  const double seed = sv[0] / p;
  double b[1] = {0.0};
  for (int t0 = max(k.a, s); t0 < k.e; t0 += kB) {
    float v[kB];
    ld8(xr, t0, k.a, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      if (t >= k.e) break;
      b[0] = (t == s) ? seed : fma(beta, b[0], alpha * (double)v[i]);
    }
  }
  const double be[1] = {beta};
  carry_fwd<1>(b, be, k.C, xw);
  double w = b[0];
  ck[(size_t)row * kAT + k.tid] = w;
  for (int t0 = k.a; t0 < k.e; t0 += kB) {
    float v[kB], o[kB];
    ld8(xr, t0, k.a, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      o[i] = __builtin_nanf("");
      if (t >= s && t < k.e) {
        w = (t == s) ? seed : fma(beta, w, alpha * (double)v[i]);
        o[i] = (float)w;
      }
    }
    st8(yr, t0, k.a, k.e, o);
  }
}

__global__ __launch_bounds__(kAT) void k_ema_bwd(const float* x, const float* g, const double* ck,
                                                 float* dx, double* dal, int rows, int n, int p,
                                                 double alphaArg, const double* alphaPtr) {
  __shared__ Xw xw;
  extern __shared__ float s_col[];
  const double alpha = alphaPtr ? *alphaPtr : alphaArg;
  const int row = blockIdx.x;
  if (row >= rows) return;
  const Chunk k = chunk_of(n);
  LdsCol r{s_col, k.C + 1};
  const float* xr = x + (size_t)row * n;
  const float* gr = g + (size_t)row * n;
  float* dr = dx + (size_t)row * n;
  const double beta = 1.0 - alpha;
  const int s = p - 1;
  // forward recompute: residual x_t - y_{t-1} of every bar after the seed
  double w = ck[(size_t)row * kAT + k.tid];
  double sv[1] = {chunk_sum(xr, k.a, min(k.e, s + 1))};
  rsum<1>(sv, xw);
  const double seed = sv[0] / p;
  for (int t0 = max(k.a, s); t0 < k.e; t0 += kB) {
    float v[kB];
    ld8(xr, t0, k.a, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      if (t >= k.e) break;
      if (t == s) {
        w = seed;
        continue;
      }
      const double xv = (double)v[i];
      r.at(k.tid, t - k.a) = (float)(xv - w);
      w = fma(beta, w, alpha * xv);
    }
  }
  // adjoint: lambda_t = g_t + beta*lambda_{t+1} for t >= s
  const int lo = max(k.a, s);
  double b[1] = {0.0};
  for (int t1 = k.e - 1; t1 >= lo; t1 -= kB) {
    float v[kB];
    ld8(gr, t1 - (kB - 1), lo, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      if (t1 - i < lo) break;
      b[0] = fma(beta, b[0], (double)v[kB - 1 - i]);
    }
  }
  const double be[1] = {beta};
  carry_bwd<1>(b, be, k.C, xw);
  double lam = b[0];
  double red[2] = {0.0, 0.0};  // d alpha, lambda at the seed bar
  for (int t1 = k.e - 1; t1 >= k.a; t1 -= kB) {
    float v[kB], o[kB];
    ld8(gr, t1 - (kB - 1), k.a, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t1 - i;
      o[kB - 1 - i] = 0.0f;
      if (t < s || t < k.a) continue;
      lam = fma(beta, lam, (double)v[kB - 1 - i]);
      if (t == s) {
        red[1] = lam;
      } else {
        o[kB - 1 - i] = (float)(alpha * lam);
        red[0] += lam * (double)r.at(k.tid, t - k.a);
      }
    }
    st8(dr, t1 - (kB - 1), k.a, k.e, o);
  }
  rsum<2>(red, xw);
  // the seed is the mean of bars 0..s
  for (int t = k.a; t < k.e && t <= s; ++t) dr[t] = (float)(red[1] / p);
  if (k.tid == 0) dal[row] = red[0];
}

// =================================================================== RSI ==
// Gains and losses of the bars [lo, hi) of a thread's chunk (bar t uses
// x_{t-1}), in bar order: f(t, gain, loss, d) for every t >= 1.
template <class F>
__device__ __forceinline__ void walk_moves(const float* xr, int lo, int hi, F&& f) {
  float prev = lo >= 1 ? xr[lo - 1] : 0.0f;
  for (int t0 = lo; t0 < hi; t0 += kB) {
    float v[kB];
    ld8(xr, t0, lo, hi, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      if (t >= hi) break;
      if (t >= 1) {
        const double d = (double)v[i] - (double)prev;
        const double gd = d > 0.0 ? d : 0.0;
        f(t, gd, gd - d, d);
      }
      prev = v[i];
    }
  }
}

// The seed averages (sum of the first p gains / losses) * (1/p).
__device__ __forceinline__ void rsi_seeds(const float* xr, const Chunk& k, int s, double inv,
                                          Xw& xw, double& sG, double& sL) {
  double ps[2] = {0.0, 0.0};
  walk_moves(xr, k.a, min(k.e, s + 1), [&](int, double gd, double ld, double) {
    ps[0] += gd;
    ps[1] += ld;
  });
  rsum<2>(ps, xw);
  sG = ps[0] * inv;
  sL = ps[1] * inv;
}

__global__ __launch_bounds__(kAT) void k_rsi_fwd(const float* x, float* y, double* ck, int rows,
                                                 int n, int p, double alphaArg,
                                                 const double* alphaPtr) {
  __shared__ Xw xw;
  const double alpha = alphaPtr ? *alphaPtr : alphaArg;
  const int row = blockIdx.x;
  if (row >= rows) return;
  const Chunk k = chunk_of(n);
  const float* xr = x + (size_t)row * n;
  float* yr = y + (size_t)row * n;
  const double beta = 1.0 - alpha, inv = 1.0 / (double)p;
  const int s = p;
  double sG, sL;
  rsi_seeds(xr, k, s, inv, xw, sG, sL);
  double b[2] = {0.0, 0.0};
  walk_moves(xr, k.a, k.e, [&](int t, double gd, double ld, double) {
    if (t == s) {
      b[0] = sG;
      b[1] = sL;
    } else if (t > s) {
      b[0] = fma(beta, b[0], alpha * gd);
      b[1] = fma(beta, b[1], alpha * ld);
    }
  });
  const double be[2] = {beta, beta};
  carry_fwd<2>(b, be, k.C, xw);
  double G = b[0], L = b[1];
  ck[((size_t)row * kAT + k.tid) * 2] = G;
  ck[((size_t)row * kAT + k.tid) * 2 + 1] = L;
  float prev = k.a >= 1 ? xr[k.a - 1] : 0.0f;
  for (int t0 = k.a; t0 < k.e; t0 += kB) {
    float v[kB], o[kB];
    ld8(xr, t0, k.a, k.e, v);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      o[i] = __builtin_nanf("");
      if (t >= k.e) break;
      if (t >= s) {
        const double d = (double)v[i] - (double)prev;
        const double gd = d > 0.0 ? d : 0.0, ld = gd - d;
        if (t == s) {
          G = sG;
          L = sL;
        } else {
          G = fma(beta, G, alpha * gd);
          L = fma(beta, L, alpha * ld);
        }
        const double sum = G + L;
        o[i] = (float)(sum > 0.0 ? 100.0 * (G / sum) : 50.0);
      }
      prev = v[i];
    }
    st8(yr, t0, k.a, k.e, o);
  }
}

// Wilder step of RSI_alpha at bar t (gd, ld: the bar's gain and loss).
__device__ __forceinline__ void rsi_step(int t, int s, double sG, double sL, double beta,
                                         double alpha, double gd, double ld, double& G,
                                         double& L) {
  if (t == s) {
    G = sG;
    L = sL;
  } else if (t > s) {
    G = fma(beta, G, alpha * gd);
    L = fma(beta, L, alpha * ld);
  }
}

// 1/v: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rcp_d(double v) {
  double y = __builtin_amdgcn_rcp(v);
#pragma unroll
  for (int k = 0; k < 2; ++k) y = fma(y, fma(-v, y, 1.0), y);
  return y;
}

// d(100 G/(G+L)) / dG and / dL at the state (G, L).
__device__ __forceinline__ void rsi_coef(double G, double L, double& cg, double& cl) {
  const double sum = G + L;
  if (sum > 0.0) {
    const double r = rcp_d(sum), q = 100.0 * (r * r);
    cg = q * L;
    cl = -q * G;
  } else {
    cg = cl = 0.0;
  }
}

// The adjoint walks each thread's chunk backwards in kB-bar blocks.  The
// forward states inside a block are recomputed from the state entering the
// block, kept in LDS for every block (2 doubles per thread and block), so a
// row needs 32 KB of LDS at 8192 bars instead of its whole state history.
__global__ __launch_bounds__(kAT) void k_rsi_bwd(const float* x, const float* g, const double* ck,
                                                 float* dx, double* dal, int rows, int n, int p,
                                                 double alphaArg, const double* alphaPtr) {
  __shared__ Xw xw;
  __shared__ float s_first[kAT];
  extern __shared__ double s_ck[];
  const double alpha = alphaPtr ? *alphaPtr : alphaArg;
  const int row = blockIdx.x;
  if (row >= rows) return;
  const Chunk k = chunk_of(n);
  const int nb = (k.C + kB - 1) / kB;
  double* ckG = s_ck;             // [block][thread]
  double* ckL = s_ck + nb * kAT;  // [block][thread]
  const float* xr = x + (size_t)row * n;
  const float* gr = g + (size_t)row * n;
  float* dr = dx + (size_t)row * n;
  const double beta = 1.0 - alpha, inv = 1.0 / (double)p;
  const int s = p;
  double sG, sL;
  rsi_seeds(xr, k, s, inv, xw, sG, sL);
  // forward: the state entering every block of the chunk
  {
    double G = ck[((size_t)row * kAT + k.tid) * 2], L = ck[((size_t)row * kAT + k.tid) * 2 + 1];
    float prev = k.a >= 1 ? xr[k.a - 1] : 0.0f;
    for (int j = 0; j < nb; ++j) {
      const int t0 = k.a + j * kB;
      ckG[j * kAT + k.tid] = G;
      ckL[j * kAT + k.tid] = L;
      if (t0 >= k.e) continue;
      float v[kB];
      ld8(xr, t0, k.a, k.e, v);
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int t = t0 + i;
        if (t >= k.e) break;
        if (t >= 1) {
          const double d = (double)v[i] - (double)prev;
          const double gd = d > 0.0 ? d : 0.0;
          rsi_step(t, s, sG, sL, beta, alpha, gd, gd - d, G, L);
        }
        prev = v[i];
      }
    }
  }
  // Recomputes block j: Ga[i], La[i] = state after bar t0+i; gd/ld of its
  // bars; xv the block's prices.
  auto block = [&](int j, float (&xv)[kB], double (&Ga)[kB], double (&La)[kB], double (&gdv)[kB],
                   double (&dv)[kB]) {
    const int t0 = k.a + j * kB;
    ld8(xr, t0, 0, n, xv);
    float prev = t0 >= 1 ? xr[t0 - 1] : 0.0f;
    double G = ckG[j * kAT + k.tid], L = ckL[j * kAT + k.tid];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      double d = 0.0;
      if (t >= 1 && t < k.e) d = (double)xv[i] - (double)prev;
      const double gd = d > 0.0 ? d : 0.0;
      if (t >= 1 && t < k.e) rsi_step(t, s, sG, sL, beta, alpha, gd, gd - d, G, L);
      Ga[i] = G;
      La[i] = L;
      gdv[i] = gd;
      dv[i] = d;
      prev = xv[i];
    }
  };
  // lambdaG_t = cG_t g_t + beta*lambdaG_{t+1} (t >= s), same for L
  const int lo = max(k.a, s);
  double bq[2] = {0.0, 0.0};
  for (int j = nb - 1; j >= 0; --j) {
    const int t0 = k.a + j * kB;
    if (t0 >= k.e || t0 + kB <= lo) continue;
    float xv[kB], gv[kB];
    double Ga[kB], La[kB], gdv[kB], dv[kB];
    block(j, xv, Ga, La, gdv, dv);
    ld8(gr, t0, lo, k.e, gv);
#pragma unroll
    for (int i = kB - 1; i >= 0; --i) {
      const int t = t0 + i;
      if (t < lo || t >= k.e) continue;
      double cg, cl;
      rsi_coef(Ga[i], La[i], cg, cl);
      bq[0] = fma(beta, bq[0], cg * (double)gv[i]);
      bq[1] = fma(beta, bq[1], cl * (double)gv[i]);
    }
  }
  const double be[2] = {beta, beta};
  carry_bwd<2>(bq, be, k.C, xw);
  double lg = bq[0], lq = bq[1];
  // d(gain)_t and d(loss)_t adjoints, then through d_t = x_t - x_{t-1}.
  // dd_t (the adjoint of d_t) goes to this thread's own part of dx first.
  double da[1] = {0.0}, lS[2] = {0.0, 0.0};
  for (int j = nb - 1; j >= 0; --j) {
    const int t0 = k.a + j * kB;
    if (t0 >= k.e) continue;
    float xv[kB], gv[kB], o[kB];
    double Ga[kB], La[kB], gdv[kB], dv[kB];
    block(j, xv, Ga, La, gdv, dv);
    ld8(gr, t0, lo, k.e, gv);
    const double G0 = ckG[j * kAT + k.tid], L0 = ckL[j * kAT + k.tid];
#pragma unroll
    for (int i = kB - 1; i >= 0; --i) {
      const int t = t0 + i;
      double ddt = 0.0;
      if (t >= lo && t < k.e) {
        double cg, cl;
        rsi_coef(Ga[i], La[i], cg, cl);
        const double gvv = (double)gv[i];
        lg = fma(beta, lg, cg * gvv);
        lq = fma(beta, lq, cl * gvv);
        if (t == s) {
          lS[0] = lg;
          lS[1] = lq;
        } else {
          const double d = dv[i], gd = gdv[i], ld = gd - d;
          const double Gp = i > 0 ? Ga[i - 1] : G0, Lp = i > 0 ? La[i - 1] : L0;
          ddt = (d > 0.0 ? alpha * lg : 0.0) - (d < 0.0 ? alpha * lq : 0.0);
          da[0] += lg * (gd - Gp) + lq * (ld - Lp);
        }
      }
      o[i] = (float)ddt;
    }
    st8(dr, t0, k.a, k.e, o);
  }
  rsum<2>(lS, xw);
  // the seeds are (sum of the first p gains / losses) * (1/p)
  for (int t = max(k.a, 1); t < k.e && t <= s; ++t) {
    const double d = (double)xr[t] - (double)xr[t - 1];
    dr[t] = (float)((d > 0.0 ? lS[0] * inv : 0.0) - (d < 0.0 ? lS[1] * inv : 0.0));
  }
  // dx_t = dd_t - dd_{t+1}; the next chunk's first dd comes from thread+1
  const float first = k.a < k.e ? dr[k.a] : 0.0f;
  s_first[k.tid] = first;
  __syncthreads();
  const float tail = (k.tid + 1 < kAT && k.e < n) ? s_first[k.tid + 1] : 0.0f;
  float cur = first;
  for (int t0 = k.a; t0 < k.e; t0 += kB) {
    float nx[kB], o[kB];
    ld8(dr, t0 + 1, k.a, k.e, nx);  // dd_{t+1} inside the chunk
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t0 + i;
      const float nxt = (t + 1 < k.e) ? nx[i] : tail;
      o[i] = (t == 0 ? 0.0f : cur) - nxt;
      cur = nxt;
    }
    st8(dr, t0, k.a, k.e, o);
  }
  rsum<1>(da, xw);
  if (k.tid == 0) dal[row] = da[0];
}

// ================================================================== MACD ==
// Adjoint of ta_MACD.c (linear in the prices).  gm, gs, gh: gradients of the
// MACD line, signal and histogram outputs.
__global__ __launch_bounds__(kAT) void k_macd_bwd(const float* gm, const float* gs,
                                                  const float* gh, float* dx, int rows, int n,
                                                  int fast, int slow, int sig) {
  __shared__ Xw xw;
  extern __shared__ float s_col[];
  const int row = blockIdx.x;
  if (row >= rows) return;
  const Chunk k = chunk_of(n);
  LdsCol dm{s_col, k.C + 1};
  const size_t ro = (size_t)row * n;
  double fb, fk, sb, sk, gb, gk;
  {
    auto coef = [](int p, double& b, double& kk) {
      double bb = (double)(p - 1) / (double)(p + 1);
      const double q = 1.0 - bb;
      if (bb < 0.5) bb = 1.0 - q;
      b = bb;
      kk = q;
    };
    coef(fast, fb, fk);
    coef(slow, sb, sk);
    coef(sig, gb, gk);
  }
  if (sig == 1) {
    gb = 0.0;
    gk = 1.0;
  }
  const int L1 = slow - 1, L2 = L1 + sig - 1;
  // signal adjoint mu_t = (gs - gh)_t + gb*mu_{t+1} for t >= L2
  const float *gmr = gm + ro, *gsr = gs + ro, *ghr = gh + ro;
  const int lo2 = max(k.a, L2);
  double b[1] = {0.0};
  for (int t1 = k.e - 1; t1 >= lo2; t1 -= kB) {
    float vs[kB], vh[kB];
    ld8(gsr, t1 - (kB - 1), lo2, k.e, vs);
    ld8(ghr, t1 - (kB - 1), lo2, k.e, vh);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      if (t1 - i < lo2) break;
      b[0] = fma(gb, b[0], (double)vs[kB - 1 - i] - (double)vh[kB - 1 - i]);
    }
  }
  {
    const double be[1] = {gb};
    carry_bwd<1>(b, be, k.C, xw);
  }
  double mu = b[0], muS[1] = {0.0};
  for (int t1 = k.e - 1; t1 >= k.a; t1 -= kB) {
    float vs[kB], vh[kB], vm[kB];
    ld8(gsr, t1 - (kB - 1), lo2, k.e, vs);
    ld8(ghr, t1 - (kB - 1), lo2, k.e, vh);
    ld8(gmr, t1 - (kB - 1), lo2, k.e, vm);
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t1 - i;
      if (t < k.a) break;
      const int j = kB - 1 - i;
      double d = 0.0;
      if (t >= L2) {
        mu = fma(gb, mu, (double)vs[j] - (double)vh[j]);
        d = (double)vm[j] + (double)vh[j];
        if (t == L2) muS[0] = mu;
        else d += gk * mu;
      }
      dm.at(k.tid, t - k.a) = (float)d;
    }
  }
  rsum<1>(muS, xw);
  for (int t = max(k.a, L1); t < k.e && t <= L2; ++t)
    dm.at(k.tid, t - k.a) += (float)(muS[0] / sig);
  // fast / slow adjoints from dm (macd = fast - slow) for t >= L1
  double bfs[2] = {0.0, 0.0};
  for (int t = k.e - 1; t >= k.a && t >= L1; --t) {
    const double d = dm.at(k.tid, t - k.a);
    bfs[0] = fma(fb, bfs[0], d);
    bfs[1] = fma(sb, bfs[1], -d);
  }
  {
    const double be[2] = {fb, sb};
    carry_bwd<2>(bfs, be, k.C, xw);
  }
  double nf = bfs[0], ns = bfs[1], nS[2] = {0.0, 0.0};
  for (int t1 = k.e - 1; t1 >= k.a; t1 -= kB) {
    float o[kB];
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int t = t1 - i;
      double out = 0.0;
      if (t >= L1 && t >= k.a) {
        const double d = dm.at(k.tid, t - k.a);
        nf = fma(fb, nf, d);
        ns = fma(sb, ns, -d);
        if (t == L1) {
          nS[0] = nf;
          nS[1] = ns;
        } else {
          out = fk * nf + sk * ns;
        }
      }
      o[kB - 1 - i] = (float)out;
    }
    st8(dx + ro, t1 - (kB - 1), k.a, k.e, o);
  }
  rsum<2>(nS, xw);
  for (int t = k.a; t < k.e && t <= L1; ++t) {
    double add = nS[1] / slow;
    if (t > L1 - fast) add += nS[0] / fast;
    dx[ro + t] = (float)((double)dx[ro + t] + add);
  }
}

// ================================================================ BBANDS ==
// Window adjoint: out_t = mid_t + ku*sd_t (upper), mid_t, mid_t - kd*sd_t,
// mid = window mean, sd = population deviation.  dL/dx_j sums, over the
// windows t = j..j+p-1 holding bar j, A_t + B_t (x_j - mid_t) with
// A_t = (gu+gm+gl)_t/p and B_t = (ku*gu - kd*gl)_t/(p*sd_t).
//
// One workgroup of kBT threads per kBB-bar tile.  The tile's windows come
// from prefix sums of (x - c) and (x - c)^2 over its bars (c: the tile's
// first bar, so the sums stay small next to the window moments), and every
// bar's sum over its p windows from prefix sums of A, B and B*(mid - c):
// O(1) a bar after two tile scans.  Threads own contiguous runs of bars and
// of windows for the scans; 1/sd is a refined hardware reciprocal square
// root.
constexpr int kBT = 256;           // threads per tile
constexpr int kBB = 512;           // bars per tile
constexpr int kBBScanMaxP = 1024;  // longer periods take the direct kernel
constexpr int kBBRun = 8;          // windows per thread at most: (kBB + kBBScanMaxP - 1) / kBT

// Exclusive sum scan over the workgroup's kBT threads (every thread calls
// it; it holds barriers).
template <int R>
__device__ __forceinline__ void blk_escan(double (&v)[R], double (*sh)[3]) {
  const int lane = lane_id(), w = (int)threadIdx.x / kWave;
  wave_iscan_add<R>(v);
  double ex[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (lane == kWave - 1) sh[w][r] = v[r];
    ex[r] = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, v[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double e = 0.0;
    for (int k = 0; k < w; ++k) e += sh[k][r];
    v[r] = ex[r] + e;
  }
  __syncthreads();
}

// The tile's LDS: the bar prefix sums (2 x (L + 1) doubles), later the
// window prefix sums in the same place (3 x (W + 1)), then the staged bars.
__host__ __device__ size_t bb_scan_doubles(int p) {
  const size_t L = kBB + 2 * (size_t)p - 2, W = kBB + (size_t)p - 1;
  return 2 * (L + 1) > 3 * (W + 1) ? 2 * (L + 1) : 3 * (W + 1);
}
size_t bb_scan_lds(int p) {
  return bb_scan_doubles(p) * sizeof(double) + (kBB + 2 * (size_t)p - 2) * sizeof(float);
}

// 1/sqrt(v) for v > 0: the hardware estimate and two Newton steps.
__device__ __forceinline__ double rsqrt_d(double v) {
  double y = __builtin_amdgcn_rsq(v);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const double e = fma(-v * y, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

__global__ __launch_bounds__(kBT) void k_bbands_bwd_scan(const float* x, const float* gu,
                                                         const float* gm, const float* gl,
                                                         float* dx, int rows, int n, int p,
                                                         double ku, double kd) {
  extern __shared__ double s_bb[];
  __shared__ double sh[kBT / kWave][3];
  const int tiles = (n + kBB - 1) / kBB;
  const int row = blockIdx.x / tiles, tile = blockIdx.x % tiles;
  if (row >= rows) return;
  const int tid = (int)threadIdx.x;
  const int Lmax = kBB + 2 * p - 2, Wmax = kBB + p - 1;
  const int j0 = tile * kBB;
  const int T0 = j0, T1 = min(n, j0 + kBB + p - 1);  // windows used by this tile
  const int X0 = max(0, T0 - p + 1);                  // bars those windows read
  const int L = T1 - X0, W = T1 - T0;
  double* P1 = s_bb;             // [L + 1] prefix sums of x - c
  double* P2 = P1 + (Lmax + 1);  // [L + 1] prefix sums of (x - c)^2
  double* QA = s_bb;             // [W + 1] prefix sums of A (after P1, P2 are read)
  double* QB = QA + (Wmax + 1);  // [W + 1] prefix sums of B
  double* QM = QB + (Wmax + 1);  // [W + 1] prefix sums of B (mid - c)
  float* xs = reinterpret_cast<float*>(s_bb + bb_scan_doubles(p));  // [L]
  const size_t ro = (size_t)row * n;
  for (int i = tid; i < L; i += kBT) xs[i] = x[ro + X0 + i];
  __syncthreads();
  const double c = (double)xs[T0 - X0];
  // prefix sums over the staged bars, thread runs of m bars
  {
    const int m = (L + kBT - 1) / kBT;
    const int r0 = min(L, tid * m), r1 = min(L, r0 + m);
    double v[2] = {0.0, 0.0};
    for (int i = r0; i < r1; ++i) {
      const double d = (double)xs[i] - c;
      v[0] += d;
      v[1] += d * d;
    }
    blk_escan<2>(v, sh);
    if (tid == 0) P1[0] = P2[0] = 0.0;
    for (int i = r0; i < r1; ++i) {
      const double d = (double)xs[i] - c;
      v[0] += d;
      v[1] += d * d;
      P1[i + 1] = v[0];
      P2[i + 1] = v[1];
    }
  }
  __syncthreads();
  // every window's A, B, B (mid - c) for a run of windows per thread, in
  // registers; then their prefix sums over the tile
  {
    const double inv = 1.0 / (double)p;
    const int m = (W + kBT - 1) / kBT;
    const int r0 = min(W, tid * m), cnt = min(W, r0 + m) - r0;
    double A[kBBRun], B[kBBRun], BM[kBBRun];
    double v[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kBBRun; ++k) {
      A[k] = B[k] = BM[k] = 0.0;
      const int t = T0 + r0 + k;
      if (k < cnt && t >= p - 1) {
        const int lo = t - p + 1 - X0, hi = t + 1 - X0;
        const double mean = (P1[hi] - P1[lo]) * inv, e2 = (P2[hi] - P2[lo]) * inv;
        const double var = e2 - mean * mean;
        const double u = gu[ro + t], mm = gm[ro + t], l = gl[ro + t];
        A[k] = (u + mm + l) * inv;
        if (var > 0.0 && !(var < 1e-12 * e2)) B[k] = (ku * u - kd * l) * inv * rsqrt_d(var);
        BM[k] = B[k] * mean;
      }
      v[0] += A[k];
      v[1] += B[k];
      v[2] += BM[k];
    }
    blk_escan<3>(v, sh);  // its barriers also end every read of P1 and P2
    if (tid == 0) QA[0] = QB[0] = QM[0] = 0.0;
#pragma unroll
    for (int k = 0; k < kBBRun; ++k) {
      if (k >= cnt) break;
      v[0] += A[k];
      v[1] += B[k];
      v[2] += BM[k];
      QA[r0 + k + 1] = v[0];
      QB[r0 + k + 1] = v[1];
      QM[r0 + k + 1] = v[2];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kBB / kBT; ++q) {
    const int j = j0 + q * kBT + tid;
    if (j >= n) break;
    const int tlo = max(j, p - 1), thi = min(n - 1, j + p - 1);
    double acc = 0.0;
    if (tlo <= thi) {
      const int i0 = tlo - T0, i1 = thi - T0 + 1;
      const double sa = QA[i1] - QA[i0], sb = QB[i1] - QB[i0], sm = QM[i1] - QM[i0];
      acc = sa + ((double)xs[j - X0] - c) * sb - sm;
    }
    dx[ro + j] = (float)acc;
  }
}

// Periods above kBBScanMaxP: one thread per bar of a 256-bar tile, every
// window's moments and every bar's window sum evaluated directly; the
// tile's per-window coefficients are built once in LDS.
__global__ __launch_bounds__(256) void k_bbands_bwd(const float* x, const float* gu,
                                                    const float* gm, const float* gl, float* dx,
                                                    int rows, int n, int p, double ku,
                                                    double kd) {
  extern __shared__ double s_bb[];
  const int tiles = (n + 255) / 256;
  const int row = blockIdx.x / tiles, tile = blockIdx.x % tiles;
  if (row >= rows) return;
  const int j0 = tile * 256;
  const int T0 = j0, T1 = min(n, j0 + 256 + p - 1);  // windows used by this tile
  const int X0 = max(0, T0 - p + 1);                  // bars those windows read
  float* xs = reinterpret_cast<float*>(s_bb + 3 * (256 + p));
  const size_t ro = (size_t)row * n;
  for (int i = threadIdx.x; X0 + i < T1; i += 256) xs[i] = x[ro + X0 + i];
  __syncthreads();
  double* A = s_bb;            // (gu+gm+gl)/p
  double* B = s_bb + 256 + p;  // (ku*gu - kd*gl)/(p*sd)
  double* M = s_bb + 2 * (256 + p);
  for (int i = threadIdx.x; T0 + i < T1; i += 256) {
    const int t = T0 + i;
    double a = 0.0, bcoef = 0.0, mid = 0.0;
    if (t >= p - 1) {
      double sum = 0.0, s1 = 0.0, s2 = 0.0;
      const double shift = xs[t - p + 1 - X0];
      for (int q = t - p + 1; q <= t; ++q) {
        const double v = xs[q - X0];
        sum += v;
        const double dd = v - shift;
        s1 += dd;
        s2 += dd * dd;
      }
      mid = sum / p;
      const double inv = 1.0 / (double)p, mean = s1 * inv, e2 = s2 * inv;
      double var = e2 - mean * mean;
      if (var < 1e-12 * e2) var = 0.0;
      const double sd = var > 0.0 ? sqrt(var) : 0.0;
      const double u = gu[ro + t], m = gm[ro + t], l = gl[ro + t];
      a = (u + m + l) / p;
      bcoef = sd > 0.0 ? (ku * u - kd * l) / (p * sd) : 0.0;
    }
    A[i] = a;
    B[i] = bcoef;
    M[i] = mid;
  }
  __syncthreads();
  const int j = j0 + threadIdx.x;
  if (j >= n) return;
  const double xj = xs[j - X0];
  double acc = 0.0;
  for (int t = max(j, p - 1); t <= min(n - 1, j + p - 1); ++t) {
    const int i = t - T0;
    acc += A[i] + B[i] * (xj - M[i]);
  }
  dx[ro + j] = (float)acc;
}

ctStatus_t check(const void* a, const void* b, int rows, int n, int p, const char* fn) {
  if (!a || !b) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  if (rows <= 0 || n <= 0 || p <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: invalid shape or period", fn);
  if ((n + kAT - 1) / kAT > kMaxChunk)
    return fail(CT_STATUS_UNSUPPORTED, "%s: rows longer than %d bars are not provided", fn,
                kAT * kMaxChunk);
  return CT_STATUS_SUCCESS;
}

size_t col_bytes(int n, int arrays) {
  const int C = (n + kAT - 1) / kAT;
  return (size_t)arrays * kAT * (C + 1) * sizeof(float);
}

template <class K>
ctStatus_t allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    TCA_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return CT_STATUS_SUCCESS;
}

}  // namespace
}  // namespace tca

using namespace tca;

namespace {

ctStatus_t ema_alpha_fwd(const float* x, float* y, double* ckpt, int n, int rows, int period,
                         double alpha, const double* alphaPtr, hipStream_t s) {
  ctStatus_t r = check(x, y, rows, n, period, "ema_alpha");
  if (r != CT_STATUS_SUCCESS) return r;
  if (!ckpt) return fail(CT_STATUS_INVALID_ARGUMENT, "ema_alpha: null checkpoint");
  hipLaunchKernelGGL(k_ema_fwd, dim3(rows), dim3(kAT), 0, s, x, y, ckpt, rows, n, period, alpha,
                     alphaPtr);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t ema_alpha_bwd(const float* x, const float* gy, const double* ckpt, float* gx,
                         double* galpha, int n, int rows, int period, double alpha,
                         const double* alphaPtr, hipStream_t s) {
  ctStatus_t r = check(x, gy, rows, n, period, "ema_alpha backward");
  if (r != CT_STATUS_SUCCESS) return r;
  if (!ckpt || !gx || !galpha)
    return fail(CT_STATUS_INVALID_ARGUMENT, "ema_alpha backward: null pointer");
  const size_t lds = col_bytes(n, 1);
  r = allow_lds(k_ema_bwd, lds);
  if (r != CT_STATUS_SUCCESS) return r;
  hipLaunchKernelGGL(k_ema_bwd, dim3(rows), dim3(kAT), lds, s, x, gy, ckpt, gx, galpha, rows, n,
                     period, alpha, alphaPtr);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t rsi_alpha_fwd(const float* x, float* y, double* ckpt, int n, int rows, int period,
                         double alpha, const double* alphaPtr, hipStream_t s) {
  ctStatus_t r = check(x, y, rows, n, period, "rsi_alpha");
  if (r != CT_STATUS_SUCCESS) return r;
  if (!ckpt) return fail(CT_STATUS_INVALID_ARGUMENT, "rsi_alpha: null checkpoint");
  hipLaunchKernelGGL(k_rsi_fwd, dim3(rows), dim3(kAT), 0, s, x, y, ckpt, rows, n, period, alpha,
                     alphaPtr);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t rsi_alpha_bwd(const float* x, const float* gy, const double* ckpt, float* gx,
                         double* galpha, int n, int rows, int period, double alpha,
                         const double* alphaPtr, hipStream_t s) {
  ctStatus_t r = check(x, gy, rows, n, period, "rsi_alpha backward");
  if (r != CT_STATUS_SUCCESS) return r;
  if (!ckpt || !gx || !galpha)
    return fail(CT_STATUS_INVALID_ARGUMENT, "rsi_alpha backward: null pointer");
  // per thread and kB-bar block: the entering G and L (doubles)
  const size_t lds = (size_t)2 * kAT * (((n + kAT - 1) / kAT + kB - 1) / kB) * sizeof(double);
  r = allow_lds(k_rsi_bwd, lds);
  if (r != CT_STATUS_SUCCESS) return r;
  hipLaunchKernelGGL(k_rsi_bwd, dim3(rows), dim3(kAT), lds, s, x, gy, ckpt, gx, galpha, rows, n,
                     period, alpha, alphaPtr);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

ctStatus_t alpha_ptr_ok(const double* d_alpha, const char* what) {
  if (!d_alpha) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null device alpha", what);
  return CT_STATUS_SUCCESS;
}

}  // namespace

extern "C" {

CTAPI_EXPORT int ctx_alpha_ckpt_per_row(void) { return kAT; }

CTAPI_EXPORT ctStatus_t ctx_ema_alpha_fwd_d(const float* x, float* y, double* ckpt, int n,
                                            int rows, int period, double alpha, hipStream_t s) {
  return guarded([&] { return ema_alpha_fwd(x, y, ckpt, n, rows, period, alpha, nullptr, s); });
}

CTAPI_EXPORT ctStatus_t ctx_ema_alpha_bwd_d(const float* x, const float* gy, const double* ckpt,
                                            float* gx, double* galpha, int n, int rows,
                                            int period, double alpha, hipStream_t s) {
  return guarded([&] {
    return ema_alpha_bwd(x, gy, ckpt, gx, galpha, n, rows, period, alpha, nullptr, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_fwd_d(const float* x, float* y, double* ckpt, int n,
                                            int rows, int period, double alpha, hipStream_t s) {
  return guarded([&] { return rsi_alpha_fwd(x, y, ckpt, n, rows, period, alpha, nullptr, s); });
}

CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_bwd_d(const float* x, const float* gy, const double* ckpt,
                                            float* gx, double* galpha, int n, int rows,
                                            int period, double alpha, hipStream_t s) {
  return guarded([&] {
    return rsi_alpha_bwd(x, gy, ckpt, gx, galpha, n, rows, period, alpha, nullptr, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_ema_alpha_fwd_dp(const float* x, float* y, double* ckpt, int n,
                                             int rows, int period, const double* d_alpha,
                                             hipStream_t s) {
  return guarded([&] {
    ctStatus_t r = alpha_ptr_ok(d_alpha, "ema_alpha");
    return r != CT_STATUS_SUCCESS ? r
                                  : ema_alpha_fwd(x, y, ckpt, n, rows, period, 0.0, d_alpha, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_ema_alpha_bwd_dp(const float* x, const float* gy, const double* ckpt,
                                             float* gx, double* galpha, int n, int rows,
                                             int period, const double* d_alpha, hipStream_t s) {
  return guarded([&] {
    ctStatus_t r = alpha_ptr_ok(d_alpha, "ema_alpha backward");
    return r != CT_STATUS_SUCCESS
               ? r
               : ema_alpha_bwd(x, gy, ckpt, gx, galpha, n, rows, period, 0.0, d_alpha, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_fwd_dp(const float* x, float* y, double* ckpt, int n,
                                             int rows, int period, const double* d_alpha,
                                             hipStream_t s) {
  return guarded([&] {
    ctStatus_t r = alpha_ptr_ok(d_alpha, "rsi_alpha");
    return r != CT_STATUS_SUCCESS ? r
                                  : rsi_alpha_fwd(x, y, ckpt, n, rows, period, 0.0, d_alpha, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_bwd_dp(const float* x, const float* gy, const double* ckpt,
                                             float* gx, double* galpha, int n, int rows,
                                             int period, const double* d_alpha, hipStream_t s) {
  return guarded([&] {
    ctStatus_t r = alpha_ptr_ok(d_alpha, "rsi_alpha backward");
    return r != CT_STATUS_SUCCESS
               ? r
               : rsi_alpha_bwd(x, gy, ckpt, gx, galpha, n, rows, period, 0.0, d_alpha, s);
  });
}

CTAPI_EXPORT ctStatus_t ctx_macd_bwd_d(const float* gMacd, const float* gSignal,
                                       const float* gHist, float* gx, int n, int rows, int fast,
                                       int slow, int signal, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    ctStatus_t r = check(gMacd, gx, rows, n, 1, "macd backward");
    if (r != CT_STATUS_SUCCESS) return r;
    if (!gSignal || !gHist) return fail(CT_STATUS_INVALID_ARGUMENT, "macd backward: null pointer");
    if (fast <= 0 || slow <= 0 || signal <= 0)
      return fail(CT_STATUS_INVALID_ARGUMENT, "macd backward: invalid periods");
    if (slow < fast) std::swap(slow, fast);
    const size_t lds = col_bytes(n, 1);
    r = allow_lds(k_macd_bwd, lds);
    if (r != CT_STATUS_SUCCESS) return r;
    hipLaunchKernelGGL(k_macd_bwd, dim3(rows), dim3(kAT), lds, s, gMacd, gSignal, gHist, gx,
                       rows, n, fast, slow, signal);
    TCA_HIP(hipGetLastError());
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ctx_bbands_bwd_d(const float* x, const float* gUpper,
                                         const float* gMiddle, const float* gLower, float* gx,
                                         int n, int rows, int period, float nbDevUp,
                                         float nbDevDn, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!x || !gUpper || !gMiddle || !gLower || !gx)
      return fail(CT_STATUS_INVALID_ARGUMENT, "bbands backward: null pointer");
    if (rows <= 0 || n <= 0 || period <= 0)
      return fail(CT_STATUS_INVALID_ARGUMENT, "bbands backward: invalid shape or period");
    if (period > 4096) return fail(CT_STATUS_UNSUPPORTED, "bbands backward: period above 4096");
    if (period <= kBBScanMaxP) {
      const size_t lds = bb_scan_lds(period);
      ctStatus_t r = allow_lds(k_bbands_bwd_scan, lds);
      if (r != CT_STATUS_SUCCESS) return r;
      const int tiles = (n + kBB - 1) / kBB;
      hipLaunchKernelGGL(k_bbands_bwd_scan, dim3((unsigned)rows * tiles), dim3(kBT), lds, s, x,
                         gUpper, gMiddle, gLower, gx, rows, n, period, (double)nbDevUp,
                         (double)nbDevDn);
      TCA_HIP(hipGetLastError());
      return CT_STATUS_SUCCESS;
    }
    const size_t lds = (size_t)3 * (256 + period) * sizeof(double) +
                       (size_t)(256 + 2 * period) * sizeof(float);
    ctStatus_t r = allow_lds(k_bbands_bwd, lds);
    if (r != CT_STATUS_SUCCESS) return r;
    const int tiles = (n + 255) / 256;
    hipLaunchKernelGGL(k_bbands_bwd, dim3((unsigned)rows * tiles), dim3(256), lds, s, x, gUpper,
                       gMiddle, gLower, gx, rows, n, period, (double)nbDevUp, (double)nbDevDn);
    TCA_HIP(hipGetLastError());
    return CT_STATUS_SUCCESS;
  });
}

}  // extern "C"
