// Column tiles for the sliding-window families.
//
// A wave owns a 256-bar segment [s0, s0+256) of one row plus a halo of H
// bars before it, staged in LDS.  Bars are dealt to lanes in columns: LDS
// index e = 64*j + lane is bar s0 - H + e, so each chunk j of 64 bars is one
// lane each.  Every LDS access of a chunk touches 64 consecutive words
// (no bank conflicts) and every global load or store of a chunk is one
// coalesced 256-byte line.  Window outputs are independent, so each lane
// evaluates its bar's window directly, in the oracle's summation order:
// results are bit-identical to the sequential definition for finite and
// non-finite data alike.  Rolling extrema use a sparse table built by
// doubling in LDS (log2(p) passes instead of p compares per bar).
#pragma once
#include "engine.hpp"

namespace tca {

template <int H_>
struct ColTile {
  static constexpr int H = H_;
  static constexpr int SEG = 256;
  static constexpr int E = H + SEG;
  static constexpr int NC = E / kWave;  // chunks
  static constexpr int NH = H / kWave;  // halo chunks
  static_assert(H % kWave == 0, "halo is whole chunks");
};

// Per-wave LDS: staged inputs, one double work array, and an area that
// holds either the two float extrema tables or a second double array.
template <int E>
struct ColLds {
  float h[E], l[E], c[E];
  double d0[E];
  double w[E];
  __device__ float* mx() { return reinterpret_cast<float*>(w); }
  __device__ float* mn() { return reinterpret_cast<float*>(w) + E; }
  __device__ const float* mx() const { return reinterpret_cast<const float*>(w); }
  __device__ const float* mn() const { return reinterpret_cast<const float*>(w) + E; }
  __device__ double* d1() { return w; }
};

__device__ __forceinline__ float lds_f(const float* p, int i) {
  return ((__attribute__((address_space(3))) const float*)p)[i];
}
__device__ __forceinline__ double lds_d(const double* p, int i) {
  return ((__attribute__((address_space(3))) const double*)p)[i];
}
__device__ __forceinline__ void lds_wf(float* p, int i, float v) {
  ((__attribute__((address_space(3))) float*)p)[i] = v;
}
__device__ __forceinline__ void lds_wd(double* p, int i, double v) {
  ((__attribute__((address_space(3))) double*)p)[i] = v;
}

__device__ __forceinline__ void col_sync() {
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  __builtin_amdgcn_wave_barrier();
}

// Stage bars [s0-H, s0+SEG) of one row; bars outside [0, n) read as NaN.
template <class C>
__device__ __forceinline__ void col_stage(float* dst, const float* src, const Seg& c) {
  const int lane = lane_id();
  const float* row = src + c.rowOff;
  float v[C::NC];
#pragma unroll
  for (int j = 0; j < C::NC; ++j) {
    const int t = c.s0 - C::H + j * kWave + lane;
    v[j] = __builtin_nanf("");
    if (t >= 0 && t < c.n) v[j] = __builtin_nontemporal_load(row + t);
  }
#pragma unroll
  for (int j = 0; j < C::NC; ++j) lds_wf(dst, j * kWave + lane, v[j]);
}

// One output per lane per segment chunk; bars past the row end are skipped.
template <class C>
__device__ __forceinline__ void col_store(float* dst, const Seg& c, int j, float v) {
  const int t = c.s0 - C::H + j * kWave + lane_id();
  if (dst && t < c.s1) __builtin_nontemporal_store(v, dst + c.rowOff + t);
}

// Sparse table for window extrema of period p, in place in mx / mn:
// afterwards mx[e] = max over the 2^k bars ending at e (2^k <= p < 2^(k+1)),
// NaN samples replaced by -inf (+inf for the minimum).  The window rule of
// the oracle (start from the newest bar, replace on strict improvement)
// makes the result NaN exactly when the newest bar is NaN, otherwise the
// extremum of the non-NaN samples; combining newer-over-older halves with
// strict comparisons keeps that, ties included.
template <class C>
__device__ __forceinline__ int col_extrema_table(ColLds<C::E>& s, int p) {
  const int lane = lane_id();
#pragma unroll
  for (int j = 0; j < C::NC; ++j) {
    const int e = j * kWave + lane;
    const float a = lds_f(s.h, e), b = lds_f(s.l, e);
    lds_wf(s.mx(), e, __builtin_isnan(a) ? -__builtin_inff() : a);
    lds_wf(s.mn(), e, __builtin_isnan(b) ? __builtin_inff() : b);
  }
  int k = 1;
  for (; 2 * k <= p; k *= 2) {
    col_sync();
    // newest chunk first: chunk j reads older bars that this pass has not
    // rewritten yet
#pragma unroll
    for (int j = C::NC - 1; j >= 0; --j) {
      const int e = j * kWave + lane;
      const int o = e - k >= 0 ? e - k : 0;
      const float a = lds_f(s.mx(), e), ao = lds_f(s.mx(), o);
      const float b = lds_f(s.mn(), e), bo = lds_f(s.mn(), o);
      lds_wf(s.mx(), e, ao > a ? ao : a);
      lds_wf(s.mn(), e, bo < b ? bo : b);
    }
  }
  col_sync();
  return k;
}

// Window extrema of period p ending at e from a table of span k.
template <class C>
__device__ __forceinline__ void col_hl(const ColLds<C::E>& s, int e, int p, int k, float& hh,
                                       float& ll) {
  const int o = e - (p - k) >= 0 ? e - (p - k) : 0;
  const float a = lds_f(s.mx(), e), ao = lds_f(s.mx(), o);
  const float b = lds_f(s.mn(), e), bo = lds_f(s.mn(), o);
  hh = ao > a ? ao : a;
  ll = bo < b ? bo : b;
  if (__builtin_isnan(lds_f(s.h, e))) hh = __builtin_nanf("");
  if (__builtin_isnan(lds_f(s.l, e))) ll = __builtin_nanf("");
}

// Window variance with the oldest sample as shift (the oracle's window_var).
__device__ __forceinline__ double col_var_finish(double s1, double s2, double inv) {
  const double mean = s1 * inv;
  const double e2 = s2 * inv;
  double var = e2 - mean * mean;
  if (var < 1e-12 * e2) var = 0.0;
  if (!(var >= 0.0) && !__builtin_isnan(var)) var = 0.0;
  return var;
}

}  // namespace tca
