// Device-side engine shared by every indicator kernel.
//
// Work unit: one wave64 owns one SEGMENT of one row -- K contiguous bars per
// lane, SEG = 64*K bars.  Rows are the series of a dense [rows x len]
// matrix, of a ragged batch, or the single series of a ct_*_d call.
//
//  * Inputs are staged into LDS with a halo of H bars before the segment so
//    window families read neighbours from LDS; older bars come from global
//    memory (L2).
//  * Recurrences are affine maps y -> A*y + b.  Each lane composes its bars,
//    a DPP wave scan (row_shr 1/2/4/8, row_bcast 15/31, wave_shr 1) gives the
//    exclusive carry.  Rows longer than one segment combine segments through
//    a two-level tree in global memory: every segment publishes its
//    transform, the last wave of each group of 64 consecutive segments
//    publishes the group's composite, and a segment composes the group
//    composites and in-group transforms before it -- a fixed number of
//    memory round trips however long the row.  Slot flags carry a
//    device-side epoch, so the workspace needs no clearing between launches
//    or graph replays.
//  * Arithmetic of every recurrence is FP64; lanes replay the TA-Lib update
//    order from their carry so the outputs match the sequential reference to
//    the last float32 bit except where a carry differs in the last FP64 ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tca {

constexpr int kWave = 64;

// --------------------------------------------------------------- workspace --
// Launch bookkeeping at the head of a workspace.  Completion is counted on
// 64 striped counters (one cache line each) so blocks do not serialize on a
// single address; the block that completes the last stripe bumps the epoch.
constexpr int kStripes = 64;
struct WsHeader {
  unsigned int epoch;         // launch generation, bumped when a launch ends
  unsigned int stripesDone;   // stripes completed by the running launch
  unsigned int pad0[14];
  unsigned int stripe[kStripes][16];  // blocks completed per stripe
};

constexpr int kGroup = 64;  // segments per group of the look-back tree

struct Geometry {
  // Row r of a dense matrix starts at r*ld; ragged rows use offsets[].
  const long long* offsets;  // ragged: rows+1 row offsets (elements), or null
  const int* segStart;       // ragged: rows+1 first-segment indices, or null
  long long ld;              // dense row stride (elements)
  int len;                   // dense row length
  int rows;
  int segsPerRow;            // dense only
  int totalSegs;
};

// Look-back area: per level, one slot per segment then one per group.  A
// slot is `stride` doubles: [0] = flag (epoch), [1..2R] = (A, b) pairs.
struct Ws {
  WsHeader* hdr;
  double* area;
  int stride;
};

// ------------------------------------------------------------ lane helpers --
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double bcast_d(double v, int src) {
  unsigned long long u = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), src);
  int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <int CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_d(double old, double src) {
  unsigned long long o = __double_as_longlong(old), s = __double_as_longlong(src);
  int lo = __builtin_amdgcn_update_dpp((int)(o & 0xffffffffu), (int)(s & 0xffffffffu), CTRL,
                                       ROW_MASK, BANK_MASK, false);
  int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, ROW_MASK,
                                       BANK_MASK, false);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// DPP controls (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

// --------------------------------------------------------- affine monoid --
// (A, b) represents y -> A*y + b.  combine(e, l) applies e first, then l.
template <int R>
struct Aff {
  double A[R];
  double b[R];
};

template <int R>
__device__ __forceinline__ void aff_id(Aff<R>& v) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v.A[r] = 1.0;
    v.b[r] = 0.0;
  }
}

// Multipliers are products of decay factors in [0, 1], combined with plain
// fma so a zero factor keeps TA-Lib's literal arithmetic (0 * NaN is NaN).
// A product of nonzero factors that underflows is clamped to the smallest
// normal double, so a carried Inf or NaN survives it as it does in the
// sequential recurrence.  A structural reset -- a segment that (re)starts the
// state at its seed bar, or holds no valid bar yet -- is marked by A = -0.0
// and drops whatever came before it, which keeps rows independent.
__device__ __forceinline__ bool is_reset(double a) { return a == 0.0 && signbit(a); }
__device__ __forceinline__ double decay_mul(double a, double m) {
  const double p = a * m;
  return (p == 0.0 && a != 0.0 && m != 0.0) ? 2.2250738585072014e-308 : p;
}
__device__ __forceinline__ double fma_carry(double A, double y, double b) {
  return is_reset(A) ? b : fma(A, y, b);
}

template <int R>
__device__ __forceinline__ void aff_after(const Aff<R>& e, Aff<R>& l) {
  // l := l o e
#pragma unroll
  for (int r = 0; r < R; ++r) {
    l.b[r] = fma_carry(l.A[r], e.b[r], l.b[r]);
    l.A[r] = (is_reset(l.A[r]) || is_reset(e.A[r])) ? -0.0 : decay_mul(l.A[r], e.A[r]);
  }
}

template <int CTRL, int RM, int BM, int R>
__device__ __forceinline__ void scan_step(Aff<R>& v) {
  Aff<R> e;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    e.A[r] = dpp_d<CTRL, RM, BM>(1.0, v.A[r]);
    e.b[r] = dpp_d<CTRL, RM, BM>(0.0, v.b[r]);
  }
  aff_after(e, v);
}

// Inclusive wave scan (lane 0 first in time).
template <int R>
__device__ __forceinline__ void wave_scan_incl(Aff<R>& v) {
  scan_step<kRowShr1, 0xf, 0xf>(v);
  scan_step<kRowShr2, 0xf, 0xf>(v);
  scan_step<kRowShr4, 0xf, 0xf>(v);
  scan_step<kRowShr8, 0xf, 0xf>(v);
  scan_step<kRowBcast15, 0xa, 0xf>(v);
  scan_step<kRowBcast31, 0xc, 0xf>(v);
}

// Exclusive scan; `tot` receives the wave total (inclusive of lane 63).
template <int R>
__device__ __forceinline__ void wave_scan_excl(Aff<R>& v, Aff<R>& tot) {
  wave_scan_incl(v);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    tot.A[r] = bcast_d(v.A[r], 63);
    tot.b[r] = bcast_d(v.b[r], 63);
    v.A[r] = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, v.A[r]);
    v.b[r] = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, v.b[r]);
  }
}

template <int R>
__device__ __forceinline__ double aff_apply(const Aff<R>& t, int r, double y) {
  return fma_carry(t.A[r], y, t.b[r]);
}

__device__ __forceinline__ double qnan_d() { return __longlong_as_double(0x7ff8000000000000ll); }

// ----------------------------------------------------------- look-back --
// Slots are read and written only with relaxed agent-scope atomics, which
// are coherent across the XCDs' L2s by themselves.  Ordering payload before
// flag (and flag before payload on the reader) needs only the wave's own
// memory counters drained -- no L2 writeback or invalidate, which an
// agent-scope release/acquire would add to every publish and every poll.
__device__ __forceinline__ void vmem_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_s_waitcnt(0);
}
__device__ __forceinline__ unsigned long long ld_flag(const double* p) {
  return __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_rel_d(const double* p) {
  unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(u);
}
__device__ __forceinline__ void st_rel_d(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                     (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Segment context of the running wave.
struct Seg {
  long long rowOff;  // element offset of the row start
  int n;             // row length
  int s0, s1;        // segment bar range within the row
  int seg, nseg;     // segment index in the row, segments in the row
  int gseg;          // global segment index (ticket)
  unsigned epoch;
  Ws ws;
  int totalSegs;
  bool tree;         // this launch combines segments through the tree
};

__device__ __forceinline__ int total_groups(int totalSegs) { return totalSegs / kGroup; }

__device__ __forceinline__ double* slot_ptr(const Seg& c, int level, int idx, bool group) {
  const size_t per = (size_t)c.totalSegs + total_groups(c.totalSegs);
  return c.ws.area + ((size_t)level * per + (group ? c.totalSegs + idx : idx)) * c.ws.stride;
}

template <int R>
__device__ __forceinline__ void write_slot(const Seg& c, double* s, const Aff<R>& v) {
  if (lane_id() == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      st_rel_d(s + 1 + 2 * r, v.A[r]);
      st_rel_d(s + 2 + 2 * r, v.b[r]);
    }
    vmem_order();
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(s), (unsigned long long)c.epoch,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Reduce across the wave, lane 0 earliest in time; result broadcast.
template <int R>
__device__ __forceinline__ void wave_compose(Aff<R>& v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    Aff<R> o;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      o.A[r] = __shfl_down(v.A[r], off);
      o.b[r] = __shfl_down(v.b[r], off);
    }
    // o holds later lanes: v := o after v
    if (lane + off < kWave) {
      aff_after(v, o);
      v = o;
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    v.A[r] = bcast_d(v.A[r], 0);
    v.b[r] = bcast_d(v.b[r], 0);
  }
}

// Lane i loads slot idx0 + i when i < cnt (identity otherwise), spinning
// until it carries this launch's epoch, and the wave composes them.
template <int R>
__device__ __forceinline__ Aff<R> gather(const Seg& c, int level, int idx0, int cnt, bool group) {
  const int lane = lane_id();
  Aff<R> v;
  aff_id(v);
  const bool mine = lane < cnt;
  const double* s = mine ? slot_ptr(c, level, idx0 + lane, group) : nullptr;
  if (mine) {
    while ((unsigned)ld_flag(s) != c.epoch) __builtin_amdgcn_s_sleep(1);
  }
  vmem_order();
  if (mine) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      v.A[r] = ld_rel_d(s + 1 + 2 * r);
      v.b[r] = ld_rel_d(s + 2 + 2 * r);
    }
  }
  wave_compose(v);
  return v;
}

// Publish this segment's transform; the last wave of a full group then
// publishes the group's composite.
template <int R>
__device__ __forceinline__ void publish_tree(const Seg& c, int level, const Aff<R>& t) {
  write_slot(c, slot_ptr(c, level, c.gseg, false), t);
  const int g = c.gseg / kGroup;
  if ((c.gseg % kGroup) == kGroup - 1) {
    Aff<R> gv = gather<R>(c, level, g * kGroup, kGroup, false);
    write_slot(c, slot_ptr(c, level, g, true), gv);
  }
}

// State at bar s0-1: composite of everything from the group holding the
// row's first segment (which resets the state) up to this segment.
template <int R>
__device__ __forceinline__ void lookback(const Seg& c, int level, double (&y)[R]) {
  const int first = c.gseg - c.seg;
  const int G = c.gseg / kGroup;
  Aff<R> acc;
  aff_id(acc);
  for (int g0 = first / kGroup; g0 < G; g0 += kWave) {
    Aff<R> ch = gather<R>(c, level, g0, min(kWave, G - g0), true);
    aff_after(acc, ch);
    acc = ch;
  }
  const int i0 = max(first, G * kGroup);
  if (i0 < c.gseg) {
    Aff<R> ch = gather<R>(c, level, i0, c.gseg - i0, false);
    aff_after(acc, ch);
    acc = ch;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) y[r] = acc.b[r];
}

// Carry for one level: composes the lane-local transforms `loc` (identity on
// inactive bars) with the segment's incoming state.  `Ls` is the level's seed
// bar and `yL` its state there (valid when Ls lies in this segment).
// Returns the state entering each lane's first bar in `yin`.
template <int R>
__device__ __forceinline__ void level_carry(const Seg& c, int level, int Ls, const double (&yL)[R], Aff<R> loc,
                            double (&yin)[R]) {
  Aff<R> tot;
  wave_scan_excl(loc, tot);
  const double qnan = qnan_d();
  if (Ls >= c.s1) {
    // nothing valid yet: the row restarts from "undefined"
    Aff<R> t;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      t.A[r] = -0.0;
      t.b[r] = qnan;
      yin[r] = qnan;
    }
    if (c.tree) publish_tree(c, level, t);
    return;
  }
  if (Ls >= c.s0) {
    // the seed lies here: this segment resets the state
    Aff<R> t;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      yin[r] = aff_apply(loc, r, yL[r]);
      t.A[r] = -0.0;
      t.b[r] = aff_apply(tot, r, yL[r]);
    }
    if (c.tree) publish_tree(c, level, t);
    return;
  }
  publish_tree(c, level, tot);
  double ys[R];
  lookback<R>(c, level, ys);
#pragma unroll
  for (int r = 0; r < R; ++r) yin[r] = aff_apply(loc, r, ys[r]);
}

// ------------------------------------------------- seeded offset scans --
// A recurrence y_t = beta*y_{t-1} + u_t that starts at a seed bar: bars
// before the seed contribute nothing (offset 0), the seed bar sets y to the
// seed, later bars apply the step.  The composite of any K-bar run of a
// row is then y -> beta^K * y + b -- for the runs holding or preceding the
// seed too, because everything before the seed is exactly 0 -- so scans and
// look-back carry only the offsets b; the multipliers are known powers.
// Rows stay independent: a segment's look-back reads only its own row's
// slots (whole groups only once they lie inside the row).
//
// Slots hold (flag, b[0..R-1]).  Powers of nonzero factors that underflow
// are clamped to the smallest normal double so a carried Inf or NaN
// survives them as in the sequential recurrence; a zero factor stays zero
// (0 * NaN is NaN, TA-Lib's literal fma).
__device__ __forceinline__ double pow_clamp(double p, double beta) {
  return (p == 0.0 && beta != 0.0) ? 2.2250738585072014e-308 : p;
}

// beta^n for a small n (n >= 0), by squaring; exact for n <= 2.
__device__ __forceinline__ double pow_n(double beta, int n) {
  double r = 1.0, b = beta;
  while (n) {
    if (n & 1) r = pow_clamp(r * b, beta);
    b = pow_clamp(b * b, beta);
    n >>= 1;
  }
  return r;
}

// Inclusive lane scan of (A, b) pairs (lane 0 earliest).  A starts as each
// lane's multiplier; CLAMP guards products that may underflow.
template <int R, bool CLAMP, int CTRL, int RM, int BM>
__device__ __forceinline__ void bscan_step(double (&b)[R], double (&A)[R], const double (&beta)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double As = dpp_d<CTRL, RM, BM>(1.0, A[r]);
    const double bs = dpp_d<CTRL, RM, BM>(0.0, b[r]);
    b[r] = fma(A[r], bs, b[r]);
    A[r] = CLAMP ? pow_clamp(A[r] * As, beta[r]) : A[r] * As;
  }
}

template <int R, bool CLAMP>
__device__ __forceinline__ void bscan_incl(double (&b)[R], double (&A)[R], const double (&beta)[R]) {
  bscan_step<R, CLAMP, kRowShr1, 0xf, 0xf>(b, A, beta);
  bscan_step<R, CLAMP, kRowShr2, 0xf, 0xf>(b, A, beta);
  bscan_step<R, CLAMP, kRowShr4, 0xf, 0xf>(b, A, beta);
  bscan_step<R, CLAMP, kRowShr8, 0xf, 0xf>(b, A, beta);
  bscan_step<R, CLAMP, kRowBcast15, 0xa, 0xf>(b, A, beta);
  bscan_step<R, CLAMP, kRowBcast31, 0xc, 0xf>(b, A, beta);
}

template <int R>
__device__ __forceinline__ void write_slot_b(const Seg& c, double* s, const double (&b)[R]) {
  if (lane_id() == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) st_rel_d(s + 1 + r, b[r]);
    vmem_order();
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(s), (unsigned long long)c.epoch,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Lanes 0..cnt-1 load slots idx0.. (spinning on this launch's epoch); the
// wave folds them oldest first with the per-slot multiplier P.  Returns the
// folded offsets and P^cnt (both uniform).
template <int R>
__device__ __forceinline__ void gather_b(const Seg& c, int level, int idx0, int cnt, bool group,
                                         const double (&P)[R], const double (&beta)[R],
                                         double (&tb)[R], double (&tp)[R]) {
  const int lane = lane_id();
  const bool mine = lane < cnt;
  const double* s = mine ? slot_ptr(c, level, idx0 + lane, group) : nullptr;
  if (mine) {
    while ((unsigned)ld_flag(s) != c.epoch) __builtin_amdgcn_s_sleep(1);
  }
  vmem_order();
  double b[R], A[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    b[r] = mine ? ld_rel_d(s + 1 + r) : 0.0;
    A[r] = mine ? P[r] : 1.0;
  }
  bscan_incl<R, true>(b, A, beta);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    tb[r] = bcast_d(b[r], 63);
    tp[r] = bcast_d(A[r], 63);
  }
}

// Offsets of this segment's row up to bar s0-1 (0 before the seeds).
template <int R>
__device__ __forceinline__ void lookback_b(const Seg& c, int level, const double (&Pseg)[R],
                                           const double (&Pgrp)[R], const double (&beta)[R],
                                           double (&y)[R]) {
  const int first = c.gseg - c.seg;
  const int G = c.gseg / kGroup;
#pragma unroll
  for (int r = 0; r < R; ++r) y[r] = 0.0;
  double tb[R], tp[R];
  auto fold = [&]() {
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = fma(tp[r], y[r], tb[r]);
  };
  // leading segments of a group the row shares with the previous row
  const int hEnd = min(c.gseg, ((first + kGroup - 1) / kGroup) * kGroup);
  if (first < hEnd) {
    gather_b<R>(c, level, first, hEnd - first, false, Pseg, beta, tb, tp);
    fold();
  }
  for (int g0 = hEnd / kGroup; g0 < G; g0 += kWave) {
    gather_b<R>(c, level, g0, min(kWave, G - g0), true, Pgrp, beta, tb, tp);
    fold();
  }
  const int i0 = max(hEnd, G * kGroup);
  if (i0 < c.gseg) {
    gather_b<R>(c, level, i0, c.gseg - i0, false, Pseg, beta, tb, tp);
    fold();
  }
}

// Sum over the wave of N per-lane values (lane order), result in every lane.
template <int N>
__device__ __forceinline__ void wave_total(double (&v)[N]) {
#define TCA_ADD_STEP(CTRL, RM)                                         \
  _Pragma("unroll") for (int r = 0; r < N; ++r) v[r] += dpp_d<CTRL, RM, 0xf>(0.0, v[r]);
  TCA_ADD_STEP(kRowShr1, 0xf)
  TCA_ADD_STEP(kRowShr2, 0xf)
  TCA_ADD_STEP(kRowShr4, 0xf)
  TCA_ADD_STEP(kRowShr8, 0xf)
  TCA_ADD_STEP(kRowBcast15, 0xa)
  TCA_ADD_STEP(kRowBcast31, 0xc)
#undef TCA_ADD_STEP
#pragma unroll
  for (int r = 0; r < N; ++r) v[r] = bcast_d(v[r], 63);
}

// One level of seeded recurrences.  `b` holds each lane's K-bar offsets
// (folded from 0), `P` = beta^K per recurrence.  Returns in `yin` the state
// entering each lane's first bar.  Publishes this segment's offsets (and,
// for the last segment of a full group, the group's) when the launch chains
// segments.
template <int R>
__device__ __forceinline__ void bcarry(const Seg& c, int level, double (&b)[R],
                                       const double (&P)[R], const double (&beta)[R],
                                       double (&yin)[R]) {
  double A[R];
#pragma unroll
  for (int r = 0; r < R; ++r) A[r] = P[r];
  bscan_incl<R, false>(b, A, beta);
  double bex[R], Aex[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    bex[r] = dpp_d<kWaveShr1, 0xf, 0xf>(0.0, b[r]);
    Aex[r] = dpp_d<kWaveShr1, 0xf, 0xf>(1.0, A[r]);
  }
  if (!c.tree) {
#pragma unroll
    for (int r = 0; r < R; ++r) yin[r] = bex[r];
    return;
  }
  double tot[R], Pseg[R], Pgrp[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    tot[r] = bcast_d(b[r], 63);
    Pseg[r] = pow_clamp(bcast_d(A[r], 63), beta[r]);
    Pgrp[r] = pow_n(Pseg[r], kGroup);
    Pgrp[r] = (Pgrp[r] == 0.0 && beta[r] != 0.0) ? 2.2250738585072014e-308 : Pgrp[r];
  }
  write_slot_b(c, slot_ptr(c, level, c.gseg, false), tot);
  if ((c.gseg % kGroup) == kGroup - 1) {
    double gb[R], gp[R];
    gather_b<R>(c, level, c.gseg - (kGroup - 1), kGroup, false, Pseg, beta, gb, gp);
    write_slot_b(c, slot_ptr(c, level, c.gseg / kGroup, true), gb);
  }
  if (c.seg == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) yin[r] = bex[r];
    return;
  }
  double y[R];
  lookback_b<R>(c, level, Pseg, Pgrp, beta, y);
#pragma unroll
  for (int r = 0; r < R; ++r) yin[r] = fma(Aex[r], y[r], bex[r]);
}

// ------------------------------------------------------------ scheduling --
// Ticket -> segment.  Dense rows: consecutive tickets walk a row's segments
// in order; ragged rows use the precomputed first-segment table.
__device__ __forceinline__ bool resolve(const Geometry& g, int seg_size, int ticket, Seg& c) {
  if (ticket >= g.totalSegs) return false;
  int row, seg, nseg;
  if (g.segStart == nullptr) {
    row = ticket / g.segsPerRow;
    seg = ticket - row * g.segsPerRow;
    nseg = g.segsPerRow;
    c.rowOff = (long long)row * g.ld;
    c.n = g.len;
  } else {
    int lo = 0, hi = g.rows - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (g.segStart[mid] <= ticket)
        lo = mid;
      else
        hi = mid - 1;
    }
    row = lo;
    seg = ticket - g.segStart[row];
    nseg = g.segStart[row + 1] - g.segStart[row];
    c.rowOff = g.offsets[row];
    c.n = (int)(g.offsets[row + 1] - g.offsets[row]);
  }
  c.seg = seg;
  c.nseg = nseg;
  c.gseg = ticket;
  c.s0 = seg * seg_size;
  c.s1 = min(c.n, c.s0 + seg_size);
  c.totalSegs = g.totalSegs;
  return true;
}

// Launch epoch, read by every wave at its start.  Segments are ordered by
// workgroup id: each XCD's dispatcher issues its workgroups in id order, so
// every predecessor a wave waits for is resident or already finished.
__device__ __forceinline__ unsigned read_epoch(const Ws& ws) {
  unsigned e = 0;
  if (lane_id() == 0)
    e = __hip_atomic_load(&ws.hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return (unsigned)__builtin_amdgcn_readfirstlane((int)e);
}

// Block epilogue (thread 0, after the block's waves are done; b is the
// block's index among the nblocks blocks that share the workspace): the block
// completing the last stripe resets the counters and bumps the epoch, so the
// next launch on this workspace (ordered after this one by the stream)
// starts clean.  Relaxed atomics suffice: kernel end publishes every write.
__device__ __forceinline__ void finish_block(const Ws& ws, unsigned nblocks,
                                             unsigned b = blockIdx.x) {
  const unsigned st = b % kStripes;
  const unsigned inStripe = (nblocks - st + kStripes - 1) / kStripes;
  const unsigned d = __hip_atomic_fetch_add(&ws.hdr->stripe[st][0], 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
  if (d + 1 != inStripe) return;
  __hip_atomic_store(&ws.hdr->stripe[st][0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned nst = nblocks < (unsigned)kStripes ? nblocks : (unsigned)kStripes;
  const unsigned s2 = __hip_atomic_fetch_add(&ws.hdr->stripesDone, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
  if (s2 + 1 != nst) return;
  __hip_atomic_store(&ws.hdr->stripesDone, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned e = __hip_atomic_load(&ws.hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const unsigned ne = (e + 1u) & 0x3fffffffu;
  __hip_atomic_store(&ws.hdr->epoch, ne ? ne : 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float qnanf_() { return __int_as_float(0x7fc00000); }
__device__ __forceinline__ double qnan_() { return __longlong_as_double(0x7ff8000000000000ll); }

}  // namespace tca
