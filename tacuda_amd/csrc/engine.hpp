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
//    exclusive carry, and rows longer than one segment chain segments with a
//    decoupled look-back whose flags carry a device-side epoch, so the
//    workspace needs no clearing between launches or graph replays.
//  * Arithmetic of every recurrence is FP64; lanes replay the TA-Lib update
//    order from their carry so the outputs match the sequential reference to
//    the last float32 bit except where a carry differs in the last FP64 ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tca {

constexpr int kWave = 64;

// --------------------------------------------------------------- workspace --
struct WsHeader {
  unsigned int ticket;  // wave tickets of the running launch
  unsigned int done;    // finished waves of the running launch
  unsigned int epoch;   // launch generation, bumped by the last wave
  unsigned int pad[61];
};

// One look-back slot: flag = epoch*4 + state, payload = (A, b) per recurrence.
constexpr unsigned kAgg = 1, kIncl = 2;
constexpr int kSlotDoubles = 7;  // up to 3 recurrences (A,b) + 1 spare
struct alignas(64) Slot {
  unsigned long long flag;
  double v[kSlotDoubles];
};

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

struct Ws {
  WsHeader* hdr;
  Slot* slots;  // levels * totalSegs
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

template <int R>
__device__ __forceinline__ void aff_after(const Aff<R>& e, Aff<R>& l) {
  // l := l o e
#pragma unroll
  for (int r = 0; r < R; ++r) {
    l.b[r] = fma(l.A[r], e.b[r], l.b[r]);
    l.A[r] = l.A[r] * e.A[r];
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
  return fma(t.A[r], y, t.b[r]);
}

// ----------------------------------------------------------- look-back --
__device__ __forceinline__ unsigned long long ld_acq(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
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
};

template <int R>
__device__ void publish(const Seg& c, int level, unsigned state, const Aff<R>& v) {
  if (c.nseg <= 1) return;
  Slot* s = c.ws.slots + (size_t)level * c.totalSegs + c.gseg;
  if (lane_id() == 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      st_rel_d(&s->v[2 * r], v.A[r]);
      st_rel_d(&s->v[2 * r + 1], v.b[r]);
    }
    __hip_atomic_store(&s->flag, (unsigned long long)c.epoch * 4ull + state, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Composite of all predecessors of this segment within its row: returns the
// state at bar s0-1 per recurrence.  Stops at the nearest INCLUSIVE slot.
template <int R>
__device__ void lookback(const Seg& c, int level, double (&y)[R]) {
  const int lane = lane_id();
  const int first = c.gseg - c.seg;
  Aff<R> acc;
  aff_id(acc);
  int base = c.gseg - 1;
  const Slot* slots = c.ws.slots + (size_t)level * c.totalSegs;
  for (;;) {
    const int pred = base - lane;
    const bool valid = pred >= first;
    unsigned st = kIncl;
    if (valid) {
      const unsigned long long want = (unsigned long long)c.epoch;
      unsigned long long f;
      do {
        f = ld_acq(&slots[pred].flag);
      } while ((f >> 2) != want);
      st = (unsigned)(f & 3ull);
    }
    const unsigned long long inclMask = __ballot(st == kIncl);
    const int j = inclMask ? (__ffsll((long long)inclMask) - 1) : 64;
    Aff<R> v;
    aff_id(v);
    if (valid && lane <= j) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v.A[r] = ld_rel_d(&slots[pred].v[2 * r]);
        v.b[r] = ld_rel_d(&slots[pred].v[2 * r + 1]);
      }
    } else if (!valid && lane <= j) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        v.A[r] = 0.0;
        v.b[r] = __longlong_as_double(0x7ff8000000000000ll);
      }
    }
    // reduce: higher lanes are earlier in time
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      Aff<R> o;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        o.A[r] = __shfl_down(v.A[r], off);
        o.b[r] = __shfl_down(v.b[r], off);
      }
      if (lane + off < kWave) aff_after(o, v);
    }
    Aff<R> chunk;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      chunk.A[r] = bcast_d(v.A[r], 0);
      chunk.b[r] = bcast_d(v.b[r], 0);
    }
    aff_after(chunk, acc);
    if (j < 64) break;
    base -= 64;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) y[r] = acc.b[r];
}

// Carry for one level: composes the lane-local transforms `loc` (identity on
// inactive bars) with the segment's incoming state.  `Ls` is the level's seed
// bar and `yL` its state there (valid when Ls lies in this segment).
// Returns the state entering each lane's first bar in `yin`.
template <int R>
__device__ void level_carry(const Seg& c, int level, int Ls, const double (&yL)[R], Aff<R> loc,
                            double (&yin)[R]) {
  Aff<R> tot;
  wave_scan_excl(loc, tot);
  const double qnan = __longlong_as_double(0x7ff8000000000000ll);
  if (Ls >= c.s1) {
    Aff<R> inc;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      inc.A[r] = 0.0;
      inc.b[r] = qnan;
      yin[r] = qnan;
    }
    publish(c, level, kIncl, inc);
    return;
  }
  double ys[R];
  if (Ls >= c.s0) {
#pragma unroll
    for (int r = 0; r < R; ++r) ys[r] = yL[r];
  } else {
    publish(c, level, kAgg, tot);
    lookback<R>(c, level, ys);
  }
  Aff<R> inc;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    yin[r] = aff_apply(loc, r, ys[r]);
    inc.A[r] = 0.0;
    inc.b[r] = aff_apply(tot, r, ys[r]);
  }
  publish(c, level, kIncl, inc);
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

// Wave prologue: take a ticket (in launch order, which guarantees every
// predecessor is resident) and read the launch epoch.
__device__ __forceinline__ int take_ticket(const Ws& ws, unsigned& epoch) {
  int t = 0;
  unsigned e = 0;
  if (lane_id() == 0) {
    t = (int)__hip_atomic_fetch_add(&ws.hdr->ticket, 1u, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    e = __hip_atomic_load(&ws.hdr->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  t = __builtin_amdgcn_readfirstlane(t);
  epoch = (unsigned)__builtin_amdgcn_readfirstlane((int)e);
  return t;
}

// Wave epilogue: the last wave of the launch resets the counters and bumps
// the epoch so the next launch on this workspace starts clean.
__device__ __forceinline__ void finish_wave(const Ws& ws, unsigned totalWaves) {
  if (lane_id() == 0) {
    const unsigned d = __hip_atomic_fetch_add(&ws.hdr->done, 1u, __ATOMIC_ACQ_REL,
                                              __HIP_MEMORY_SCOPE_AGENT);
    if (d + 1 == totalWaves) {
      __hip_atomic_store(&ws.hdr->ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ws.hdr->done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned e = __hip_atomic_load(&ws.hdr->epoch, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&ws.hdr->epoch, (e + 1u) & 0x3fffffffu ? (e + 1u) & 0x3fffffffu : 1u,
                         __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__device__ __forceinline__ float qnanf_() { return __int_as_float(0x7fc00000); }
__device__ __forceinline__ double qnan_() { return __longlong_as_double(0x7ff8000000000000ll); }

}  // namespace tca
