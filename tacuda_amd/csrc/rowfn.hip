// TA-Lib 0.8.3 functions the reference serves from its formula backend
// (src/compat/Talib08.cu:47-103, one work-item per row running the TA_*
// routine in double): CUMSUM, PVT, RMA, ZLEMA, EFI, VWMA, ER, VHF, DPO,
// CHOP, ADR, AO, QSTICK, DONCHIAN, MARKETFI, WAD, MCGD, RVOL, CMF, EMV,
// FOSC, ERI, PERCENTRANK, SI, ASI_WILDER (TA_ASI), VORTEX and CMOU, each in
// host, device, batch and host-batch forms.  Reference entry points:
// include/tacuda.h (ct_<name> and ct_<name>_batch[_d], cited per
// declaration), include/tacuda_device.h (ct_<name>_d).
//
// Semantics are the backend's: per row, the leading bars where any input is
// NaN are skipped (the last bar is always kept), the routine runs in double
// from the first remaining bar, outputs are rounded to float, and every bar
// before the routine's first output is NaN.  A period outside TA-Lib's range
// is CT_STATUS_INVALID_ARGUMENT; a row no longer than the look-back is all
// NaN.
//
// These are sequential recurrences and running sums whose TA-Lib rounding
// a parallel scan would not reproduce, so each row is one work-item, as in
// the reference.  Few long rows get a wave each (lane 0 works, so the rows'
// loads do not queue behind one another in one CU); otherwise 64 rows share
// a wave, one per lane.  The state machines are
// restated from the vendored ta_*.c routines (src/compat/talib08_formulas.inc)
// in the same order of floating-point operations, with -ffp-contract=off
// and fma() where TA-Lib writes it, so the outputs equal TA-Lib's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/tacuda.h"
#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"

namespace tca {
namespace {

enum RowFn {
  kCumsum, kPvt, kRma, kZlema, kEfi, kVwma, kEr, kVhf, kDpo, kChop, kAdr, kAo, kQstick,
  kDonchian, kMarketfi, kWad, kMcgd, kRvol, kCmf, kEmv, kFosc, kEri, kPercentrank, kSi,
  kAsiWilder, kVortex, kCmou
};

struct RowSpec {
  int nin, nout, npar;
  int lo[2], hi[2];
  double dlo, dhi;  // the real parameter's range (dlo > dhi: none)
};

constexpr double kRealMax = 3e37;  // TA_REAL_MAX

// inputs, outputs, periods and their valid ranges (TA-Lib's checks)
constexpr RowSpec kSpec[] = {
    {1, 1, 0, {0, 0}, {0, 0}, 1, 0},            // CUMSUM
    {2, 1, 0, {0, 0}, {0, 0}, 1, 0},            // PVT (close, volume)
    {1, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // RMA
    {1, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // ZLEMA
    {2, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // EFI (close, volume)
    {2, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // VWMA (input, volume)
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // ER
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // VHF
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // DPO
    {3, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // CHOP (high, low, close)
    {2, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // ADR (high, low)
    {2, 1, 2, {2, 2}, {100000, 100000}, 1, 0},  // AO (high, low; fast, slow)
    {2, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // QSTICK (open, close)
    {2, 3, 1, {2, 0}, {100000, 0}, 1, 0},       // DONCHIAN (high, low -> upper, middle, lower)
    {3, 1, 0, {0, 0}, {0, 0}, 1, 0},            // MARKETFI (high, low, volume)
    {3, 1, 0, {0, 0}, {0, 0}, 1, 0},            // WAD (high, low, close)
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // MCGD
    {1, 1, 1, {1, 0}, {100000, 0}, 1, 0},       // RVOL (volume)
    {4, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // CMF (high, low, close, volume)
    {3, 1, 1, {1, 0}, {100000, 0}, 1.0, kRealMax},  // EMV (high, low, volume; divisor)
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // FOSC
    {3, 2, 1, {1, 0}, {100000, 0}, 1, 0},       // ERI (high, low, close -> bull, bear)
    {1, 1, 1, {2, 0}, {10000, 0}, 1, 0},        // PERCENTRANK
    {4, 1, 0, {0, 0}, {0, 0}, 1e-8, kRealMax},  // SI (open, high, low, close; limit move)
    {4, 1, 0, {0, 0}, {0, 0}, 1e-8, kRealMax},  // ASI_WILDER (the same)
    {3, 2, 1, {1, 0}, {100000, 0}, 1, 0},       // VORTEX (high, low, close -> VI+, VI-)
    {1, 1, 1, {2, 0}, {100000, 0}, 1, 0},       // CMOU
};

struct RowArgs {
  const float* in[4];
  float* out[3];
  int len, rows, p0, p1;
  double d0;
};

typedef __attribute__((address_space(3))) float lds_float;

// One row's view: inputs and outputs offset to the first bar after the
// leading-NaN prefix; n bars remain.  Outputs collect in an LDS run of up to
// cap values per output and go to memory when the run is full, breaks or
// ends: on gfx950 a load waits for every earlier store of its wave (one
// in-order counter), so storing each bar as it is made would put a store's
// full round trip between consecutive loads of a sequential routine.  The
// routines write each output's bars in increasing order.
struct Row {
  const float* x[4];
  float* y[3];
  int n;
  lds_float* buf;  // [3][cap]
  int cap;
  mutable int t0[3];
  mutable int cnt[3];
  __device__ double in(int j, int t) const { return (double)x[j][t]; }
  __device__ void flush(int k) const {
    for (int i = 0; i < cnt[k]; ++i) y[k][t0[k] + i] = buf[k * cap + i];
    cnt[k] = 0;
  }
  __device__ void put(int k, int t, double v) const {
    if (cnt[k] > 0 && t != t0[k] + cnt[k]) flush(k);
    if (cnt[k] == 0) t0[k] = t;
    buf[k * cap + cnt[k]] = (float)v;
    if (++cnt[k] == cap) flush(k);
  }
  __device__ void flush_all() const {
    for (int k = 0; k < 3; ++k)
      if (cnt[k] > 0) flush(k);
  }
};

__device__ __forceinline__ bool finite_d(double v) {
  return v <= 1.7976931348623157e308 && v >= -1.7976931348623157e308;
}

// EMA weights as TA-Lib forms them (ta_EMA.c): beta = (p-1)/(p+1) and k =
// 1 - beta, with beta recomputed as 1 - k when below one half.
__device__ __forceinline__ void ema_bk(int p, double& b, double& k) {
  b = (double)(p - 1) / (double)(p + 1);
  k = 1.0 - b;
  if (b < 0.5) b = 1.0 - k;
}

// ------------------------------------------------------------ routines --
// Each writes y[k][t] for t at or past its first output; the kernel has
// already filled the row with NaN.

__device__ void f_cumsum(const Row& r) {
  double total = 0.0;
  for (int t = 0; t < r.n; ++t) {
    total += r.in(0, t);
    r.put(0, t, total);
  }
}

// ta_PVT: acc += (c - prev) / prev * volume while prev != 0
__device__ void f_pvt(const Row& r) {
  double acc = 0.0, prev = r.in(0, 0);
  for (int t = 0; t < r.n; ++t) {
    const double c = r.in(0, t);
    if (prev != 0.0) acc += (c - prev) / prev * r.in(1, t);
    r.put(0, t, acc);
    prev = c;
  }
}

// ta_RMA: SMA seed of the first p bars, then fma(beta, prev, alpha x)
__device__ void f_rma(const Row& r, int p) {
  if (r.n < p) return;
  const double wb = (double)(p - 1) / (double)p, wa = 1.0 - wb;
  double s = 0.0;
  for (int t = 0; t < p; ++t) s += r.in(0, t);
  double v = s / p;
  r.put(0, p - 1, v);
  for (int t = p; t < r.n; ++t) {
    v = fma(wb, v, wa * r.in(0, t));
    r.put(0, t, v);
  }
}

// ta_ZLEMA: EMA of 2 x[t] - x[t - lag], lag = (p-1)/2, seeded by the mean of
// its first p terms
__device__ void f_zlema(const Row& r, int p) {
  const int lag = (p - 1) / 2, first = lag + p - 1;
  if (r.n <= first) return;
  if (p == 1) {
    for (int t = 0; t < r.n; ++t) r.put(0, t, r.in(0, t));
    return;
  }
  double b, k;
  ema_bk(p, b, k);
  double s = 0.0;
  for (int i = 0; i < p; ++i) s += 2.0 * r.in(0, lag + i) - r.in(0, i);
  double v = s / p;
  r.put(0, first, v);
  for (int t = first + 1; t < r.n; ++t) {
    v = fma(b, v, k * (2.0 * r.in(0, t) - r.in(0, t - lag)));
    r.put(0, t, v);
  }
}

// ta_EFI: EMA of the force (c[t] - c[t-1]) v[t], seeded by the mean of the
// first p forces; p = 1 is the raw force
__device__ void f_efi(const Row& r, int p) {
  if (r.n <= p) return;
  if (p == 1) {
    for (int t = 1; t < r.n; ++t) r.put(0, t, (r.in(0, t) - r.in(0, t - 1)) * r.in(1, t));
    return;
  }
  double b, k;
  ema_bk(p, b, k);
  double s = 0.0;
  for (int t = 1; t <= p; ++t) s += (r.in(0, t) - r.in(0, t - 1)) * r.in(1, t);
  double v = s / p;
  r.put(0, p, v);
  for (int t = p + 1; t < r.n; ++t) {
    v = fma(b, v, k * ((r.in(0, t) - r.in(0, t - 1)) * r.in(1, t)));
    r.put(0, t, v);
  }
}

// ta_VWMA: running sums of price x volume and volume; a window of zero
// volumes restarts the sums and gives 0/0
__device__ void f_vwma(const Row& r, int p) {
  if (r.n < p) return;
  if (p == 1) {
    for (int t = 0; t < r.n; ++t) r.put(0, t, r.in(0, t));
    return;
  }
  double spv = 0.0, sv = 0.0;
  int zeros = 0;
  for (int t = 0; t < p - 1; ++t) {
    const double v = r.in(1, t);
    spv += r.in(0, t) * v;
    sv += v;
    zeros += fabs(v) <= 0.0;
  }
  for (int t = p - 1; t < r.n; ++t) {
    const double v = r.in(1, t);
    spv += r.in(0, t) * v;
    sv += v;
    zeros += fabs(v) <= 0.0;
    const int o = t - (p - 1);
    const double ov = r.in(1, o), opv = r.in(0, o) * ov;
    double a = spv, c = sv;
    if (zeros >= p) {
      a = 0.0;
      c = 0.0;
      spv = 0.0 - opv;
      sv = 0.0 - ov;
    } else {
      spv -= opv;
      sv -= ov;
    }
    r.put(0, t, a / (double)p / (c / (double)p));
    zeros -= fabs(ov) <= 0.0;
  }
}

// ta_ER: |x[t] - x[t-p]| over the running sum of |one-bar changes|, clamped
// to 1; a full window of unchanged bars restarts the sum and gives 1
__device__ void f_er(const Row& r, int p) {
  if (r.n <= p) return;
  double sum = 0.0;
  int run = 0;
  for (int t = 1; t <= p; ++t) {
    const double d = r.in(0, t - 1) - r.in(0, t);
    sum += fabs(d);
    run = fabs(d) <= 0.0 ? run + 1 : 0;
  }
  double trailing = r.in(0, 0);
  {
    const double roc = r.in(0, p) - trailing;
    double o = 1.0;
    if (!(sum <= 0.0 || sum <= roc)) {
      o = fabs(roc / sum);
      if (o > 1.0) o = 1.0;
    }
    r.put(0, p, o);
  }
  for (int t = p + 1; t < r.n; ++t) {
    const double cur = r.in(0, t), old = r.in(0, t - p), prev = r.in(0, t - 1);
    const double roc = cur - old;
    sum -= fabs(trailing - old);
    sum += fabs(cur - prev);
    run = fabs(cur - prev) <= 0.0 ? run + 1 : 0;
    double o;
    if (run >= p) {
      run = p;
      sum = 0.0;
      o = 1.0;
    } else if (sum <= 0.0 || sum <= roc) {
      o = 1.0;
    } else {
      o = fabs(roc / sum);
      if (o > 1.0) o = 1.0;
    }
    r.put(0, t, o);
    trailing = old;
  }
}

// ta_VHF: (max - min) / sum |changes| over the p + 1 bars ending at t, every
// window evaluated afresh oldest first (the extremes seeded with x[t])
__device__ __forceinline__ double pt_vhf(const Row& r, int t, int p) {
  double hi = r.in(0, t), lo = hi, prev = r.in(0, t - p), chg = 0.0;
  for (int s = t - p + 1; s <= t; ++s) {
    const double v = r.in(0, s);
    chg += fabs(v - prev);
    if (v > hi) hi = v;
    if (v < lo) lo = v;
    prev = v;
  }
  return chg > 0.0 ? (hi - lo) / chg : 0.0;
}

// ta_DPO: x[t - (p/2 + 1)] minus the running p-bar mean ending at t
__device__ void f_dpo(const Row& r, int p) {
  const int shift = p / 2 + 1, first = (p - 1) > shift ? (p - 1) : shift;
  if (r.n <= first) return;
  double s = 0.0;
  for (int t = first - (p - 1); t < first; ++t) s += r.in(0, t);
  for (int t = first; t < r.n; ++t) {
    s += r.in(0, t);
    const double tot = s;
    s -= r.in(0, t - (p - 1));
    r.put(0, t, r.in(0, t - shift) - tot / (double)p);
  }
}

// ta_CHOP: 100 log10(sum TR / (max high - min low)) / log10(p) over the p
// bars ending at t (true ranges against the close before each bar), 100
// when the range or the sum is not positive
__device__ __forceinline__ double pt_chop(const Row& r, int t, int p) {
  const double lp = log10((double)p);
  double hi = r.in(0, t), lo = r.in(1, t), str = 0.0, pc = r.in(2, t - p);
  for (int s = t - p + 1; s <= t; ++s) {
    const double h = r.in(0, s), l = r.in(1, s);
    const double th = pc > h ? pc : h, tl = pc < l ? pc : l;
    str += th - tl;
    if (h > hi) hi = h;
    if (l < lo) lo = l;
    pc = r.in(2, s);
  }
  return (hi - lo <= 0.0 || str <= 0.0) ? 100.0 : 100.0 * (log10(str / (hi - lo)) / lp);
}

// ta_ADR (high - low) and ta_QSTICK (close - open): the running p-bar mean
// of a - b
__device__ void f_mean_diff(const Row& r, int ia, int ib, int p) {
  if (r.n < p) return;
  double s = 0.0;
  for (int t = 0; t < p - 1; ++t) s += r.in(ia, t) - r.in(ib, t);
  for (int t = p - 1; t < r.n; ++t) {
    s += r.in(ia, t) - r.in(ib, t);
    const double tot = s;
    const int o = t - (p - 1);
    s -= r.in(ia, o) - r.in(ib, o);
    r.put(0, t, tot / (double)p);
  }
}

// ta_AO: the fast minus the slow running mean of the median price
__device__ void f_ao(const Row& r, int f, int sl) {
  const int first = (f > sl ? f : sl) - 1;
  if (r.n <= first) return;
  double sf = 0.0, ss = 0.0;
  for (int t = 0; t < first; ++t) {
    const double m = (r.in(0, t) + r.in(1, t)) / 2.0;
    if (t >= first - (f - 1)) sf += m;
    if (t >= first - (sl - 1)) ss += m;
  }
  for (int t = first; t < r.n; ++t) {
    const double m = (r.in(0, t) + r.in(1, t)) / 2.0;
    sf += m;
    ss += m;
    const double o = sf / (double)f - ss / (double)sl;
    const int of = t - (f - 1), os = t - (sl - 1);
    sf -= (r.in(0, of) + r.in(1, of)) / 2.0;
    ss -= (r.in(0, os) + r.in(1, os)) / 2.0;
    r.put(0, t, o);
  }
}

// ta_DONCHIAN: window extremes kept with their bar index; when the extreme
// leaves the window it is searched again from the oldest bar (first strict
// improvement wins), otherwise a new bar at or beyond it replaces it.  Mid =
// (upper + lower) / 2.
__device__ void f_donchian(const Row& r, int p) {
  int hiAt = -1, loAt = -1;
  double hi = 0.0, lo = 0.0;
  for (int t = p - 1; t < r.n; ++t) {
    const int o = t - (p - 1);
    if (hiAt < o) {
      hiAt = o;
      hi = r.in(0, o);
      for (int s = o + 1; s <= t; ++s) {
        const double v = r.in(0, s);
        if (v > hi) {
          hiAt = s;
          hi = v;
        }
      }
    } else if (r.in(0, t) >= hi) {
      hiAt = t;
      hi = r.in(0, t);
    }
    if (loAt < o) {
      loAt = o;
      lo = r.in(1, o);
      for (int s = o + 1; s <= t; ++s) {
        const double v = r.in(1, s);
        if (v < lo) {
          loAt = s;
          lo = v;
        }
      }
    } else if (r.in(1, t) <= lo) {
      loAt = t;
      lo = r.in(1, t);
    }
    r.put(0, t, hi);
    r.put(1, t, (hi + lo) / 2.0);
    r.put(2, t, lo);
  }
}

__device__ __forceinline__ double pt_marketfi(const Row& r, int t) {
  const double v = r.in(2, t);
  return v != 0.0 ? (r.in(0, t) - r.in(1, t)) / v : 0.0;
}

// ta_WAD: accumulation against the true low on up bars, the true high on
// down bars, nothing on unchanged closes
__device__ void f_wad(const Row& r) {
  double sum = 0.0, pc = r.in(2, 0);
  for (int t = 0; t < r.n; ++t) {
    const double c = r.in(2, t);
    if (c > pc) {
      const double l = r.in(1, t);
      sum += c - (pc < l ? pc : l);
    } else if (c < pc) {
      const double h = r.in(0, t);
      sum += c - (pc > h ? pc : h);
    }
    r.put(0, t, sum);
    pc = c;
  }
}

// ta_MCGD: md += (x - md) / (p (x/md)^4), a non-finite step keeps md
__device__ void f_mcgd(const Row& r, int p) {
  if (r.n < p) return;
  const double per = (double)p;
  double md = r.in(0, 0);
  for (int t = 1; t < r.n; ++t) {
    const double x = r.in(0, t), q = x / md;
    const double nm = md + (x - md) / (per * (q * q * (q * q)));
    if (finite_d(nm)) md = nm;
    if (t >= p - 1) r.put(0, t, md);
  }
}

// ta_RVOL: volume over the running mean of the p bars before it; a full
// window of zero volumes resets the sum
__device__ void f_rvol(const Row& r, int p) {
  if (r.n <= p) return;
  double s = 0.0;
  int run = 0;
  for (int t = 0; t < p; ++t) {
    const double v = r.in(0, t);
    s += v;
    run = fabs(v) <= 0.0 ? run + 1 : 0;
  }
  for (int t = p; t < r.n; ++t) {
    const double base = s / (double)p, old = r.in(0, t - p), v = r.in(0, t);
    r.put(0, t, v / base);
    s -= old;
    s += v;
    run = fabs(v) <= 0.0 ? run + 1 : 0;
    if (run >= p) s = 0.0;
  }
}

// ta_CMF: running sums of money-flow volume and volume (a bar's flow is a
// function of that bar alone, so the leaving term is formed again rather
// than kept in a ring); 0 when the volume sum is not positive, and a full
// window of zero volumes restarts both sums and gives 0
__device__ __forceinline__ double cmf_flow(const Row& r, int t) {
  const double h = r.in(0, t), l = r.in(1, t), c = r.in(2, t), rg = h - l;
  return rg > 0.0 ? (c - l - (h - c)) / rg * r.in(3, t) : 0.0;
}
__device__ void f_cmf(const Row& r, int p) {
  if (r.n < p) return;
  double sm = 0.0, sv = 0.0;
  int run = 0;
  for (int t = 0; t < p; ++t) {
    const double v = r.in(3, t);
    sm += cmf_flow(r, t);
    sv += v;
    run = fabs(v) <= 0.0 ? run + 1 : 0;
  }
  r.put(0, p - 1, sv > 0.0 ? sm / sv : 0.0);
  for (int t = p; t < r.n; ++t) {
    const double v = r.in(3, t);
    sm -= cmf_flow(r, t - p);
    sv -= r.in(3, t - p);
    sm += cmf_flow(r, t);
    sv += v;
    run = fabs(v) <= 0.0 ? run + 1 : 0;
    double o;
    if (run >= p) {
      run = p;
      sm = 0.0;
      sv = 0.0;
      o = 0.0;
    } else {
      o = sv > 0.0 ? sm / sv : 0.0;
    }
    r.put(0, t, o);
  }
}

// ta_EMV: the running p-bar mean of (mid - previous mid) / box ratio, box =
// volume / divisor / range (a zero range or box contributes 0)
__device__ __forceinline__ double emv_term(const Row& r, int t, double div) {
  const double h = r.in(0, t), l = r.in(1, t), rg = h - l;
  if (rg == 0.0) return 0.0;
  const double box = r.in(2, t) / div / rg;
  if (box == 0.0) return 0.0;
  return ((h + l) / 2.0 - (r.in(0, t - 1) + r.in(1, t - 1)) / 2.0) / box;
}
__device__ void f_emv(const Row& r, int p, double div) {
  if (r.n <= p) return;
  double s = 0.0;
  for (int t = 1; t < p; ++t) s += emv_term(r, t, div);
  for (int t = p; t < r.n; ++t) {
    s += emv_term(r, t, div);
    const double tot = s;
    s -= emv_term(r, t - p + 1, div);
    r.put(0, t, tot / (double)p);
  }
}

// ta_FOSC: 100 (x - forecast) / x, the forecast fma(m, p, b) of the
// least-squares line over the p bars before t (abscissa: bars back); the
// sums move one bar at a time and are formed afresh every 32 p bars or when
// the leaving term exceeds 100 times the window's absolute sum
__device__ void f_fosc(const Row& r, int p) {
  if (r.n <= p) return;
  const double sumX = (double)p * (p - 1) * 0.5;
  const double sumXSqr = (double)p * (p - 1) * (2 * p - 1) / 6.0;
  const double div = sumX * sumX - p * sumXSqr;
  double sxy = 0, sy = 0, sa = 0;
  for (int back = p - 1; back >= 0; --back) {
    const double v = r.in(0, p - 1 - back);
    sy += v;
    sxy += (double)back * v;
    sa += fabs(v);
  }
  int left = 32 * p;
  for (int t = p; t < r.n; ++t) {
    if (t > p) {
      const double old = r.in(0, t - p - 1), wt = (double)p * old, nw = r.in(0, t - 1);
      sxy = sxy + sy - wt;
      sy = sy - old + nw;
      sa = sa - fabs(old) + fabs(nw);
      if (--left <= 0 || fabs(wt) > 100.0 * sa) {
        left = 32 * p;
        sy = 0;
        sxy = 0;
        sa = 0;
        double w = (double)(p - 1);
        for (int j = t - p; j < t; ++j) {
          const double v = r.in(0, j);
          sy += v;
          sxy += w * v;
          sa += fabs(v);
          w -= 1.0;
        }
      }
    }
    const double m = (p * sxy - sumX * sy) / div;
    const double b = (sy - m * sumX) / (double)p;
    const double c = r.in(0, t);
    r.put(0, t, c != 0.0 ? 100.0 * (c - fma(m, (double)p, b)) / c : 0.0);
  }
}

// ta_ERI: bull = high - EMA(close), bear = low - EMA(close)
__device__ void f_eri(const Row& r, int p) {
  if (r.n < p) return;
  if (p == 1) {
    for (int t = 0; t < r.n; ++t) {
      r.put(0, t, r.in(0, t) - r.in(2, t));
      r.put(1, t, r.in(1, t) - r.in(2, t));
    }
    return;
  }
  double b, k;
  ema_bk(p, b, k);
  double s = 0.0;
  for (int t = 0; t < p; ++t) s += r.in(2, t);
  double ma = s / p;
  r.put(0, p - 1, r.in(0, p - 1) - ma);
  r.put(1, p - 1, r.in(1, p - 1) - ma);
  for (int t = p; t < r.n; ++t) {
    ma = fma(b, ma, k * r.in(2, t));
    r.put(0, t, r.in(0, t) - ma);
    r.put(1, t, r.in(1, t) - ma);
  }
}

// ta_PERCENTRANK: 100 x the share of the p bars before t below x[t]
__device__ __forceinline__ double pt_percentrank(const Row& r, int t, int p) {
  const double c = r.in(0, t);
  int below = 0;
  for (int s = t - p; s < t; ++s) below += r.in(0, s) < c;
  return (double)below / (double)p * 100.0;
}

// Wilder's swing index of bar t (inputs open, high, low, close) against bar
// t - 1, limit move lim (ta_SI, ta_ASI)
__device__ __forceinline__ double swing_index(const Row& r, int t, double lim) {
  const double pc = r.in(3, t - 1), pb = pc - r.in(0, t - 1), body = r.in(3, t) - r.in(0, t);
  double num = r.in(3, t) - pc;
  num += 0.5 * body;
  num += 0.25 * pb;
  const double up = fabs(r.in(1, t) - pc), dn = fabs(r.in(2, t) - pc);
  const double rg = fabs(r.in(1, t) - r.in(2, t));
  const double kk = up > dn ? up : dn;
  const double a = up - 0.5 * dn, c = dn - 0.5 * up;
  double den = a > c ? a : c;
  den = den > rg ? den : rg;
  den += 0.25 * fabs(pb);
  return den == 0.0 ? 0.0 : 50.0 * (num / den) * (kk / lim);
}
__device__ void f_asi(const Row& r, double lim) {
  double acc = 0.0;
  r.put(0, 0, acc);
  for (int t = 1; t < r.n; ++t) {
    acc += swing_index(r, t, lim);
    r.put(0, t, acc);
  }
}

// ta_VORTEX: running sums over the p bars ending at t of the true range
// against the previous close, |high - previous low| and |low - previous
// high|; VI+ and VI- are the last two over the first (0 when it is not
// positive).  A full window of zero true ranges zeroes the true-range sum
// only.
__device__ __forceinline__ double vortex_tr(const Row& r, int i) {
  const double h = r.in(0, i), l = r.in(1, i), pc = r.in(2, i - 1);
  double tr = h - l;
  const double a = fabs(pc - h);
  if (a > tr) tr = a;
  const double b = fabs(pc - l);
  if (b > tr) tr = b;
  return tr;
}
__device__ void f_vortex(const Row& r, int p) {
  if (r.n <= p) return;
  double st = 0.0, sp = 0.0, sm = 0.0;
  int run = 0;
  for (int i = 1; i < p; ++i) {
    const double tr = vortex_tr(r, i);
    st += tr;
    sp += fabs(r.in(0, i) - r.in(1, i - 1));
    sm += fabs(r.in(1, i) - r.in(0, i - 1));
    run = tr <= 0.0 ? run + 1 : 0;
  }
  for (int t = p; t < r.n; ++t) {
    const double tr = vortex_tr(r, t);
    st += tr;
    sp += fabs(r.in(0, t) - r.in(1, t - 1));
    sm += fabs(r.in(1, t) - r.in(0, t - 1));
    run = tr <= 0.0 ? run + 1 : 0;
    if (run >= p) {
      run = p;
      st = 0.0;
    }
    const double ct = st, cp = sp, cm = sm;
    const int o = t - p + 1;
    st -= vortex_tr(r, o);
    sp -= fabs(r.in(0, o) - r.in(1, o - 1));
    sm -= fabs(r.in(1, o) - r.in(0, o - 1));
    r.put(0, t, ct > 0.0 ? cp / ct : 0.0);
    r.put(1, t, ct > 0.0 ? cm / ct : 0.0);
  }
}

// ta_CMOU: running sums of the up and down moves of the p changes ending at
// t (up = (d + |d|) / 2, down = up - d); 100 (up - down) / (up + down), 0
// when that sum is not positive; a full window of unchanged bars restarts
// the sums and gives 0
__device__ void f_cmou(const Row& r, int p) {
  if (r.n <= p) return;
  double us = 0.0, ds = 0.0;
  int run = 0;
  auto add = [&](int t) {
    const double d = r.in(0, t) - r.in(0, t - 1), up = (d + fabs(d)) * 0.5;
    us += up;
    ds += up - d;
    return d;
  };
  auto drop = [&](int t) {
    const double d = r.in(0, t) - r.in(0, t - 1), up = (d + fabs(d)) * 0.5;
    us -= up;
    ds -= up - d;
  };
  for (int t = 1; t <= p; ++t) run = fabs(add(t)) <= 0.0 ? run + 1 : 0;
  {
    const double sum = us + ds;
    r.put(0, p, sum > 0.0 ? 100.0 * (us - ds) / sum : 0.0);
  }
  drop(1);
  for (int t = p + 1; t < r.n; ++t) {
    run = fabs(add(t)) <= 0.0 ? run + 1 : 0;
    double o = 0.0;
    if (run >= p) {
      run = p;
      us = 0.0;
      ds = 0.0;
    } else {
      const double sum = us + ds;
      o = sum > 0.0 ? 100.0 * (us - ds) / sum : 0.0;
    }
    r.put(0, t, o);
    drop(t - p + 1);
  }
}

template <int F>
__device__ void run_row(const Row& r, int p0, int p1, double d0) {
  if constexpr (F == kCumsum) f_cumsum(r);
  else if constexpr (F == kPvt) f_pvt(r);
  else if constexpr (F == kRma) f_rma(r, p0);
  else if constexpr (F == kZlema) f_zlema(r, p0);
  else if constexpr (F == kEfi) f_efi(r, p0);
  else if constexpr (F == kVwma) f_vwma(r, p0);
  else if constexpr (F == kEr) f_er(r, p0);
  else if constexpr (F == kDpo) f_dpo(r, p0);
  else if constexpr (F == kAdr) f_mean_diff(r, 0, 1, p0);
  else if constexpr (F == kAo) f_ao(r, p0, p1);
  else if constexpr (F == kQstick) f_mean_diff(r, 1, 0, p0);
  else if constexpr (F == kDonchian) f_donchian(r, p0);
  else if constexpr (F == kWad) f_wad(r);
  else if constexpr (F == kMcgd) f_mcgd(r, p0);
  else if constexpr (F == kRvol) f_rvol(r, p0);
  else if constexpr (F == kCmf) f_cmf(r, p0);
  else if constexpr (F == kEmv) f_emv(r, p0, d0);
  else if constexpr (F == kFosc) f_fosc(r, p0);
  else if constexpr (F == kEri) f_eri(r, p0);
  else if constexpr (F == kAsiWilder) f_asi(r, d0);
  else if constexpr (F == kVortex) f_vortex(r, p0);
  else f_cmou(r, p0);
}

// The functions whose every output is a function of a window of inputs
// alone (no state carried between bars): one work-item per output.
// DONCHIAN joins them for NaN-free rows: its kept extremes are then the
// window's maximum and minimum, which the first strict improvement from the
// oldest bar finds as TA-Lib's re-search does; a row with a NaN after its
// prefix runs the sequential routine afterwards (k_rows_fix), whose
// NaN-dependent re-search points the direct form cannot know.
constexpr bool pointwise(int F) {
  return F == kVhf || F == kChop || F == kMarketfi || F == kPercentrank || F == kSi ||
         F == kDonchian;
}

__device__ __forceinline__ void pt_donchian(const Row& r, int t, int p, double& hi, double& lo) {
  hi = r.in(0, t - p + 1);
  lo = r.in(1, t - p + 1);
  for (int s = t - p + 2; s <= t; ++s) {
    const double h = r.in(0, s), l = r.in(1, s);
    if (h > hi) hi = h;
    if (l < lo) lo = l;
  }
}

// Output at bar t of the row view r (NaN before the first output).
template <int F>
__device__ __forceinline__ double pt_eval(const Row& r, int t, int p0, double d0) {
  const double nan_ = __longlong_as_double(0x7ff8000000000000ll);
  if constexpr (F == kVhf) return t >= p0 ? pt_vhf(r, t, p0) : nan_;
  else if constexpr (F == kChop) return t >= p0 ? pt_chop(r, t, p0) : nan_;
  else if constexpr (F == kMarketfi) return pt_marketfi(r, t);
  else if constexpr (F == kPercentrank) return t >= p0 ? pt_percentrank(r, t, p0) : nan_;
  else return t >= 1 ? swing_index(r, t, d0) : nan_;
}

// Leading bars of a row where any input is NaN (at most len - 1), found by
// the whole block a chunk at a time.
template <int NIN>
__device__ int block_prefix(const float* const* in, size_t base, int len, int* s_f) {
  if (threadIdx.x == 0) *s_f = len - 1;
  __syncthreads();
  for (int c = 0; c < len - 1; c += (int)blockDim.x) {
    const int i = c + (int)threadIdx.x;
    if (i < len - 1) {
      bool any = false;
      for (int j = 0; j < NIN; ++j) any |= __builtin_isnan(in[j][base + i]);
      if (!any) atomicMin(s_f, i);
    }
    __syncthreads();
    if (*s_f < c + (int)blockDim.x) break;
  }
  const int f = *s_f;
  __syncthreads();
  return f;
}

// Pointwise functions: a block of 256 outputs of one row.
template <int F>
__global__ __launch_bounds__(256) void k_rows_pt(RowArgs a, int tilesPerRow) {
  constexpr RowSpec S = kSpec[F];
  __shared__ int s_f;
  const long long row = blockIdx.x / tilesPerRow;
  const int t = (int)(blockIdx.x % tilesPerRow) * 256 + (int)threadIdx.x;
  const size_t base = (size_t)row * (size_t)a.len;
  const int f = block_prefix<S.nin>(a.in, base, a.len, &s_f);
  if (t >= a.len) return;
  Row r{};
  for (int j = 0; j < 4; ++j) r.x[j] = j < S.nin ? a.in[j] + base + f : nullptr;
  r.n = a.len - f;
  const float nanf_ = __int_as_float(0x7fc00000);
  if constexpr (F == kDonchian) {
    float u = nanf_, m = nanf_, l = nanf_;
    if (t - f >= a.p0 - 1) {
      double hi, lo;
      pt_donchian(r, t - f, a.p0, hi, lo);
      u = (float)hi;
      m = (float)((hi + lo) / 2.0);
      l = (float)lo;
    }
    a.out[0][base + t] = u;
    a.out[1][base + t] = m;
    a.out[2][base + t] = l;
  } else {
    a.out[0][base + t] = t < f ? nanf_ : (float)pt_eval<F>(r, t - f, a.p0, a.d0);
  }
}

// Rows with a NaN after the prefix: the sequential routine over the row, one
// block of four waves per row (the block finds the prefix and looks for a
// NaN, four loads in flight per lane; lane 0 runs the routine, overwriting
// every output from the first one on).
template <int F>
__global__ __launch_bounds__(256) void k_rows_fix(RowArgs a) {
  constexpr RowSpec S = kSpec[F];
  constexpr int CAP = 256;
  __shared__ float s_buf[3 * CAP];
  __shared__ int s_f;
  __shared__ int s_nan;
  const long long row = blockIdx.x;
  const size_t base = (size_t)row * (size_t)a.len;
  const int f = block_prefix<S.nin>(a.in, base, a.len, &s_f);
  if (threadIdx.x == 0) s_nan = 0;
  __syncthreads();
  bool any = false;
  for (int t = f + (int)threadIdx.x; t < a.len; t += 4 * 256) {
    float v[S.nin][4];
#pragma unroll
    for (int j = 0; j < S.nin; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int u = t + q * 256;
        v[j][q] = u < a.len ? a.in[j][base + u] : 0.0f;
      }
#pragma unroll
    for (int j = 0; j < S.nin; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) any |= __builtin_isnan(v[j][q]);
  }
  if (any) s_nan = 1;
  __syncthreads();
  if (!s_nan || threadIdx.x != 0) return;
  Row r{};
  for (int j = 0; j < 4; ++j) r.x[j] = j < S.nin ? a.in[j] + base + f : nullptr;
  for (int k = 0; k < 3; ++k) r.y[k] = k < S.nout ? a.out[k] + base + f : nullptr;
  r.n = a.len - f;
  r.buf = (lds_float*)s_buf;
  r.cap = CAP;
  run_row<F>(r, a.p0, a.p1, a.d0);
  r.flush_all();
}

// SPREAD: one row per wave (the wave fills the row with NaN and finds its
// prefix, then lane 0 runs the routine); else one row per lane.
template <int F, bool SPREAD>
__global__ __launch_bounds__(64) void k_rows(RowArgs a) {
  constexpr RowSpec S = kSpec[F];
  constexpr int CAP = SPREAD ? 256 : 16;
  __shared__ float s_buf[(SPREAD ? 1 : 64) * 3 * CAP];
  __shared__ int s_f;
  const float nanf_ = __int_as_float(0x7fc00000);
  const long long row = SPREAD ? (long long)blockIdx.x : (long long)blockIdx.x * 64 + threadIdx.x;
  const size_t base = (size_t)row * (size_t)a.len;
  int f = 0;
  if constexpr (SPREAD) {
    for (int k = 0; k < S.nout; ++k)
      for (int t = (int)threadIdx.x; t < a.len; t += 64) a.out[k][base + t] = nanf_;
    f = block_prefix<S.nin>(a.in, base, a.len, &s_f);  // also orders the fill
    if (threadIdx.x != 0) return;
  } else {
    if (row >= a.rows) return;
    for (int k = 0; k < S.nout; ++k)
      for (int t = 0; t < a.len; ++t) a.out[k][base + t] = nanf_;
    for (; f < a.len - 1; ++f) {
      bool any = false;
      for (int j = 0; j < S.nin; ++j) any |= __builtin_isnan(a.in[j][base + f]);
      if (!any) break;
    }
  }
  Row r{};
  for (int j = 0; j < 4; ++j) r.x[j] = j < S.nin ? a.in[j] + base + f : nullptr;
  for (int k = 0; k < 3; ++k) r.y[k] = k < S.nout ? a.out[k] + base + f : nullptr;
  r.n = a.len - f;
  r.buf = (lds_float*)(s_buf + (SPREAD ? 0 : threadIdx.x * 3 * CAP));
  r.cap = CAP;
  run_row<F>(r, a.p0, a.p1, a.d0);
  r.flush_all();
}

template <int F>
ctStatus_t run_rows(const char* fn, const float* const* in, float* const* out, int len,
                    int rows, int p0, int p1, double d0, hipStream_t s) {
  constexpr RowSpec S = kSpec[F];
  if (len <= 0 || rows <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen and numSeries must be positive",
                fn);
  RowArgs a{};
  for (int j = 0; j < S.nin; ++j) {
    if (!in[j]) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
    a.in[j] = in[j];
  }
  for (int k = 0; k < S.nout; ++k) {
    if (!out[k]) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
    a.out[k] = out[k];
  }
  const int par[2] = {p0, p1};
  for (int k = 0; k < S.npar; ++k)
    if (par[k] < S.lo[k] || par[k] > S.hi[k])
      return fail(CT_STATUS_INVALID_ARGUMENT, "%s: period out of range", fn);
  if (S.dlo <= S.dhi && !(d0 >= S.dlo && d0 <= S.dhi))
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: parameter out of range", fn);
  a.d0 = d0;
  a.len = len;
  a.rows = rows;
  a.p0 = p0;
  a.p1 = p1;
  if constexpr (pointwise(F)) {
    const int tiles = (len + 255) / 256;
    if ((long long)tiles * rows >= (1ll << 31))
      return fail(CT_STATUS_UNSUPPORTED, "%s: batch too large", fn);
    hipLaunchKernelGGL(k_rows_pt<F>, dim3((unsigned)(tiles * rows)), dim3(256), 0, s, a, tiles);
    TCA_HIP(hipGetLastError());
    if constexpr (F == kDonchian) {
      hipLaunchKernelGGL(k_rows_fix<F>, dim3((unsigned)rows), dim3(256), 0, s, a);
      TCA_HIP(hipGetLastError());
    }
    return CT_STATUS_SUCCESS;
  }
  // few long rows: a wave each, so the rows' loads do not share one CU
  if (rows <= 2048 && len >= 4096)
    hipLaunchKernelGGL((k_rows<F, true>), dim3(rows), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_rows<F, false>), dim3((rows + 63) / 64), dim3(64), 0, s, a);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

#define RF_RET(x)                               \
  do {                                          \
    const ctStatus_t r_ = (x);                  \
    if (r_ != CT_STATUS_SUCCESS) return r_;     \
  } while (0)

// Device calls synchronize on the legacy default stream, as the reference's
// backend does (Talib08.cu:143-144).
template <int F>
ctStatus_t dev_rows(const char* fn, const float* const* in, float* const* out, int len, int rows,
                    int p0, int p1, double d0, hipStream_t s) {
  RF_RET(run_rows<F>(fn, in, out, len, rows, p0, p1, d0, s));
  return finish_batch(s);
}

template <int F>
ctStatus_t host_rows(const char* fn, const float* const* hin, float* const* hout, int len,
                     int rows, int p0, int p1, double d0, hipStream_t s) {
  constexpr RowSpec S = kSpec[F];
  if (len <= 0 || rows <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: size/seriesLen and numSeries must be positive",
                fn);
  for (int j = 0; j < S.nin; ++j)
    if (!hin[j]) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  for (int k = 0; k < S.nout; ++k)
    if (!hout[k]) return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer", fn);
  const size_t n = (size_t)len * (size_t)rows;
  HostRun hr(s);
  const float* d_in[4] = {nullptr, nullptr, nullptr, nullptr};
  float* d_out[3] = {nullptr, nullptr, nullptr};
  for (int j = 0; j < S.nin; ++j) RF_RET(hr.in(hin[j], n, &d_in[j]));
  for (int k = 0; k < S.nout; ++k) RF_RET(hr.out(hout[k], n, &d_out[k]));
  RF_RET(run_rows<F>(fn, d_in, d_out, len, rows, p0, p1, d0, s));
  return hr.finish();
}

}  // namespace
}  // namespace tca

using namespace tca;

// Parameter lists by shape: inputs, outputs and periods.
#define RF_IN1 const float* i0
#define RF_IN2 const float* i0, const float* i1
#define RF_IN3 const float* i0, const float* i1, const float* i2
#define RF_IN4 const float* i0, const float* i1, const float* i2, const float* i3
#define RF_INA1 {i0, nullptr, nullptr, nullptr}
#define RF_INA2 {i0, i1, nullptr, nullptr}
#define RF_INA3 {i0, i1, i2, nullptr}
#define RF_INA4 {i0, i1, i2, i3}
#define RF_OUT1 float* o0
#define RF_OUT2 float* o0, float* o1
#define RF_OUT3 float* o0, float* o1, float* o2
#define RF_OUTA1 {o0, nullptr, nullptr}
#define RF_OUTA2 {o0, o1, nullptr}
#define RF_OUTA3 {o0, o1, o2}
#define RF_P0
#define RF_P1 , int p0
#define RF_P2 , int p0, int p1
#define RF_PD , double d0
#define RF_P1D , int p0, double d0
#define RF_PV0 0, 0, 0.0
#define RF_PV1 p0, 0, 0.0
#define RF_PV2 p0, p1, 0.0
#define RF_PVD 0, 0, d0
#define RF_PV1D p0, 0, d0

#define TCA_ROWFN(NAME, F, NI, NO, NP)                                                           \
  extern "C" {                                                                                   \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_d(RF_IN##NI, RF_OUT##NO, int n RF_P##NP, hipStream_t s) {  \
    const float* in[4] = RF_INA##NI;                                                             \
    float* out[3] = RF_OUTA##NO;                                                                 \
    return guarded([&] { return dev_rows<F>("ct_" #NAME "_d", in, out, n, 1, RF_PV##NP, s); });  \
  }                                                                                              \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch_d(RF_IN##NI, RF_OUT##NO, int len,                    \
                                              int rows RF_P##NP, int, hipStream_t s) {           \
    const float* in[4] = RF_INA##NI;                                                             \
    float* out[3] = RF_OUTA##NO;                                                                 \
    return guarded(                                                                              \
        [&] { return dev_rows<F>("ct_" #NAME "_batch_d", in, out, len, rows, RF_PV##NP, s); });  \
  }                                                                                              \
  CTAPI_EXPORT ctStatus_t ct_##NAME(RF_IN##NI, RF_OUT##NO, int n RF_P##NP, hipStream_t s) {      \
    const float* in[4] = RF_INA##NI;                                                             \
    float* out[3] = RF_OUTA##NO;                                                                 \
    return guarded([&] { return host_rows<F>("ct_" #NAME, in, out, n, 1, RF_PV##NP, s); });      \
  }                                                                                              \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch(RF_IN##NI, RF_OUT##NO, int len, int rows RF_P##NP,   \
                                            int, hipStream_t s) {                                \
    const float* in[4] = RF_INA##NI;                                                             \
    float* out[3] = RF_OUTA##NO;                                                                 \
    return guarded(                                                                              \
        [&] { return host_rows<F>("ct_" #NAME "_batch", in, out, len, rows, RF_PV##NP, s); });   \
  }                                                                                              \
  }

TCA_ROWFN(cumsum, kCumsum, 1, 1, 0)
TCA_ROWFN(pvt, kPvt, 2, 1, 0)
TCA_ROWFN(rma, kRma, 1, 1, 1)
TCA_ROWFN(zlema, kZlema, 1, 1, 1)
TCA_ROWFN(efi, kEfi, 2, 1, 1)
TCA_ROWFN(vwma, kVwma, 2, 1, 1)
TCA_ROWFN(er, kEr, 1, 1, 1)
TCA_ROWFN(vhf, kVhf, 1, 1, 1)
TCA_ROWFN(dpo, kDpo, 1, 1, 1)
TCA_ROWFN(chop, kChop, 3, 1, 1)
TCA_ROWFN(adr, kAdr, 2, 1, 1)
TCA_ROWFN(ao, kAo, 2, 1, 2)
TCA_ROWFN(qstick, kQstick, 2, 1, 1)
TCA_ROWFN(donchian, kDonchian, 2, 3, 1)
TCA_ROWFN(marketfi, kMarketfi, 3, 1, 0)
TCA_ROWFN(wad, kWad, 3, 1, 0)
TCA_ROWFN(mcgd, kMcgd, 1, 1, 1)
TCA_ROWFN(rvol, kRvol, 1, 1, 1)
TCA_ROWFN(cmf, kCmf, 4, 1, 1)
TCA_ROWFN(emv, kEmv, 3, 1, 1D)
TCA_ROWFN(fosc, kFosc, 1, 1, 1)
TCA_ROWFN(eri, kEri, 3, 2, 1)
TCA_ROWFN(percentrank, kPercentrank, 1, 1, 1)
TCA_ROWFN(si, kSi, 4, 1, D)
TCA_ROWFN(asi_wilder, kAsiWilder, 4, 1, D)
TCA_ROWFN(vortex, kVortex, 3, 2, 1)
TCA_ROWFN(cmou, kCmou, 1, 1, 1)
