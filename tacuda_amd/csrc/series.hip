// Kernel template, launch policy and the C ABI for single-series (ct_*_d),
// host (ct_*) and dense batch (ct_*_batch[_d]) calls of the indicator core.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/tacuda_amd.h"
#include "../../include/tacuda_device.h"
#include "families.hpp"
#include "host.hpp"
#include "series.hpp"

namespace tca {

// One wave per segment; W waves per block share nothing but the block.
template <class F, int K, int H, int NIN, int W>
__global__ __launch_bounds__(W* kWave) void k_seg(Geometry g, Ws ws, int useLB,
                                                   typename F::P prm, IO io) {
  using T = Tile<K, H>;
  __shared__ __align__(16) float s_in[W][NIN * T::PLEN];
  __shared__ __align__(16) float s_out[W][T::POUT];
  const int w = threadIdx.x / kWave;
  Seg c;
  c.tree = useLB != 0;
  c.epoch = useLB ? read_epoch(ws) : 0u;
  const int segIdx = blockIdx.x * W + w;
  if (resolve(g, T::SEG, segIdx, c)) {
    c.ws = ws;
    F::template run<T>(c, s_in[w], s_out[w], prm, io);
  }
  if (useLB) {
    __syncthreads();
    if (threadIdx.x == 0) finish_block(ws, gridDim.x);
  }
}

// Look-back area of a launch: levels x (segments + groups) slots of
// 2R+1 doubles.
inline size_t tree_bytes(int levels, int maxR, int totalSegs) {
  return (size_t)levels * ((size_t)totalSegs + totalSegs / kGroup) * (2 * maxR + 1) * 8;
}

template <class F, int K, int H, int NIN, int W>
static ctStatus_t launch_geom(Geometry g, int maxSegsPerRow, hipStream_t s, typename F::P prm,
                              IO io) {
  if (g.totalSegs <= 0) return CT_STATUS_SUCCESS;
  const bool useLB = F::kScan && maxSegsPerRow > 1;
  Ws ws{nullptr, nullptr, 2 * (F::kMaxR > 0 ? F::kMaxR : 1) + 1};
  if (useLB) {
    ctStatus_t st = workspace_for(s, tree_bytes(F::kLevels, F::kMaxR, g.totalSegs), &ws);
    if (st != CT_STATUS_SUCCESS) return st;
  }
  const int blocks = div_up(g.totalSegs, W);
  hipLaunchKernelGGL((k_seg<F, K, H, NIN, W>), dim3(blocks), dim3(W * kWave), 0, s, g, ws,
                     useLB ? 1 : 0, prm, io);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

// Short rows (<= 256 bars) use 4 bars per lane so a row fills one wave;
// longer rows use the family's long tile and chain segments by look-back.
template <class F, int KL, int HL, int KS, int HS, int NIN, int W>
static ctStatus_t launch_dense(int len, int rows, hipStream_t s, typename F::P prm, IO io) {
  if (t_ragged) {
    // ragged batch: the short tile when every series fits one wave
    Geometry g;
    int maxSegs = 0;
    if (t_ragged->maxLen <= KS * kWave) {
      if (!ragged_geometry(t_ragged, KS * kWave, &g, &maxSegs))
        return fail(CT_STATUS_UNSUPPORTED, "ragged: no table for this tile");
      return launch_geom<F, KS, HS, NIN, W>(g, maxSegs, s, prm, io);
    }
    if (!ragged_geometry(t_ragged, KL * kWave, &g, &maxSegs))
      return fail(CT_STATUS_UNSUPPORTED, "ragged: no table for this tile");
    return launch_geom<F, KL, HL, NIN, W>(g, maxSegs, s, prm, io);
  }
  if (len <= KS * kWave) {
    Geometry g = dense_geometry(len, rows, KS * kWave);
    return launch_geom<F, KS, HS, NIN, W>(g, g.segsPerRow, s, prm, io);
  }
  Geometry g = dense_geometry(len, rows, KL * kWave);
  return launch_geom<F, KL, HL, NIN, W>(g, g.segsPerRow, s, prm, io);
}

// ----------------------------------------------------- per-family policy --
#ifndef TCA_KL
#define TCA_KL 8
#endif
#define TCA_POLICY(FAM, KL, HL, KS, HS, NIN, W)                                            \
  ctStatus_t run_##FAM(int len, int rows, hipStream_t s, const F##FAM::P& p, IO io) {     \
    return launch_dense<F##FAM, KL, HL, KS, HS, NIN, W>(len, rows, s, p, io);              \
  }

TCA_POLICY(Ema, TCA_KL, 16, 4, 4, 1, 4)
TCA_POLICY(Dema, TCA_KL, 16, 4, 4, 1, 4)
TCA_POLICY(Tema, TCA_KL, 16, 4, 4, 1, 4)
TCA_POLICY(T3, 8, 8, 4, 4, 1, 4)
TCA_POLICY(Macd, TCA_KL, 16, 4, 4, 1, 4)
TCA_POLICY(Rsi, TCA_KL, 16, 4, 4, 1, 4)
TCA_POLICY(Kama, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Atr, TCA_KL, 16, 4, 4, 3, 2)
TCA_POLICY(Natr, TCA_KL, 16, 4, 4, 3, 2)
TCA_POLICY(Sma, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Sum, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Wma, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Var, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Stddev, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Bbands, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Max, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Min, TCA_KL, 64, 4, 64, 1, 4)
TCA_POLICY(Willr, TCA_KL, 64, 4, 64, 2, 2)
TCA_POLICY(Cci, 8, 64, 4, 64, 3, 2)
TCA_POLICY(Aroon, TCA_KL, 64, 4, 64, 2, 2)
TCA_POLICY(AroonOsc, TCA_KL, 64, 4, 64, 2, 2)
TCA_POLICY(Trange, TCA_KL, 16, 4, 4, 3, 2)
TCA_POLICY(Stoch, 8, 64, 4, 64, 3, 2)
TCA_POLICY(Obv, TCA_KL, 16, 4, 4, 2, 4)
TCA_POLICY(Ad, TCA_KL, 16, 4, 4, 4, 2)
#undef TCA_POLICY

#define TCA_DIR_POLICY(NAME, KIND)                                                          \
  ctStatus_t run_##NAME(int len, int rows, hipStream_t s, int p, IO io) {                   \
    typename FDir<KIND>::P pr{p};                                                           \
    return launch_dense<FDir<KIND>, TCA_KL, TCA_KL, 4, 4, 3, 2>(len, rows, s, pr, io);              \
  }
TCA_DIR_POLICY(PlusDm, kPDM)
TCA_DIR_POLICY(MinusDm, kMDM)
TCA_DIR_POLICY(PlusDi, kPDI)
TCA_DIR_POLICY(MinusDi, kMDI)
TCA_DIR_POLICY(Dx, kDX)
TCA_DIR_POLICY(Adx, kADX)
#undef TCA_DIR_POLICY

}  // namespace tca

// =================================================================== C ABI ==
using namespace tca;

namespace {

inline IO io1(const float* a, float* o0, float* o1 = nullptr, float* o2 = nullptr) {
  IO io{};
  io.in[0] = a;
  io.out[0] = o0;
  io.out[1] = o1;
  io.out[2] = o2;
  return io;
}
inline IO io3(const float* h, const float* l, const float* c, float* o0, float* o1 = nullptr) {
  IO io{};
  io.in[0] = h;
  io.in[1] = l;
  io.in[2] = c;
  io.out[0] = o0;
  io.out[1] = o1;
  return io;
}

ctStatus_t bad(const char* fn, const char* why) {
  return fail(CT_STATUS_INVALID_ARGUMENT, "%s: %s", fn, why);
}

// Shape checks shared by single-series and batch entry points.
ctStatus_t check_shape(const char* fn, int len, int rows) {
  if (len <= 0) return bad(fn, "size/seriesLen must be positive");
  if (rows <= 0) return bad(fn, "numSeries must be positive");
  if ((long long)len * rows > (long long)INT32_MAX * 16)
    return bad(fn, "matrix too large");
  return CT_STATUS_SUCCESS;
}

}  // namespace

// Core device-pointer dispatch used by _d, _batch_d and the host wrappers.
namespace tca {

#define CHECK_PTR(p, fn) \
  if (!(p)) return bad(fn, "null pointer")
#define CHECK_RET(x)                            \
  do {                                          \
    ctStatus_t st_ = (x);                       \
    if (st_ != CT_STATUS_SUCCESS) return st_;   \
  } while (0)

// Dense batches of long rows of EMA, KAMA, T3 and SMA run through the fused
// scan (scan_serves): the same oracle-following results from one pass that
// reads each bar once, with one look-back record per four segments.
ctStatus_t via_scan(const float* x, int len, int rows, hipStream_t s, ctxScanOutputs so,
                    ctxScanParams sp) {
  return scan_run(x, len, rows, &sp, &so, s);
}
// Dense batches of long rows of RSI, MACD and ATR likewise run through the
// fused pass's two streaming passes (pipe_family), with every other pack
// period at a valid placeholder.
ctxPackParams pipe_params() {
  ctxPackParams p{};
  p.sma_period = p.ema_period = p.wma_period = p.dema_period = p.tema_period = 14;
  p.rsi_period = p.atr_period = p.adx_period = p.willr_period = p.cci_period = 14;
  p.macd_fast = 12;
  p.macd_slow = 26;
  p.macd_signal = 9;
  p.bb_period = 20;
  p.bb_up = p.bb_dn = 2.0f;
  p.stoch_k = 14;
  p.stoch_d = 3;
  return p;
}

ctxScanParams scan_params() {
  ctxScanParams sp{};
  sp.ema_period = 2;
  sp.kama_period = 2;
  sp.kama_fast = 2;
  sp.kama_slow = 30;
  sp.t3_period = 2;
  sp.t3_vfactor = 0.7f;
  sp.sma_period = 1;
  return sp;
}

ctStatus_t dev_sma(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "sma"); CHECK_PTR(y, "sma");
  CHECK_RET(check_shape("sma", len, rows));
  if (p <= 0) return bad("sma", "invalid period");
  if (!t_ragged && scan_serves(len, rows, p - 1)) {
    ctxScanParams sp = scan_params();
    sp.sma_period = p;
    ctxScanOutputs so{};
    so.sma = y;
    return via_scan(x, len, rows, s, so, sp);
  }
  {
    ctxPackParams pp = pipe_params();
    pp.sma_period = p;
    ctxPackOutputs po{};
    po.sma = y;
    bool served = false;
    CHECK_RET(win_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Sma(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_sum(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "sum"); CHECK_PTR(y, "sum");
  CHECK_RET(check_shape("sum", len, rows));
  if (p <= 0) return bad("sum", "invalid period");
  return run_Sum(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_wma(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "wma"); CHECK_PTR(y, "wma");
  CHECK_RET(check_shape("wma", len, rows));
  if (p <= 0) return bad("wma", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.wma_period = p;
    ctxPackOutputs po{};
    po.wma = y;
    bool served = false;
    CHECK_RET(win_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Wma(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_ema(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "ema"); CHECK_PTR(y, "ema");
  CHECK_RET(check_shape("ema", len, rows));
  if (p <= 0) return bad("ema", "invalid period");
  if (!t_ragged && scan_serves(len, rows, 0)) {
    ctxScanParams sp = scan_params();
    sp.ema_period = p;
    ctxScanOutputs so{};
    so.ema = y;
    return via_scan(x, len, rows, s, so, sp);
  }
  {
    ctxPackParams pp = pipe_params();
    pp.ema_period = p;
    ctxPackOutputs po{};
    po.ema = y;
    bool served = false;
    CHECK_RET(pipe_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Ema(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_dema(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "dema"); CHECK_PTR(y, "dema");
  CHECK_RET(check_shape("dema", len, rows));
  if (p <= 0) return bad("dema", "invalid period");
  return run_Dema(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_tema(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "tema"); CHECK_PTR(y, "tema");
  CHECK_RET(check_shape("tema", len, rows));
  if (p <= 0) return bad("tema", "invalid period");
  return run_Tema(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_t3(const float* x, float* y, int len, int rows, int p, float v, hipStream_t s) {
  CHECK_PTR(x, "t3"); CHECK_PTR(y, "t3");
  CHECK_RET(check_shape("t3", len, rows));
  if (p < 2) return bad("t3", "invalid period (T3 needs period >= 2)");
  if (!t_ragged && scan_serves(len, rows, 0)) {
    ctxScanParams sp = scan_params();
    sp.t3_period = p;
    sp.t3_vfactor = v;
    ctxScanOutputs so{};
    so.t3 = y;
    return via_scan(x, len, rows, s, so, sp);
  }
  return run_T3(len, rows, s, {p, (double)v}, io1(x, y));
}
ctStatus_t dev_kama(const float* x, float* y, int len, int rows, int p, int fast, int slow,
                    hipStream_t s) {
  CHECK_PTR(x, "kama"); CHECK_PTR(y, "kama");
  CHECK_RET(check_shape("kama", len, rows));
  if (p <= 0 || fast <= 0 || slow <= 0) return bad("kama", "invalid period");
  if (!t_ragged && scan_serves(len, rows, p + 1)) {
    ctxScanParams sp = scan_params();
    sp.kama_period = p;
    sp.kama_fast = fast;
    sp.kama_slow = slow;
    ctxScanOutputs so{};
    so.kama = y;
    return via_scan(x, len, rows, s, so, sp);
  }
  return run_Kama(len, rows, s, {p, fast, slow}, io1(x, y));
}
ctStatus_t dev_rsi(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "rsi"); CHECK_PTR(y, "rsi");
  CHECK_RET(check_shape("rsi", len, rows));
  if (p <= 0) return bad("rsi", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.rsi_period = p;
    ctxPackOutputs po{};
    po.rsi = y;
    bool served = false;
    CHECK_RET(pipe_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Rsi(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_macd(const float* x, float* m, float* sg, float* h, int len, int rows, int fast,
                    int slow, int sig, int maType, hipStream_t s) {
  CHECK_PTR(x, "macd"); CHECK_PTR(m, "macd"); CHECK_PTR(sg, "macd"); CHECK_PTR(h, "macd");
  CHECK_RET(check_shape("macd", len, rows));
  if (fast <= 0 || slow <= 0 || sig <= 0) return bad("macd", "invalid periods");
  // other moving-average types: TA-Lib MACDEXT with that type on all three
  // lines (reference MACDEXT.cu:164-235)
  if (maType != CT_MA_EMA && maType != CT_MA_DEFAULT)
    return dev_macdext(x, m, sg, h, len, rows, fast, maType, slow, maType, sig, maType, s);
  if (slow < fast) std::swap(slow, fast);
  {
    ctxPackParams pp = pipe_params();
    pp.macd_fast = fast;
    pp.macd_slow = slow;
    pp.macd_signal = sig;
    ctxPackOutputs po{};
    po.macd = m;
    po.macd_signal = sg;
    po.macd_hist = h;
    bool served = false;
    CHECK_RET(pipe_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Macd(len, rows, s, {fast, slow, sig}, io1(x, m, sg, h));
}
ctStatus_t dev_bbands(const float* x, float* u, float* m, float* l, int len, int rows, int p,
                      float up, float dn, hipStream_t s) {
  CHECK_PTR(x, "bbands"); CHECK_PTR(u, "bbands"); CHECK_PTR(m, "bbands"); CHECK_PTR(l, "bbands");
  CHECK_RET(check_shape("bbands", len, rows));
  if (p <= 0) return bad("bbands", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.bb_period = p;
    pp.bb_up = up;
    pp.bb_dn = dn;
    ctxPackOutputs po{};
    po.bb_upper = u;
    po.bb_middle = m;
    po.bb_lower = l;
    bool served = false;
    CHECK_RET(win_family(pp, po, x, x, x, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Bbands(len, rows, s, {p, (double)up, (double)dn}, io1(x, u, m, l));
}
ctStatus_t dev_var(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "var"); CHECK_PTR(y, "var");
  CHECK_RET(check_shape("var", len, rows));
  if (p <= 0) return bad("var", "invalid period");
  return run_Var(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_stddev(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "stddev"); CHECK_PTR(y, "stddev");
  CHECK_RET(check_shape("stddev", len, rows));
  if (p <= 0) return bad("stddev", "invalid period");
  return run_Stddev(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_max(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "max"); CHECK_PTR(y, "max");
  CHECK_RET(check_shape("max", len, rows));
  if (p <= 0) return bad("max", "invalid period");
  return run_Max(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_min(const float* x, float* y, int len, int rows, int p, hipStream_t s) {
  CHECK_PTR(x, "min"); CHECK_PTR(y, "min");
  CHECK_RET(check_shape("min", len, rows));
  if (p <= 0) return bad("min", "invalid period");
  return run_Min(len, rows, s, {p}, io1(x, y));
}
ctStatus_t dev_atr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                   int p, hipStream_t s) {
  CHECK_PTR(h, "atr"); CHECK_PTR(l, "atr"); CHECK_PTR(c, "atr"); CHECK_PTR(y, "atr");
  CHECK_RET(check_shape("atr", len, rows));
  if (p <= 0) return bad("atr", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.atr_period = p;
    ctxPackOutputs po{};
    po.atr = y;
    bool served = false;
    CHECK_RET(pipe_family(pp, po, h, l, c, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Atr(len, rows, s, {p}, io3(h, l, c, y));
}
ctStatus_t dev_natr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                    int p, hipStream_t s) {
  CHECK_PTR(h, "natr"); CHECK_PTR(l, "natr"); CHECK_PTR(c, "natr"); CHECK_PTR(y, "natr");
  CHECK_RET(check_shape("natr", len, rows));
  if (p <= 0) return bad("natr", "invalid period");
  return run_Natr(len, rows, s, {p}, io3(h, l, c, y));
}
ctStatus_t dev_obv(const float* c, const float* v, float* y, int len, int rows, hipStream_t s) {
  CHECK_PTR(c, "obv"); CHECK_PTR(v, "obv"); CHECK_PTR(y, "obv");
  CHECK_RET(check_shape("obv", len, rows));
  IO io{};
  io.in[0] = c;
  io.in[1] = v;
  io.out[0] = y;
  return run_Obv(len, rows, s, {0}, io);
}
ctStatus_t dev_ad(const float* h, const float* l, const float* c, const float* v, float* y,
                  int len, int rows, hipStream_t s) {
  CHECK_PTR(h, "ad"); CHECK_PTR(l, "ad"); CHECK_PTR(c, "ad"); CHECK_PTR(v, "ad"); CHECK_PTR(y, "ad");
  CHECK_RET(check_shape("ad", len, rows));
  IO io{};
  io.in[0] = h;
  io.in[1] = l;
  io.in[2] = c;
  io.in[3] = v;
  io.out[0] = y;
  return run_Ad(len, rows, s, {0}, io);
}
ctStatus_t dev_trange(const float* h, const float* l, const float* c, float* y, int len, int rows,
                      hipStream_t s) {
  CHECK_PTR(h, "trange"); CHECK_PTR(l, "trange"); CHECK_PTR(c, "trange"); CHECK_PTR(y, "trange");
  CHECK_RET(check_shape("trange", len, rows));
  return run_Trange(len, rows, s, {0}, io3(h, l, c, y));
}
ctStatus_t dev_stoch(const float* h, const float* l, const float* c, float* k, float* d, int len,
                     int rows, int kp, int dp, hipStream_t s) {
  CHECK_PTR(h, "stochastic"); CHECK_PTR(l, "stochastic"); CHECK_PTR(c, "stochastic");
  CHECK_PTR(k, "stochastic"); CHECK_PTR(d, "stochastic");
  CHECK_RET(check_shape("stochastic", len, rows));
  if (kp <= 0 || dp <= 0) return bad("stochastic", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.stoch_k = kp;
    pp.stoch_d = dp;
    ctxPackOutputs po{};
    po.stoch_k = k;
    po.stoch_d = d;
    bool served = false;
    CHECK_RET(win_family(pp, po, h, l, c, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Stoch(len, rows, s, {kp, dp}, io3(h, l, c, k, d));
}
ctStatus_t dev_willr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                     int p, hipStream_t s) {
  CHECK_PTR(h, "willr"); CHECK_PTR(l, "willr"); CHECK_PTR(c, "willr"); CHECK_PTR(y, "willr");
  CHECK_RET(check_shape("willr", len, rows));
  if (p <= 0) return bad("willr", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.willr_period = p;
    ctxPackOutputs po{};
    po.willr = y;
    bool served = false;
    CHECK_RET(win_family(pp, po, h, l, c, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Willr(len, rows, s, {p}, io3(h, l, c, y));
}
ctStatus_t dev_cci(const float* h, const float* l, const float* c, float* y, int len, int rows,
                   int p, hipStream_t s) {
  CHECK_PTR(h, "cci"); CHECK_PTR(l, "cci"); CHECK_PTR(c, "cci"); CHECK_PTR(y, "cci");
  CHECK_RET(check_shape("cci", len, rows));
  if (p <= 0) return bad("cci", "invalid period");
  {
    ctxPackParams pp = pipe_params();
    pp.cci_period = p;
    ctxPackOutputs po{};
    po.cci = y;
    bool served = false;
    CHECK_RET(win_family(pp, po, h, l, c, len, rows, s, &served));
    if (served) return CT_STATUS_SUCCESS;
  }
  return run_Cci(len, rows, s, {p}, io3(h, l, c, y));
}
ctStatus_t dev_aroon(const float* h, const float* l, float* dn, float* up, int len, int rows,
                     int p, hipStream_t s) {
  CHECK_PTR(h, "aroon"); CHECK_PTR(l, "aroon"); CHECK_PTR(dn, "aroon"); CHECK_PTR(up, "aroon");
  CHECK_RET(check_shape("aroon", len, rows));
  if (p <= 0) return bad("aroon", "invalid period");
  return run_Aroon(len, rows, s, {p}, io3(h, l, nullptr, dn, up));
}
ctStatus_t dev_aroonosc(const float* h, const float* l, float* y, int len, int rows, int p,
                        hipStream_t s) {
  CHECK_PTR(h, "aroonosc"); CHECK_PTR(l, "aroonosc"); CHECK_PTR(y, "aroonosc");
  CHECK_RET(check_shape("aroonosc", len, rows));
  if (p <= 0) return bad("aroonosc", "invalid period");
  return run_AroonOsc(len, rows, s, {p}, io3(h, l, nullptr, y));
}
#define DIR_DEV(NAME, RUN, NEEDC)                                                            \
  ctStatus_t dev_##NAME(const float* h, const float* l, const float* c, float* y, int len,  \
                        int rows, int p, hipStream_t s) {                                   \
    CHECK_PTR(h, #NAME); CHECK_PTR(l, #NAME); CHECK_PTR(y, #NAME);                          \
    if (NEEDC) CHECK_PTR(c, #NAME);                                                         \
    CHECK_RET(check_shape(#NAME, len, rows));                                               \
    if (p <= 0) return bad(#NAME, "invalid period");                                        \
    return RUN(len, rows, s, p, io3(h, l, c, y));                                           \
  }
DIR_DEV(plus_dm, run_PlusDm, false)
DIR_DEV(minus_dm, run_MinusDm, false)
DIR_DEV(plus_di, run_PlusDi, true)
DIR_DEV(minus_di, run_MinusDi, true)
DIR_DEV(dx, run_Dx, true)
DIR_DEV(adx, run_Adx, true)
#undef DIR_DEV

}  // namespace tca

#define R_(x) CHECK_RET(x)

extern "C" {

// ---------------------------------------------------------- device (_d) --
CTAPI_EXPORT ctStatus_t ct_sma_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_sma(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_sum_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_sum(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_wma_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_wma(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_ema_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_ema(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_dema_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_dema(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_tema_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_tema(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_t3_d(const float* x, float* y, int n, int p, float v, hipStream_t s) { return guarded([&] { return dev_t3(x, y, n, 1, p, v, s); }); }
CTAPI_EXPORT ctStatus_t ct_kama_d(const float* x, float* y, int n, int p, int fast, int slow, hipStream_t s) { return guarded([&] { return dev_kama(x, y, n, 1, p, fast, slow, s); }); }
CTAPI_EXPORT ctStatus_t ct_rsi_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_rsi(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_macd_d(const float* x, float* m, float* sg, float* h, int n, int fast, int slow, int sig, int maType, hipStream_t s) { return guarded([&] { return dev_macd(x, m, sg, h, n, 1, fast, slow, sig, maType, s); }); }
CTAPI_EXPORT ctStatus_t ct_bbands_d(const float* x, float* u, float* m, float* l, int n, int p, float up, float dn, hipStream_t s) { return guarded([&] { return dev_bbands(x, u, m, l, n, 1, p, up, dn, s); }); }
CTAPI_EXPORT ctStatus_t ct_var_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_var(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_stddev_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_stddev(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_max_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_max(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_min_d(const float* x, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_min(x, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_atr_d(const float* h, const float* l, const float* c, float* y, int n, int p, float, hipStream_t s) { return guarded([&] { return dev_atr(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_natr_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_natr(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_trange_d(const float* h, const float* l, const float* c, float* y, int n, hipStream_t s) { return guarded([&] { return dev_trange(h, l, c, y, n, 1, s); }); }
CTAPI_EXPORT ctStatus_t ct_stochastic_d(const float* h, const float* l, const float* c, float* k, float* d, int n, int kp, int dp, hipStream_t s) { return guarded([&] { return dev_stoch(h, l, c, k, d, n, 1, kp, dp, s); }); }
CTAPI_EXPORT ctStatus_t ct_willr_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_willr(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_cci_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_cci(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_aroon_d(const float* h, const float* l, float* dn, float* up, int n, int p, hipStream_t s) { return guarded([&] { return dev_aroon(h, l, dn, up, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_aroonosc_d(const float* h, const float* l, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_aroonosc(h, l, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_plus_dm_d(const float* h, const float* l, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_plus_dm(h, l, nullptr, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_minus_dm_d(const float* h, const float* l, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_minus_dm(h, l, nullptr, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_plus_di_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_plus_di(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_minus_di_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_minus_di(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_dx_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_dx(h, l, c, y, n, 1, p, s); }); }
CTAPI_EXPORT ctStatus_t ct_adx_d(const float* h, const float* l, const float* c, float* y, int n, int p, hipStream_t s) { return guarded([&] { return dev_adx(h, l, c, y, n, 1, p, s); }); }

// ------------------------------------------------------------ batch (_d) --
#define B1D(NAME)                                                                                     \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch_d(const float* x, float* y, int len, int rows, int p,     \
                                              int, hipStream_t s) {                                  \
    return guarded([&]() -> ctStatus_t {                                                              \
      R_(dev_##NAME(x, y, len, rows, p, s));                                                          \
      return finish_batch(s);                                                                         \
    });                                                                                               \
  }
B1D(sma) B1D(ema) B1D(rsi) B1D(wma) B1D(dema) B1D(tema) B1D(stddev) B1D(var) B1D(sum) B1D(max) B1D(min)
#undef B1D

#define BHLCD(NAME)                                                                                   \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_batch_d(const float* h, const float* l, const float* c,         \
                                              float* y, int len, int rows, int p, int,               \
                                              hipStream_t s) {                                       \
    return guarded([&]() -> ctStatus_t {                                                              \
      R_(dev_##NAME(h, l, c, y, len, rows, p, s));                                                    \
      return finish_batch(s);                                                                         \
    });                                                                                               \
  }
BHLCD(natr) BHLCD(adx) BHLCD(dx) BHLCD(plus_di) BHLCD(minus_di) BHLCD(willr) BHLCD(cci)
#undef BHLCD

CTAPI_EXPORT ctStatus_t ct_atr_batch_d(const float* h, const float* l, const float* c, float* y, int len, int rows, int p, float, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_atr(h, l, c, y, len, rows, p, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_macd_batch_d(const float* x, float* m, float* sg, float* h, int len, int rows, int fast, int slow, int sig, int maType, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_macd(x, m, sg, h, len, rows, fast, slow, sig, maType, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_bbands_batch_d(const float* x, float* u, float* m, float* l, int len, int rows, int p, float up, float dn, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_bbands(x, u, m, l, len, rows, p, up, dn, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_stochastic_batch_d(const float* h, const float* l, const float* c, float* k, float* d, int len, int rows, int kp, int dp, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_stoch(h, l, c, k, d, len, rows, kp, dp, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_kama_batch_d(const float* x, float* y, int len, int rows, int p, int fast, int slow, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_kama(x, y, len, rows, p, fast, slow, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_t3_batch_d(const float* x, float* y, int len, int rows, int p, float v, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_t3(x, y, len, rows, p, v, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_obv_d(const float* c, const float* v, float* y, int n, hipStream_t s) { return guarded([&] { return dev_obv(c, v, y, n, 1, s); }); }
CTAPI_EXPORT ctStatus_t ct_obv_batch_d(const float* c, const float* v, float* y, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_obv(c, v, y, len, rows, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_ad_d(const float* h, const float* l, const float* c, const float* v, float* y, int n, hipStream_t s) { return guarded([&] { return dev_ad(h, l, c, v, y, n, 1, s); }); }
CTAPI_EXPORT ctStatus_t ct_ad_batch_d(const float* h, const float* l, const float* c, const float* v, float* y, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_ad(h, l, c, v, y, len, rows, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_trange_batch_d(const float* h, const float* l, const float* c, float* y, int len, int rows, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_trange(h, l, c, y, len, rows, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_aroonosc_batch_d(const float* h, const float* l, float* y, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_aroonosc(h, l, y, len, rows, p, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_aroon_batch_d(const float* h, const float* l, float* dn, float* up, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_aroon(h, l, dn, up, len, rows, p, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_plus_dm_batch_d(const float* h, const float* l, float* y, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_plus_dm(h, l, nullptr, y, len, rows, p, s)); return finish_batch(s); });
}
CTAPI_EXPORT ctStatus_t ct_minus_dm_batch_d(const float* h, const float* l, float* y, int len, int rows, int p, int, hipStream_t s) {
  return guarded([&]() -> ctStatus_t { R_(dev_minus_dm(h, l, nullptr, y, len, rows, p, s)); return finish_batch(s); });
}

}  // extern "C"

// ------------------------------------------------------------- ragged --
namespace {
template <class Fn>
ctStatus_t ragged_call(ctxRagged layout, const char* fn, Fn&& body) {
  return guarded([&]() -> ctStatus_t {
    const Ragged* r = ragged_of(layout);
    if (!r) return bad(fn, "null layout");
    RaggedScope scope(r);
    return body(r->maxLen, r->rows);
  });
}
}  // namespace

extern "C" {
#define RAG1(name)                                                                            \
  CTAPI_EXPORT ctStatus_t ctx_##name##_ragged_d(const float* x, float* y, ctxRagged lay, int p, \
                                                hipStream_t s) {                              \
    return ragged_call(lay, #name, [&](int n, int rows) { return dev_##name(x, y, n, rows, p, s); }); \
  }
RAG1(sma)
RAG1(ema)
RAG1(wma)
RAG1(dema)
RAG1(tema)
RAG1(rsi)
RAG1(var)
RAG1(stddev)
RAG1(sum)
RAG1(max)
RAG1(min)
#undef RAG1
#define RAGHLC(name)                                                                         \
  CTAPI_EXPORT ctStatus_t ctx_##name##_ragged_d(const float* h, const float* l, const float* c, \
                                                float* y, ctxRagged lay, int p, hipStream_t s) { \
    return ragged_call(lay, #name,                                                          \
                       [&](int n, int rows) { return dev_##name(h, l, c, y, n, rows, p, s); }); \
  }
RAGHLC(atr)
RAGHLC(natr)
RAGHLC(adx)
RAGHLC(dx)
RAGHLC(plus_di)
RAGHLC(minus_di)
RAGHLC(willr)
RAGHLC(cci)
#undef RAGHLC
CTAPI_EXPORT ctStatus_t ctx_t3_ragged_d(const float* x, float* y, ctxRagged lay, int p, float v,
                                        hipStream_t s) {
  return ragged_call(lay, "t3", [&](int n, int rows) { return dev_t3(x, y, n, rows, p, v, s); });
}
CTAPI_EXPORT ctStatus_t ctx_kama_ragged_d(const float* x, float* y, ctxRagged lay, int p, int fast,
                                          int slow, hipStream_t s) {
  return ragged_call(lay, "kama",
                     [&](int n, int rows) { return dev_kama(x, y, n, rows, p, fast, slow, s); });
}
CTAPI_EXPORT ctStatus_t ctx_macd_ragged_d(const float* x, float* m, float* sg, float* h,
                                          ctxRagged lay, int fast, int slow, int sig,
                                          hipStream_t s) {
  return ragged_call(lay, "macd", [&](int n, int rows) {
    return dev_macd(x, m, sg, h, n, rows, fast, slow, sig, CT_MA_EMA, s);
  });
}
CTAPI_EXPORT ctStatus_t ctx_bbands_ragged_d(const float* x, float* u, float* m, float* l,
                                            ctxRagged lay, int p, float up, float dn,
                                            hipStream_t s) {
  return ragged_call(lay, "bbands",
                     [&](int n, int rows) { return dev_bbands(x, u, m, l, n, rows, p, up, dn, s); });
}
CTAPI_EXPORT ctStatus_t ctx_stochastic_ragged_d(const float* h, const float* l, const float* c,
                                                float* k, float* d, ctxRagged lay, int kp, int dp,
                                                hipStream_t s) {
  return ragged_call(lay, "stochastic",
                     [&](int n, int rows) { return dev_stoch(h, l, c, k, d, n, rows, kp, dp, s); });
}
CTAPI_EXPORT ctStatus_t ctx_trange_ragged_d(const float* h, const float* l, const float* c,
                                            float* y, ctxRagged lay, hipStream_t s) {
  return ragged_call(lay, "trange",
                     [&](int n, int rows) { return dev_trange(h, l, c, y, n, rows, s); });
}
}  // extern "C"
