/*
 * Extensions of this MI355X build that have no counterpart in the reference
 * ABI: workspace reservation before graph capture, the seeded OHLCV
 * generator used by tests and benchmarks, ragged batches, the fused
 * multi-indicator plan, int32 candlestick outputs and the adjoint kernels of
 * the differentiable indicators.
 */
#ifndef TACUDA_AMD_EXT_H
#define TACUDA_AMD_EXT_H

#include <stdint.h>

#include "tacuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reserve `bytes` of the per-stream look-back workspace so a later call can
 * be captured into a hipGraph without allocating (an eager call of the same
 * shape on the stream reserves it too).  A graph captured on a stream keeps
 * the workspace it was captured with: a later, larger call on that stream
 * moves the stream to a new workspace and leaves the old one allocated until
 * ct_device_buffer_pool_cleanup, after which such graphs must not be
 * replayed.  Replays of one captured graph, and eager calls on its capture
 * stream, share that workspace and must not overlap: replay such a graph on
 * one stream at a time (ctx_plan_launch orders the replays of a plan
 * itself). */
CTAPI_EXPORT ctStatus_t ctx_workspace_reserve(hipStream_t stream, size_t bytes);

/* How the host calls (ct_<name>, ct_<name>_batch) move host arrays (no
 * counterpart in the reference, which page-locks host matrices of 1 MB or
 * more): 0 = by size (arrays of 1 MiB or more through a pair of pinned 4 MiB
 * staging chunks whose transfers overlap the host-side copies, smaller ones
 * copied directly; the default), 1 = always direct copies of pageable
 * memory, 2 = always pinned staging. */
CTAPI_EXPORT ctStatus_t ctx_host_staging(int mode);

/* Kernel timing (no counterpart in the reference; for benchmarks and
 * tuning).  While enabled, the fused-pass, long-scan and candle launches
 * record a pair of HIP events around each kernel on the caller's stream;
 * launches on a stream that is being captured record nothing.
 * ctx_kernel_times waits for the recorded launches, writes up to maxN of
 * them (kernel names separated by '\n' into `names`, elapsed milliseconds
 * into `ms`), sets *count, and clears the list. */
CTAPI_EXPORT ctStatus_t ctx_kernel_timing(int enable);
CTAPI_EXPORT ctStatus_t ctx_kernel_times(char* names, int namesLen, float* ms, int maxN,
                                         int* count);

/* Seeded synthetic OHLCV, float32, series-major [numSeries x n]:
 * close_0 = 100, close_t = close_{t-1} * exp(0.01*z_r); open_t = close_{t-1};
 * high = max(o,c)*(1+|0.005*z_h|), low = min(o,c)*(1-|0.005*z_l|),
 * volume = exp(10 + z_v).  Series s draws from the stream
 * (seed, firstSeries + s) of a splitmix64/xoshiro256** generator with
 * Box-Muller normals in double, draw order z_r, z_h, z_l, z_v per bar.
 * Any output pointer may be null.  Runs on the host over `threads` threads
 * (0 = hardware concurrency). */
CTAPI_EXPORT ctStatus_t ctx_synth_ohlcv(unsigned long long seed, int firstSeries, int numSeries,
                                        int n, float* open, float* high, float* low,
                                        float* close, float* volume, int threads);

/* ---------------------------------------------------------------------------
 * Fused multi-indicator pack: one launch computes any subset of SMA, EMA,
 * WMA, DEMA, TEMA, RSI, MACD, BBANDS(SMA), ATR, ADX, STOCH (SMA/SMA), WILLR
 * and CCI over every row, reading high/low/close once.  A null output skips
 * that indicator; EMA, DEMA and TEMA of one pack share their period.  Each
 * output has the layout and NaN warm-up of its ct_*_batch_d counterpart and
 * agrees with it within float32 rounding of the FP64 recurrences.
 * ------------------------------------------------------------------------- */
typedef struct ctxPackParams {
  int sma_period, ema_period, wma_period, dema_period, tema_period, rsi_period;
  int macd_fast, macd_slow, macd_signal;
  int bb_period;
  float bb_up, bb_dn;
  int atr_period, adx_period;
  int stoch_k, stoch_d;
  int willr_period, cci_period;
} ctxPackParams;

typedef struct ctxPackOutputs {
  float *sma, *ema, *wma, *dema, *tema, *rsi;
  float *macd, *macd_signal, *macd_hist;
  float *bb_upper, *bb_middle, *bb_lower;
  float *atr, *adx, *stoch_k, *stoch_d, *willr, *cci;
} ctxPackOutputs;

/* Dense series-major rows (numSeries x seriesLen), device pointers.  A single
 * series is numSeries = 1.  Stream 0 blocks until done; other streams only
 * enqueue and can be captured. */
CTAPI_EXPORT ctStatus_t ctx_pack_batch_d(const float* d_high, const float* d_low,
                                         const float* d_close, int seriesLen, int numSeries,
                                         const ctxPackParams* params,
                                         const ctxPackOutputs* outputs, hipStream_t stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same pass and outputs as ctx_pack_batch_d, with
 * its look-back scratch in a fixed array of 3 x 8 MiB per GPU held until
 * the process ends, instead of the per-stream workspace cache: the call
 * allocates nothing.  A pass whose scratch exceeds 8 MiB per part (about
 * 3 x 10^7 bars of long rows) returns CT_STATUS_INVALID_ARGUMENT (outputs
 * undefined).  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (several streams or threads) race. */
CTAPI_EXPORT ctStatus_t ctx_pack_batch_noalloc_d(const float* d_high, const float* d_low,
                                                 const float* d_close, int seriesLen,
                                                 int numSeries, const ctxPackParams* params,
                                                 const ctxPackOutputs* outputs,
                                                 hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Ragged batches: series r occupies elements [offsets[r], offsets[r+1]) of
 * concatenated input and output arrays (offsets are int64, host memory,
 * numSeries+1 entries, non-decreasing; empty series are allowed).  A layout
 * is built once (it uploads its tables synchronously) and reused by any
 * number of calls, including calls captured into a hipGraph.  Each series
 * gets exactly the result of its single-series ct_*_d call.
 * ------------------------------------------------------------------------- */
typedef struct ctxRagged_st* ctxRagged;
CTAPI_EXPORT ctStatus_t ctx_ragged_create(const long long* offsets, int numSeries,
                                          ctxRagged* layout);
CTAPI_EXPORT ctStatus_t ctx_ragged_destroy(ctxRagged layout);

CTAPI_EXPORT ctStatus_t ctx_pack_ragged_d(const float* d_high, const float* d_low,
                                          const float* d_close, ctxRagged layout,
                                          const ctxPackParams* params,
                                          const ctxPackOutputs* outputs, hipStream_t stream);

/* Fused long-series scan (no counterpart in the reference, whose families run
 * as separate calls ct_ema_batch_d, ct_kama_batch_d, ct_t3_batch_d,
 * ct_sma_batch_d at include/tacuda.h and src/api_batch.cu): EMA, KAMA, T3
 * and SMA of the same close rows from one read-twice, write-once launch
 * sequence.  Each output equals the family call's; a null output is skipped.
 * Window look-backs (SMA period - 1, KAMA period + 1) up to 1024 bars. */
typedef struct ctxScanParams {
  int ema_period;
  int kama_period, kama_fast, kama_slow;
  int t3_period;
  float t3_vfactor;
  int sma_period;
} ctxScanParams;

typedef struct ctxScanOutputs {
  float *ema, *kama, *t3, *sma;
} ctxScanOutputs;

CTAPI_EXPORT ctStatus_t ctx_scan_batch_d(const float* d_close, int seriesLen, int numSeries,
                                         const ctxScanParams* params,
                                         const ctxScanOutputs* outputs, hipStream_t stream);

/* Single-family ragged calls; arguments as the family's ct_*_d call. */
#define CTX_RAGGED1(name) \
  CTAPI_EXPORT ctStatus_t ctx_##name##_ragged_d(const float* in, float* out, ctxRagged layout, \
                                                int period, hipStream_t stream);
CTX_RAGGED1(sma)
CTX_RAGGED1(ema)
CTX_RAGGED1(wma)
CTX_RAGGED1(dema)
CTX_RAGGED1(tema)
CTX_RAGGED1(rsi)
CTX_RAGGED1(var)
CTX_RAGGED1(stddev)
CTX_RAGGED1(sum)
CTX_RAGGED1(max)
CTX_RAGGED1(min)
#undef CTX_RAGGED1
#define CTX_RAGGED_HLC(name)                                                              \
  CTAPI_EXPORT ctStatus_t ctx_##name##_ragged_d(const float* high, const float* low,     \
                                                const float* close, float* out,           \
                                                ctxRagged layout, int period, hipStream_t stream);
CTX_RAGGED_HLC(atr)
CTX_RAGGED_HLC(natr)
CTX_RAGGED_HLC(adx)
CTX_RAGGED_HLC(dx)
CTX_RAGGED_HLC(plus_di)
CTX_RAGGED_HLC(minus_di)
CTX_RAGGED_HLC(willr)
CTX_RAGGED_HLC(cci)
#undef CTX_RAGGED_HLC
CTAPI_EXPORT ctStatus_t ctx_t3_ragged_d(const float* in, float* out, ctxRagged layout,
                                        int period, float vFactor, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_kama_ragged_d(const float* in, float* out, ctxRagged layout,
                                          int period, int fastPeriod, int slowPeriod,
                                          hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_macd_ragged_d(const float* in, float* macd, float* signal,
                                          float* hist, ctxRagged layout, int fastPeriod,
                                          int slowPeriod, int signalPeriod, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_bbands_ragged_d(const float* in, float* upper, float* middle,
                                            float* lower, ctxRagged layout, int period,
                                            float nbDevUp, float nbDevDn, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_stochastic_ragged_d(const float* high, const float* low,
                                                const float* close, float* k, float* d,
                                                ctxRagged layout, int kPeriod, int dPeriod,
                                                hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_trange_ragged_d(const float* high, const float* low,
                                            const float* close, float* out, ctxRagged layout,
                                            hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Persistent tuned plans: one fused pack pass over fixed device buffers
 * (dense seriesLen x numSeries when `layout` is NULL, else the ragged
 * layout), captured into a hipGraph.  With `tune` set, creation times every
 * tile configuration (bars per lane of the scan, segments per workgroup) of
 * each recurrence kernel on these buffers, then the whole pass as three
 * launches with the best tiles against one launch of 4 or 8 bars per lane,
 * and keeps the fastest; the choice is cached per shape for the process.
 * ctx_plan_info reports the tile of the launch that runs (the close
 * kernel's when there are several).  Creation runs the pass several times
 * (outputs are overwritten).  A plan's launches must be serialized (they
 * share the plan's look-back workspace); stream 0 blocks until the pass is
 * done.
 * ------------------------------------------------------------------------- */
typedef struct ctxPlan_st* ctxPlan;
CTAPI_EXPORT ctStatus_t ctx_pack_plan_create(const float* d_high, const float* d_low,
                                             const float* d_close, int seriesLen, int numSeries,
                                             ctxRagged layout, const ctxPackParams* params,
                                             const ctxPackOutputs* outputs, int tune,
                                             ctxPlan* plan);
CTAPI_EXPORT ctStatus_t ctx_plan_launch(ctxPlan plan, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_plan_info(ctxPlan plan, int* barsPerLane, int* wavesPerGroup,
                                      float* tunedUs, float* defaultUs);
CTAPI_EXPORT ctStatus_t ctx_plan_destroy(ctxPlan plan);

/* ---------------------------------------------------------------------------
 * Differentiable indicators: forward and adjoint kernels used by the torch
 * autograd binding.  Dense series-major rows (numSeries x seriesLen, at most
 * 32768 bars per row), device pointers, enqueued on `stream`.
 *
 * EMA_alpha / RSI_alpha take the smoothing weight alpha as a continuous
 * parameter and keep the TA-Lib seed of `period` samples (EMA: mean of bars
 * 0..p-1 at bar p-1; RSI: (sum of the first p gains/losses)/p at bar p);
 * alpha = 2/(p+1) (EMA) or 1/p (RSI) gives the TA-Lib indicator.  The
 * forward writes a checkpoint of the recurrence state at
 * ctx_alpha_ckpt_per_row() points per row (EMA: that many doubles per row,
 * RSI: twice as many); the adjoint recomputes the rest.
 * Adjoints write the gradient w.r.t. the prices (gx) and, per row, the
 * gradient w.r.t. alpha (galpha, doubles; sum over rows for a shared alpha).
 * MACD (ta_MACD.c, EMA type) and BBANDS (SMA middle) adjoints take the
 * output gradients; their forwards are ct_macd_batch_d / ct_bbands_batch_d.
 * Inputs must be finite.
 * ------------------------------------------------------------------------- */
CTAPI_EXPORT int ctx_alpha_ckpt_per_row(void);
CTAPI_EXPORT ctStatus_t ctx_ema_alpha_fwd_d(const float* x, float* y, double* ckpt,
                                            int seriesLen, int numSeries, int period,
                                            double alpha, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_ema_alpha_bwd_d(const float* x, const float* gy, const double* ckpt,
                                            float* gx, double* galpha, int seriesLen,
                                            int numSeries, int period, double alpha,
                                            hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_fwd_d(const float* x, float* y, double* ckpt,
                                            int seriesLen, int numSeries, int period,
                                            double alpha, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_bwd_d(const float* x, const float* gy, const double* ckpt,
                                            float* gx, double* galpha, int seriesLen,
                                            int numSeries, int period, double alpha,
                                            hipStream_t stream);
/* The same four with alpha read on the device from d_alpha (one double):
 * no host round trip for a learned alpha, so they can be captured into a
 * graph.  d_alpha is not range-checked (a NaN alpha gives NaN outputs). */
CTAPI_EXPORT ctStatus_t ctx_ema_alpha_fwd_dp(const float* x, float* y, double* ckpt,
                                             int seriesLen, int numSeries, int period,
                                             const double* d_alpha, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_ema_alpha_bwd_dp(const float* x, const float* gy, const double* ckpt,
                                             float* gx, double* galpha, int seriesLen,
                                             int numSeries, int period, const double* d_alpha,
                                             hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_fwd_dp(const float* x, float* y, double* ckpt,
                                             int seriesLen, int numSeries, int period,
                                             const double* d_alpha, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_rsi_alpha_bwd_dp(const float* x, const float* gy, const double* ckpt,
                                             float* gx, double* galpha, int seriesLen,
                                             int numSeries, int period, const double* d_alpha,
                                             hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_macd_bwd_d(const float* gMacd, const float* gSignal,
                                       const float* gHist, float* gx, int seriesLen,
                                       int numSeries, int fastPeriod, int slowPeriod,
                                       int signalPeriod, hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_bbands_bwd_d(const float* x, const float* gUpper,
                                         const float* gMiddle, const float* gLower, float* gx,
                                         int seriesLen, int numSeries, int period,
                                         float nbDevUp, float nbDevDn, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Candlestick patterns with int32 signals.  Patterns are numbered 0..60 in
 * the order ctx_cdl_name() reports (the reference's ct_cdl_<name>_d names).
 * ctx_cdl_all_batch_d evaluates all 61 in one launch, reading OHLC once and
 * writing plane k (numSeries x seriesLen int32) at out + k*numSeries*seriesLen.
 * ------------------------------------------------------------------------- */
CTAPI_EXPORT int ctx_cdl_count(void);
CTAPI_EXPORT const char* ctx_cdl_name(int pattern);
CTAPI_EXPORT ctStatus_t ctx_cdl_batch_d(int pattern, const float* d_open, const float* d_high,
                                        const float* d_low, const float* d_close,
                                        int32_t* d_output, int seriesLen, int numSeries,
                                        hipStream_t stream);
CTAPI_EXPORT ctStatus_t ctx_cdl_all_batch_d(const float* d_open, const float* d_high,
                                            const float* d_low, const float* d_close,
                                            int32_t* d_outputs, int seriesLen, int numSeries,
                                            hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_EXT_H */
