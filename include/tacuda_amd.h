/*
 * Extensions of this MI355X build that have no counterpart in the reference
 * ABI: workspace reservation before graph capture, the seeded OHLCV
 * generator used by tests and benchmarks, ragged batches, the fused
 * multi-indicator plan, int32 candlestick outputs and the adjoint kernels of
 * the differentiable indicators.
 */
#ifndef TACUDA_AMD_EXT_H
#define TACUDA_AMD_EXT_H

#include "tacuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reserve `bytes` of the per-stream look-back workspace so a later call can
 * be captured into a hipGraph without allocating (an eager call of the same
 * shape on the stream reserves it too). */
CTAPI_EXPORT ctStatus_t ctx_workspace_reserve(hipStream_t stream, size_t bytes);

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

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_EXT_H */
