/*
 * Streaming (incremental update) API of the MI355X build: the reference's
 * streaming surface (reference include/tacuda_stream.h:42-259) with the same
 * names, argument lists, status codes and warm-up semantics.
 *
 * A stream keeps one indicator's state over one series; appending bars
 * costs O(1) per bar (O(period) for the window families SMA, WMA, BBANDS)
 * and reproduces, bar for bar, the corresponding full-series call on the
 * concatenated history (the same arithmetic in the same order).
 * Single-series streams run on the host in double precision.  Batch streams
 * hold numSeries independent states; their appends run on the host or, with
 * the state resident in HBM, on the GPU (one work-item per series).
 *
 * Handles are opaque tokens (never addresses); a destroyed token is never
 * reused while the library is loaded, and using it afterwards returns
 * CT_STATUS_INVALID_ARGUMENT.  One handle must not be used by two threads at
 * once; distinct handles are independent.
 */
#ifndef TACUDA_STREAM_H
#define TACUDA_STREAM_H

#include "tacuda.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ctStream ctStream_t;

/* lifetime and introspection (tacuda_stream.h:49-82) */
CTAPI_EXPORT ctStatus_t ct_stream_destroy(ctStream_t* stream);
CTAPI_EXPORT ctStatus_t ct_stream_reset(ctStream_t* stream);
CTAPI_EXPORT ctStatus_t ct_stream_lookback(const ctStream_t* stream, int* outLookback);
CTAPI_EXPORT ctStatus_t ct_stream_bars(const ctStream_t* stream, long long* outBars);
CTAPI_EXPORT ctStatus_t ct_stream_outputs(const ctStream_t* stream, int* outCount);

/* creation (tacuda_stream.h:88-114) */
CTAPI_EXPORT ctStatus_t ct_sma_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_ema_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_wma_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_rsi_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_atr_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_natr_stream_create(int period, ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_obv_stream_create(ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_ad_stream_create(ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_macd_stream_create(int fastPeriod, int slowPeriod, int signalPeriod,
                                              ctStream_t** outStream);
CTAPI_EXPORT ctStatus_t ct_bbands_stream_create(int period, float upperMul, float lowerMul,
                                                ctStream_t** outStream);

/* appends of n bars (tacuda_stream.h:116-153) */
CTAPI_EXPORT ctStatus_t ct_sma_stream_append(ctStream_t* stream, const float* close, int n,
                                             float* out);
CTAPI_EXPORT ctStatus_t ct_ema_stream_append(ctStream_t* stream, const float* close, int n,
                                             float* out);
CTAPI_EXPORT ctStatus_t ct_wma_stream_append(ctStream_t* stream, const float* close, int n,
                                             float* out);
CTAPI_EXPORT ctStatus_t ct_rsi_stream_append(ctStream_t* stream, const float* close, int n,
                                             float* out);
CTAPI_EXPORT ctStatus_t ct_atr_stream_append(ctStream_t* stream, const float* high,
                                             const float* low, const float* close, int n,
                                             float* out);
CTAPI_EXPORT ctStatus_t ct_natr_stream_append(ctStream_t* stream, const float* high,
                                              const float* low, const float* close, int n,
                                              float* out);
CTAPI_EXPORT ctStatus_t ct_obv_stream_append(ctStream_t* stream, const float* close,
                                             const float* volume, int n, float* out);
CTAPI_EXPORT ctStatus_t ct_ad_stream_append(ctStream_t* stream, const float* high,
                                            const float* low, const float* close,
                                            const float* volume, int n, float* out);
CTAPI_EXPORT ctStatus_t ct_macd_stream_append(ctStream_t* stream, const float* close, int n,
                                              float* outMacd, float* outSignal, float* outHist);
CTAPI_EXPORT ctStatus_t ct_bbands_stream_append(ctStream_t* stream, const float* close, int n,
                                                float* outUpper, float* outMiddle,
                                                float* outLower);

/* ---------------------------------------------------------------------------
 * Batch (portfolio) streams (tacuda_stream.h:155-259).  Inputs and outputs
 * are series-major host matrices: series s's n bars at [s*n, s*n+n).
 * ------------------------------------------------------------------------- */
typedef struct ctStreamBatch ctStreamBatch_t;

typedef enum ctStreamExec {
  CT_STREAM_EXEC_AUTO = 0, /* GPU from CT_STREAM_AUTO_GPU_SERIES series up */
  CT_STREAM_EXEC_CPU = 1,
  CT_STREAM_EXEC_GPU = 2
} ctStreamExec_t;

#define CT_STREAM_AUTO_GPU_SERIES 2048

CTAPI_EXPORT ctStatus_t ct_stream_batch_destroy(ctStreamBatch_t* batch);
CTAPI_EXPORT ctStatus_t ct_stream_batch_reset(ctStreamBatch_t* batch);
CTAPI_EXPORT ctStatus_t ct_stream_batch_lookback(const ctStreamBatch_t* batch, int* outLookback);
CTAPI_EXPORT ctStatus_t ct_stream_batch_series(const ctStreamBatch_t* batch, int* outSeries);
CTAPI_EXPORT ctStatus_t ct_stream_batch_outputs(const ctStreamBatch_t* batch, int* outCount);
CTAPI_EXPORT ctStatus_t ct_stream_batch_bars(const ctStreamBatch_t* batch, long long* outBars);
CTAPI_EXPORT ctStatus_t ct_stream_batch_set_exec(ctStreamBatch_t* batch, ctStreamExec_t exec);

CTAPI_EXPORT ctStatus_t ct_sma_stream_batch_create(int period, int numSeries,
                                                   ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_ema_stream_batch_create(int period, int numSeries,
                                                   ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_wma_stream_batch_create(int period, int numSeries,
                                                   ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_rsi_stream_batch_create(int period, int numSeries,
                                                   ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_atr_stream_batch_create(int period, int numSeries,
                                                   ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_natr_stream_batch_create(int period, int numSeries,
                                                    ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_obv_stream_batch_create(int numSeries, ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_ad_stream_batch_create(int numSeries, ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_macd_stream_batch_create(int fastPeriod, int slowPeriod,
                                                    int signalPeriod, int numSeries,
                                                    ctStreamBatch_t** outBatch);
CTAPI_EXPORT ctStatus_t ct_bbands_stream_batch_create(int period, float upperMul,
                                                      float lowerMul, int numSeries,
                                                      ctStreamBatch_t** outBatch);

CTAPI_EXPORT ctStatus_t ct_sma_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                   int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_ema_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                   int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_wma_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                   int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_rsi_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                   int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_atr_stream_batch_append(ctStreamBatch_t* batch, const float* highs,
                                                   const float* lows, const float* closes, int n,
                                                   float* outs);
CTAPI_EXPORT ctStatus_t ct_natr_stream_batch_append(ctStreamBatch_t* batch, const float* highs,
                                                    const float* lows, const float* closes,
                                                    int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_obv_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                   const float* volumes, int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_ad_stream_batch_append(ctStreamBatch_t* batch, const float* highs,
                                                  const float* lows, const float* closes,
                                                  const float* volumes, int n, float* outs);
CTAPI_EXPORT ctStatus_t ct_macd_stream_batch_append(ctStreamBatch_t* batch, const float* closes,
                                                    int n, float* outMacd, float* outSignal,
                                                    float* outHist);
CTAPI_EXPORT ctStatus_t ct_bbands_stream_batch_append(ctStreamBatch_t* batch,
                                                      const float* closes, int n,
                                                      float* outUpper, float* outMiddle,
                                                      float* outLower);

/* Extension (no reference counterpart): batch append from device memory on
 * `stream`, the state staying in HBM -- for callers whose bars already live
 * on the GPU.  Inputs the family does not read may be NULL; outputs beyond
 * the family's count are ignored.  Moves the batch to GPU execution. */
CTAPI_EXPORT ctStatus_t ctx_stream_batch_append_d(ctStreamBatch_t* batch, const float* d_high,
                                                  const float* d_low, const float* d_close,
                                                  const float* d_volume, int n, float* d_out0,
                                                  float* d_out1, float* d_out2,
                                                  hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_STREAM_H */
