/*
 * Device API: memory helpers and ct_*_d calls on GPU-resident float32 series.
 * Same names, argument order and semantics as the reference's
 * include/tacuda_device.h (line numbers cited); hipStream_t replaces
 * cudaStream_t.  Single-series calls only enqueue work on `stream` and never
 * synchronize (reference docs/BATCH_API.md:324-336).
 */
#ifndef TACUDA_AMD_TACUDA_DEVICE_H
#define TACUDA_AMD_TACUDA_DEVICE_H

#include "tacuda.h"

#ifdef __cplusplus
extern "C" {
#endif

/* :81-124 memory management */
CTAPI_EXPORT float* ct_device_alloc(int count);
CTAPI_EXPORT void* ct_device_alloc_ex(int count, ctDataType_t storageType);
CTAPI_EXPORT void ct_device_free(float* d_ptr);
CTAPI_EXPORT void ct_device_free_ex(void* d_ptr);
CTAPI_EXPORT ctStatus_t ct_device_upload(float* d_dst, const float* h_src, int count, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_device_download(float* h_dst, const float* d_src, int count, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_device_sync(hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :130-171 moving averages */
CTAPI_EXPORT ctStatus_t ct_sma_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ema_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_wma_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_dema_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_tema_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_kama_d(const float* d_input, float* d_output, int size, int period, int fastPeriod, int slowPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_t3_d(const float* d_input, float* d_output, int size, int period, float vFactor, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :176-196 momentum */
CTAPI_EXPORT ctStatus_t ct_rsi_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :201-219 volatility, HLC input (`percentile` is the reference's unused
 * ATR argument, kept for ABI compatibility) */
CTAPI_EXPORT ctStatus_t ct_atr_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, float percentile, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_natr_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_trange_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :224-241 trend */
CTAPI_EXPORT ctStatus_t ct_adx_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_dx_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_plus_di_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_minus_di_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* generated device calls (tacuda_device.h:438+) */
CTAPI_EXPORT ctStatus_t ct_plus_dm_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_minus_dm_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :246-271 oscillators */
CTAPI_EXPORT ctStatus_t ct_macd_d(const float* d_input, float* d_macd, float* d_signal, float* d_hist, int size, int fastPeriod, int slowPeriod, int signalPeriod, int maType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_stochastic_d(const float* d_high, const float* d_low, const float* d_close, float* d_slowK, float* d_slowD, int size, int kPeriod, int dPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_willr_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_cci_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :280 bands */
CTAPI_EXPORT ctStatus_t ct_bbands_d(const float* d_input, float* d_upper, float* d_middle, float* d_lower, int size, int period, float nbDevUp, float nbDevDn, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :313-326 statistical */
CTAPI_EXPORT ctStatus_t ct_stddev_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_var_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_sum_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_max_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_min_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :361-366 aroon (down before up, as in the reference) */
CTAPI_EXPORT ctStatus_t ct_aroon_d(const float* d_high, const float* d_low, float* d_aroonDown, float* d_aroonUp, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_aroonosc_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_TACUDA_DEVICE_H */
