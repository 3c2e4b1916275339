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
/* :101, :115 -- element size from the storage type (2, 4 or 8 bytes) */
CTAPI_EXPORT ctStatus_t ct_device_upload_ex(void* d_dst, const void* h_src, int count, ctDataType_t storageType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_device_download_ex(void* h_dst, const void* d_src, int count, ctDataType_t storageType, hipStream_t stream CTAPI_STREAM_DEFAULT);
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

/* :468 -- moving average selected by type (see ct_ma in tacuda.h) */
CTAPI_EXPORT ctStatus_t ct_ma_d(const float* d_input, float* d_output, int size, int period, int maType, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :151 -- TA-Lib TRIMA (triangular weights), window semantics for NaN */
CTAPI_EXPORT ctStatus_t ct_trima_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :388 -- TA-Lib MACDEXT with a moving-average type per line (ctMaType_t:
 * SMA, EMA, WMA, DEMA, TEMA, TRIMA, KAMA, T3; MAMA and the types above T3
 * return CT_STATUS_UNSUPPORTED).  Fast and slow periods from 2, signal from
 * 1; reversed periods swap with their types; a leading NaN prefix is
 * skipped.  ct_macd_d with maType other than EMA is MACDEXT with that type
 * on all three lines. */
CTAPI_EXPORT ctStatus_t ct_macdext_d(const float* d_input, float* d_macd, float* d_signal, float* d_hist, int size, int fastPeriod, int fastType, int slowPeriod, int slowType, int signalPeriod, int signalType, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* Typed variants (:132, :138, :167, :197, :244, :283).  Float32 storage with
 * float32 or float64 compute runs the float32 path (its recurrences, window
 * sums and ratios are evaluated in float64 and rounded once per output);
 * half and double storage return CT_STATUS_UNSUPPORTED, as do the pairs the
 * reference rejects. */
CTAPI_EXPORT ctStatus_t ct_sma_ex_d(const void* d_input, void* d_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ema_ex_d(const void* d_input, void* d_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rsi_ex_d(const void* d_input, void* d_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_atr_ex_d(const void* d_high, const void* d_low, const void* d_close, void* d_output, int size, int period, float percentile, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_macd_ex_d(const void* d_input, void* d_macd, void* d_signal, void* d_hist, int size, int fastPeriod, int slowPeriod, int signalPeriod, int maType, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_bbands_ex_d(const void* d_input, void* d_upper, void* d_middle, void* d_lower, int size, int period, float nbDevUp, float nbDevDn, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :313-326 statistical */
CTAPI_EXPORT ctStatus_t ct_stddev_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_var_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_sum_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_max_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_min_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :361-366 aroon (down before up, as in the reference) */
CTAPI_EXPORT ctStatus_t ct_aroon_d(const float* d_high, const float* d_low, float* d_aroonDown, float* d_aroonUp, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_aroonosc_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* The 61 TA-Lib candlestick patterns (tacuda_device.h:438+, tacuda.h:949+):
 * legacy float signals {0, +-80, +-100, +-200}, 0 during each pattern's
 * TA-Lib look-back.  ct_cdl_<name>_batch_d runs numSeries series-major rows
 * (maxConcurrency is accepted and unused); the host forms ct_cdl_<name>
 * (tacuda.h:432+) and ct_cdl_<name>_batch (tacuda.h:1870+) stage host
 * arrays through the host-call buffers and wait for the stream.  int32
 * outputs and the all-61 launch: ctx_cdl_batch_d / ctx_cdl_all_batch_d
 * (tacuda_amd.h). */
#define CT_CDL_DECL(name)                                                                      \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_d(const float* d_open, const float* d_high,         \
                                            const float* d_low, const float* d_close,          \
                                            float* d_output, int size,                         \
                                            hipStream_t stream CTAPI_STREAM_DEFAULT);          \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch_d(const float* d_opens, const float* d_highs,  \
                                                  const float* d_lows, const float* d_closes,  \
                                                  float* d_outputs, int seriesLen,             \
                                                  int numSeries, int maxConcurrency,           \
                                                  hipStream_t stream CTAPI_STREAM_DEFAULT);    \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name(const float* host_open, const float* host_high,       \
                                        const float* host_low, const float* host_close,        \
                                        float* host_output, int size,                          \
                                        hipStream_t stream CTAPI_STREAM_DEFAULT);              \
  CTAPI_EXPORT ctStatus_t ct_cdl_##name##_batch(const float* host_opens,                      \
                                                const float* host_highs,                      \
                                                const float* host_lows,                       \
                                                const float* host_closes,                     \
                                                float* host_outputs, int seriesLen,           \
                                                int numSeries, int maxConcurrency,            \
                                                hipStream_t stream CTAPI_STREAM_DEFAULT);
CT_CDL_DECL(two_crows)
CT_CDL_DECL(three_black_crows)
CT_CDL_DECL(three_inside)
CT_CDL_DECL(three_line_strike)
CT_CDL_DECL(three_outside)
CT_CDL_DECL(three_stars_in_south)
CT_CDL_DECL(three_white_soldiers)
CT_CDL_DECL(abandoned_baby)
CT_CDL_DECL(advance_block)
CT_CDL_DECL(belt_hold)
CT_CDL_DECL(breakaway)
CT_CDL_DECL(closing_marubozu)
CT_CDL_DECL(conceal_baby_swallow)
CT_CDL_DECL(counterattack)
CT_CDL_DECL(dark_cloud_cover)
CT_CDL_DECL(doji)
CT_CDL_DECL(doji_star)
CT_CDL_DECL(dragonfly_doji)
CT_CDL_DECL(engulfing)
CT_CDL_DECL(evening_doji_star)
CT_CDL_DECL(evening_star)
CT_CDL_DECL(gap_side_side_white)
CT_CDL_DECL(gravestone_doji)
CT_CDL_DECL(hammer)
CT_CDL_DECL(hanging_man)
CT_CDL_DECL(harami)
CT_CDL_DECL(harami_cross)
CT_CDL_DECL(high_wave)
CT_CDL_DECL(hikkake)
CT_CDL_DECL(hikkake_mod)
CT_CDL_DECL(homing_pigeon)
CT_CDL_DECL(identical_three_crows)
CT_CDL_DECL(in_neck)
CT_CDL_DECL(inverted_hammer)
CT_CDL_DECL(kicking)
CT_CDL_DECL(kicking_by_length)
CT_CDL_DECL(ladder_bottom)
CT_CDL_DECL(long_legged_doji)
CT_CDL_DECL(long_line)
CT_CDL_DECL(marubozu)
CT_CDL_DECL(matching_low)
CT_CDL_DECL(mat_hold)
CT_CDL_DECL(morning_doji_star)
CT_CDL_DECL(morning_star)
CT_CDL_DECL(on_neck)
CT_CDL_DECL(piercing)
CT_CDL_DECL(rickshaw_man)
CT_CDL_DECL(rise_fall3_methods)
CT_CDL_DECL(separating_lines)
CT_CDL_DECL(shooting_star)
CT_CDL_DECL(short_line)
CT_CDL_DECL(spinning_top)
CT_CDL_DECL(stalled_pattern)
CT_CDL_DECL(stick_sandwich)
CT_CDL_DECL(takuri)
CT_CDL_DECL(tasuki_gap)
CT_CDL_DECL(thrusting)
CT_CDL_DECL(tristar)
CT_CDL_DECL(unique_3_river)
CT_CDL_DECL(upside_gap_2_crows)
CT_CDL_DECL(xside_gap_3_methods)
/* The reference's legacy engulfing names (tacuda.h:447-452, :1917-1952;
 * tacuda_device.h:571-582): both return TA-Lib's signed engulfing signal
 * (+-100, or +-80 when an open equals the other candle's close), written
 * from bar 1 on (0 at bar 0). */
CT_CDL_DECL(bullish_engulfing)
CT_CDL_DECL(bearish_engulfing)
#undef CT_CDL_DECL


/* Price transforms, momentum, window midpoints and linear regression on
 * device pointers (csrc/transforms.hip; enqueued on the stream). */
/* :171 */ CTAPI_EXPORT ctStatus_t ct_momentum_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :174 */ CTAPI_EXPORT ctStatus_t ct_roc_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :177 */ CTAPI_EXPORT ctStatus_t ct_rocp_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :180 */ CTAPI_EXPORT ctStatus_t ct_rocr_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :183 */ CTAPI_EXPORT ctStatus_t ct_rocr100_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :447 */ CTAPI_EXPORT ctStatus_t ct_midpoint_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :357 */ CTAPI_EXPORT ctStatus_t ct_linearreg_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :360 */ CTAPI_EXPORT ctStatus_t ct_linearreg_slope_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :363 */ CTAPI_EXPORT ctStatus_t ct_linearreg_intercept_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :366 */ CTAPI_EXPORT ctStatus_t ct_linearreg_angle_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :369 */ CTAPI_EXPORT ctStatus_t ct_tsf_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :338 */ CTAPI_EXPORT ctStatus_t ct_avgprice_d(const float* d_open, const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :342 */ CTAPI_EXPORT ctStatus_t ct_medprice_d(const float* d_high, const float* d_low, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :345 */ CTAPI_EXPORT ctStatus_t ct_typprice_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :349 */ CTAPI_EXPORT ctStatus_t ct_wclprice_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :511 */ CTAPI_EXPORT ctStatus_t ct_midprice_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :399 */ CTAPI_EXPORT ctStatus_t ct_acos_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :401 */ CTAPI_EXPORT ctStatus_t ct_asin_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :403 */ CTAPI_EXPORT ctStatus_t ct_atan_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :405 */ CTAPI_EXPORT ctStatus_t ct_ceil_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :407 */ CTAPI_EXPORT ctStatus_t ct_cos_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :409 */ CTAPI_EXPORT ctStatus_t ct_cosh_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :411 */ CTAPI_EXPORT ctStatus_t ct_exp_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :413 */ CTAPI_EXPORT ctStatus_t ct_floor_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :415 */ CTAPI_EXPORT ctStatus_t ct_ln_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :417 */ CTAPI_EXPORT ctStatus_t ct_log10_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :419 */ CTAPI_EXPORT ctStatus_t ct_sin_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :421 */ CTAPI_EXPORT ctStatus_t ct_sinh_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :423 */ CTAPI_EXPORT ctStatus_t ct_sqrt_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :425 */ CTAPI_EXPORT ctStatus_t ct_tan_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :427 */ CTAPI_EXPORT ctStatus_t ct_tanh_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :429 */ CTAPI_EXPORT ctStatus_t ct_add_d(const float* d_input1, const float* d_input2, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :431 */ CTAPI_EXPORT ctStatus_t ct_sub_d(const float* d_input1, const float* d_input2, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :433 */ CTAPI_EXPORT ctStatus_t ct_mult_d(const float* d_input1, const float* d_input2, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :435 */ CTAPI_EXPORT ctStatus_t ct_div_d(const float* d_input1, const float* d_input2, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :551 */ CTAPI_EXPORT ctStatus_t ct_bop_d(const float* d_open, const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1120 */ CTAPI_EXPORT ctStatus_t ct_ibs_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :444 */ CTAPI_EXPORT ctStatus_t ct_change_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :450 */ CTAPI_EXPORT ctStatus_t ct_avgdev_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :495 */ CTAPI_EXPORT ctStatus_t ct_minmax_d(const float* d_input, float* d_min, float* d_max, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1064 */ CTAPI_EXPORT ctStatus_t ct_cumsum_d(const float* d_input, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1206 */ CTAPI_EXPORT ctStatus_t ct_pvt_d(const float* d_close, const float* d_volume, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1212 */ CTAPI_EXPORT ctStatus_t ct_rma_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1329 */ CTAPI_EXPORT ctStatus_t ct_zlema_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1084 */ CTAPI_EXPORT ctStatus_t ct_efi_d(const float* d_close, const float* d_volume, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1311 */ CTAPI_EXPORT ctStatus_t ct_vwma_d(const float* d_input, const float* d_volume, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1091 */ CTAPI_EXPORT ctStatus_t ct_er_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1297 */ CTAPI_EXPORT ctStatus_t ct_vhf_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1081 */ CTAPI_EXPORT ctStatus_t ct_dpo_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1033 */ CTAPI_EXPORT ctStatus_t ct_chop_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :998 */ CTAPI_EXPORT ctStatus_t ct_adr_d(const float* d_high, const float* d_low, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1005 */ CTAPI_EXPORT ctStatus_t ct_ao_d(const float* d_high, const float* d_low, float* d_output, int size, int fastPeriod, int slowPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1209 */ CTAPI_EXPORT ctStatus_t ct_qstick_d(const float* d_open, const float* d_close, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1071 */ CTAPI_EXPORT ctStatus_t ct_donchian_d(const float* d_high, const float* d_low, float* d_realUpperBand, float* d_realMiddleBand, float* d_realLowerBand, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1169 */ CTAPI_EXPORT ctStatus_t ct_marketfi_d(const float* d_high, const float* d_low, const float* d_volume, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1314 */ CTAPI_EXPORT ctStatus_t ct_wad_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1177 */ CTAPI_EXPORT ctStatus_t ct_mcgd_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1229 */ CTAPI_EXPORT ctStatus_t ct_rvol_d(const float* d_volume, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1046 */ CTAPI_EXPORT ctStatus_t ct_cmf_d(const float* d_high, const float* d_low, const float* d_close, const float* d_volume, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1087 */ CTAPI_EXPORT ctStatus_t ct_emv_d(const float* d_high, const float* d_low, const float* d_volume, float* d_output, int size, int period, double volumeDivisor, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1102 */ CTAPI_EXPORT ctStatus_t ct_fosc_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1094 */ CTAPI_EXPORT ctStatus_t ct_eri_d(const float* d_high, const float* d_low, const float* d_close, float* d_bullPower, float* d_bearPower, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1191 */ CTAPI_EXPORT ctStatus_t ct_percentrank_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1232 */ CTAPI_EXPORT ctStatus_t ct_si_d(const float* d_open, const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, double limitMove, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1013 */ CTAPI_EXPORT ctStatus_t ct_asi_wilder_d(const float* d_open, const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, double limitMove, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1303 */ CTAPI_EXPORT ctStatus_t ct_vortex_d(const float* d_high, const float* d_low, const float* d_close, float* d_plusVI, float* d_minusVI, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1050 */ CTAPI_EXPORT ctStatus_t ct_cmou_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :307 */ CTAPI_EXPORT ctStatus_t ct_obv_d(const float* d_close, const float* d_volume, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :298 */ CTAPI_EXPORT ctStatus_t ct_ad_d(const float* d_high, const float* d_low, const float* d_close, const float* d_volume, float* d_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :270 */ CTAPI_EXPORT ctStatus_t ct_ultosc_d(const float* d_high, const float* d_low, const float* d_close, float* d_output, int size, int period1, int period2, int period3, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :310 */ CTAPI_EXPORT ctStatus_t ct_mfi_d(const float* d_high, const float* d_low, const float* d_close, const float* d_volume, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :981 */ CTAPI_EXPORT ctStatus_t ct_rolling_mean_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :985 */ CTAPI_EXPORT ctStatus_t ct_rolling_max_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :988 */ CTAPI_EXPORT ctStatus_t ct_rolling_min_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :991 */ CTAPI_EXPORT ctStatus_t ct_rolling_std_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :972 */ CTAPI_EXPORT ctStatus_t ct_pivot_high_d(const float* d_input, float* d_output, int size, int strength, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :975 */ CTAPI_EXPORT ctStatus_t ct_pivot_low_d(const float* d_input, float* d_output, int size, int strength, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :966 */ CTAPI_EXPORT ctStatus_t ct_skew_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :960 */ CTAPI_EXPORT ctStatus_t ct_trendscore_d(const float* d_input, float* d_output, int size, int lookback, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :956 */ CTAPI_EXPORT ctStatus_t ct_market_strength_d(const float* d_input, float* d_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_TACUDA_DEVICE_H */
