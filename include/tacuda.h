/*
 * tacuda_amd -- MI355X-native indicator core behind the TACUDA C ABI.
 *
 * Status codes, error state, version, host (ct_*) calls and series-major
 * batch calls (ct_*_batch, ct_*_batch_d) for the hot-path families.  Every
 * declaration keeps the reference's name, argument order and meaning; the
 * only change is hipStream_t in place of cudaStream_t (an opaque pointer of
 * the same size).  Citations are to the reference's include/tacuda.h.
 */
#ifndef TACUDA_AMD_TACUDA_H
#define TACUDA_AMD_TACUDA_H

#include <stddef.h>
#include <stdint.h>

#ifndef HIP_INCLUDE_HIP_HIP_RUNTIME_API_H
typedef struct ihipStream_t* hipStream_t;
#endif

#if defined(_WIN32)
#define CTAPI_EXPORT __declspec(dllexport)
#else
#define CTAPI_EXPORT __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
#define CTAPI_STREAM_DEFAULT = 0
extern "C" {
#else
#define CTAPI_STREAM_DEFAULT
#endif

/* tacuda.h:32-40 */
typedef enum ctStatus {
  CT_STATUS_SUCCESS = 0,
  CT_STATUS_ALLOC_FAILED = 1,
  CT_STATUS_COPY_FAILED = 2,
  CT_STATUS_KERNEL_FAILED = 3,
  CT_STATUS_INVALID_ARGUMENT = 4,
  CT_STATUS_DEVICE_ERROR = 5,
  CT_STATUS_UNSUPPORTED = 6,
} ctStatus_t;

/* tacuda.h:45 */
CTAPI_EXPORT const char* ct_status_string(ctStatus_t status);
/* tacuda.h:54 -- thread-local, cleared on entry of every call */
CTAPI_EXPORT const char* ct_get_last_error(void);

#define CT_VERSION_MAJOR 0
#define CT_VERSION_MINOR 9
#define CT_VERSION_PATCH 3
#define CT_VERSION_STRING "0.9.3"
/* tacuda.h:70 */
CTAPI_EXPORT const char* ct_version(void);

/* tacuda.h:72-89 */
typedef enum ctMaType {
  CT_MA_SMA = 0,
  CT_MA_EMA = 1,
  CT_MA_WMA = 2,
  CT_MA_DEMA = 3,
  CT_MA_TEMA = 4,
  CT_MA_TRIMA = 5,
  CT_MA_KAMA = 6,
  CT_MA_MAMA = 7,
  CT_MA_T3 = 8,
  CT_MA_HMA = 9,
  CT_MA_DISABLED = 10,
  CT_MA_DEFAULT = 11,
  CT_MA_ZLEMA = 12,
  CT_MA_RMA = 13,
  CT_MA_VIDYA = 14,
  CT_MA_ALMA = 15,
} ctMaType_t;

/* tacuda.h:91-95 */
typedef enum ctDataType {
  CT_DTYPE_F16 = 0,
  CT_DTYPE_F32 = 1,
  CT_DTYPE_F64 = 2,
} ctDataType_t;

/* ---------------------------------------------------------------------------
 * Host calls: upload, run the device kernel, download, synchronize the stream
 * (reference api.cpp:230-287).  Device buffers come from a pool and arrays of
 * 1 MB or more are copied through pinned staging chunks, so a call allocates
 * nothing once the pool holds buffers of its sizes.  Lines cite tacuda.h.
 * ------------------------------------------------------------------------- */
/* :98 */ CTAPI_EXPORT ctStatus_t ct_sma(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :109 */ CTAPI_EXPORT ctStatus_t ct_wma(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :123 */ CTAPI_EXPORT ctStatus_t ct_ema(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :128 */ CTAPI_EXPORT ctStatus_t ct_dema(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :130 */ CTAPI_EXPORT ctStatus_t ct_tema(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :132 */ CTAPI_EXPORT ctStatus_t ct_t3(const float* host_input, float* host_output, int size, int period, float vFactor, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :138 */ CTAPI_EXPORT ctStatus_t ct_max(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :140 */ CTAPI_EXPORT ctStatus_t ct_min(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :151 */ CTAPI_EXPORT ctStatus_t ct_stddev(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :155 */ CTAPI_EXPORT ctStatus_t ct_var(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :157 */ CTAPI_EXPORT ctStatus_t ct_sum(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :159 */ CTAPI_EXPORT ctStatus_t ct_rsi(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :164 */ CTAPI_EXPORT ctStatus_t ct_kama(const float* host_input, float* host_output, int size, int period, int fastPeriod, int slowPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :171 */ CTAPI_EXPORT ctStatus_t ct_macd(const float* host_input, float* host_macd, float* host_signal, float* host_hist, int size, int fastPeriod, int slowPeriod, int signalPeriod, ctMaType_t type, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :299 */ CTAPI_EXPORT ctStatus_t ct_bbands(const float* host_input, float* host_upper, float* host_middle, float* host_lower, int size, int period, float upperMul, float lowerMul, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :308 */ CTAPI_EXPORT ctStatus_t ct_atr(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, float initial, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :317 */ CTAPI_EXPORT ctStatus_t ct_natr(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :321 */ CTAPI_EXPORT ctStatus_t ct_trange(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :324 */ CTAPI_EXPORT ctStatus_t ct_stochastic(const float* host_high, const float* host_low, const float* host_close, float* host_k, float* host_d, int size, int kPeriod, int dPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :337 */ CTAPI_EXPORT ctStatus_t ct_cci(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :340 */ CTAPI_EXPORT ctStatus_t ct_adx(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :346 */ CTAPI_EXPORT ctStatus_t ct_plus_dm(const float* host_high, const float* host_low, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :349 */ CTAPI_EXPORT ctStatus_t ct_minus_dm(const float* host_high, const float* host_low, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :352 */ CTAPI_EXPORT ctStatus_t ct_plus_di(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :356 */ CTAPI_EXPORT ctStatus_t ct_minus_di(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :385 */ CTAPI_EXPORT ctStatus_t ct_aroon(const float* host_high, const float* host_low, float* host_up, float* host_down, float* host_osc, int size, int upPeriod, int downPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :389 */ CTAPI_EXPORT ctStatus_t ct_aroonosc(const float* host_high, const float* host_low, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :419 */ CTAPI_EXPORT ctStatus_t ct_willr(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :739 */ CTAPI_EXPORT ctStatus_t ct_dx(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :103 -- moving average selected by type: SMA, EMA, WMA, DEMA, TEMA, KAMA
 * (fast 2, slow 30) and T3 (vFactor 0.7); other types return
 * CT_STATUS_UNSUPPORTED. */
CTAPI_EXPORT ctStatus_t ct_ma(const float* host_input, float* host_output, int size, int period, ctMaType_t type, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* Typed host variants (:100, :125, :161, :177, :303, :312): float32 storage
 * with float32 or float64 compute; other storage types return
 * CT_STATUS_UNSUPPORTED (see the _ex_d calls in tacuda_device.h). */
CTAPI_EXPORT ctStatus_t ct_sma_ex(const void* host_input, void* host_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ema_ex(const void* host_input, void* host_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rsi_ex(const void* host_input, void* host_output, int size, int period, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_macd_ex(const void* host_input, void* host_macd, void* host_signal, void* host_hist, int size, int fastPeriod, int slowPeriod, int signalPeriod, ctMaType_t type, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_bbands_ex(const void* host_input, void* host_upper, void* host_middle, void* host_lower, int size, int period, float upperMul, float lowerMul, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_atr_ex(const void* host_high, const void* host_low, const void* host_close, void* host_output, int size, int period, float initial, ctDataType_t storageType, ctDataType_t computeType, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * Device-pointer variants (:202-290): the same calls as ct_<name>_d
 * (tacuda_device.h), enqueued on the stream without waiting.
 * ------------------------------------------------------------------------- */
CTAPI_EXPORT ctStatus_t ct_sma_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ma_device(const float* device_input, float* device_output, int size, int period, ctMaType_t type, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_wma_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ema_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_dema_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_tema_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_t3_device(const float* device_input, float* device_output, int size, int period, float vFactor, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_max_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_min_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_stddev_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_var_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_sum_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rsi_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_kama_device(const float* device_input, float* device_output, int size, int period, int fastPeriod, int slowPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* :134, :243 -- TRIMA; :770 -- MACDEXT (see ct_macdext_d in tacuda_device.h) */
CTAPI_EXPORT ctStatus_t ct_trima(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_trima_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_macdext(const float* host_input, float* host_macd, float* host_signal, float* host_hist, int size, int fastPeriod, ctMaType_t fastType, int slowPeriod, ctMaType_t slowType, int signalPeriod, ctMaType_t signalType, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* tacuda.h:824-827 -- the workspace cache of this build stands in for the
 * reference's DeviceBufferPool.  Cleanup frees every look-back workspace,
 * including those that graphs captured earlier still reference: call it only
 * when no such graph will be replayed again. */
CTAPI_EXPORT void ct_device_buffer_pool_cleanup(void);
CTAPI_EXPORT size_t ct_device_buffer_pool_allocation_count(void);
CTAPI_EXPORT size_t ct_device_buffer_pool_set_max_cache_bytes(size_t bytes);

/* ---------------------------------------------------------------------------
 * Batch calls over dense series-major matrices: numSeries rows of seriesLen
 * contiguous floats (tacuda.h:830-847).  Every family runs as one fused
 * launch; maxConcurrency is accepted and ignored.  Stream 0 returns when the
 * batch is done; a non-null stream only enqueues (capturable).
 * ------------------------------------------------------------------------- */
#define CT_BATCH1(name, line)                                                                 \
  CTAPI_EXPORT ctStatus_t ct_##name##_batch(const float* host_inputs, float* host_outputs,    \
                                            int seriesLen, int numSeries, int period,        \
                                            int maxConcurrency,                              \
                                            hipStream_t stream CTAPI_STREAM_DEFAULT);        \
  CTAPI_EXPORT ctStatus_t ct_##name##_batch_d(const float* d_inputs, float* d_outputs,        \
                                              int seriesLen, int numSeries, int period,      \
                                              int maxConcurrency,                            \
                                              hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :848/:904 */ CT_BATCH1(sma, 848)
/* :853/:909 */ CT_BATCH1(ema, 853)
/* :858/:914 */ CT_BATCH1(rsi, 858)
/* :953/:958 */ CT_BATCH1(wma, 953)
/* :962/:967 */ CT_BATCH1(dema, 962)
/* :972/:977 */ CT_BATCH1(tema, 972)
/* :1050/:1055 */ CT_BATCH1(stddev, 1050)
/* :1060/:1065 */ CT_BATCH1(var, 1060)
/* :1069/:1074 */ CT_BATCH1(sum, 1069)
/* :1078/:1083 */ CT_BATCH1(max, 1078)
/* :1087/:1092 */ CT_BATCH1(min, 1087)
#undef CT_BATCH1

#define CT_BATCH_HLC(name)                                                                    \
  CTAPI_EXPORT ctStatus_t ct_##name##_batch(const float* host_highs, const float* host_lows,  \
                                            const float* host_closes, float* host_outputs,   \
                                            int seriesLen, int numSeries, int period,        \
                                            int maxConcurrency,                              \
                                            hipStream_t stream CTAPI_STREAM_DEFAULT);        \
  CTAPI_EXPORT ctStatus_t ct_##name##_batch_d(const float* d_highs, const float* d_lows,      \
                                              const float* d_closes, float* d_outputs,       \
                                              int seriesLen, int numSeries, int period,      \
                                              int maxConcurrency,                            \
                                              hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1173/:1180 */ CT_BATCH_HLC(natr)
/* :1187/:1194 */ CT_BATCH_HLC(adx)
/* :1214/:1221 */ CT_BATCH_HLC(dx)
/* :1227/:1234 */ CT_BATCH_HLC(plus_di)
/* :1241/:1248 */ CT_BATCH_HLC(minus_di)
/* :1255/:1262 */ CT_BATCH_HLC(willr)
/* :1269/:1276 */ CT_BATCH_HLC(cci)
#undef CT_BATCH_HLC

/* :863 / :919 */
CTAPI_EXPORT ctStatus_t ct_atr_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int period, float initial, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_atr_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int period, float initial, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :870 / :926 */
CTAPI_EXPORT ctStatus_t ct_macd_batch(const float* host_inputs, float* host_macds, float* host_signals, float* host_hists, int seriesLen, int numSeries, int fastPeriod, int slowPeriod, int signalPeriod, int maType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_macd_batch_d(const float* d_inputs, float* d_macds, float* d_signals, float* d_hists, int seriesLen, int numSeries, int fastPeriod, int slowPeriod, int signalPeriod, int maType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :877 / :933 */
CTAPI_EXPORT ctStatus_t ct_bbands_batch(const float* host_inputs, float* host_uppers, float* host_middles, float* host_lowers, int seriesLen, int numSeries, int period, float upperMul, float lowerMul, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_bbands_batch_d(const float* d_inputs, float* d_uppers, float* d_middles, float* d_lowers, int seriesLen, int numSeries, int period, float upperMul, float lowerMul, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :885 / :940 */
CTAPI_EXPORT ctStatus_t ct_stochastic_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_ks, float* host_ds, int seriesLen, int numSeries, int kPeriod, int dPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_stochastic_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_ks, float* d_ds, int seriesLen, int numSeries, int kPeriod, int dPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1152 / :1158 */
CTAPI_EXPORT ctStatus_t ct_kama_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int fastPeriod, int slowPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_kama_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int fastPeriod, int slowPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1164 / :1169 */
CTAPI_EXPORT ctStatus_t ct_t3_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, float vFactor, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_t3_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, float vFactor, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1282 / :1288 */
CTAPI_EXPORT ctStatus_t ct_trange_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_trange_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1344 / :1350 */
CTAPI_EXPORT ctStatus_t ct_aroonosc_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_aroonosc_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1356 / :1363 (down before up, as in the reference) */
CTAPI_EXPORT ctStatus_t ct_aroon_batch(const float* host_highs, const float* host_lows, float* host_aroon_downs, float* host_aroon_ups, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_aroon_batch_d(const float* d_highs, const float* d_lows, float* d_aroon_downs, float* d_aroon_ups, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1581 / :1586 -- moving average selected by type (see ct_ma) */
CTAPI_EXPORT ctStatus_t ct_ma_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_ma_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :982 / :987 */
CTAPI_EXPORT ctStatus_t ct_trima_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_trima_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1387 / :1396 */
CTAPI_EXPORT ctStatus_t ct_macdext_batch(const float* host_inputs, float* host_macds, float* host_signals, float* host_hists, int seriesLen, int numSeries, int fastPeriod, int fastType, int slowPeriod, int slowType, int signalPeriod, int signalType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_macdext_batch_d(const float* d_inputs, float* d_macds, float* d_signals, float* d_hists, int seriesLen, int numSeries, int fastPeriod, int fastType, int slowPeriod, int slowType, int signalPeriod, int signalType, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1733 / :1739, :1745 / :1751 */
CTAPI_EXPORT ctStatus_t ct_plus_dm_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_plus_dm_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_minus_dm_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_minus_dm_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);


/* ---------------------------------------------------------------------------
 * Price transforms, the momentum family, window midpoints and the
 * linear-regression family (csrc/transforms.hip): host, device-alias and
 * batch forms.  Arithmetic in double with TA-Lib 0.8.3's formulas; a
 * period of 0 or less is CT_STATUS_INVALID_ARGUMENT, a period beyond the
 * series (MOMENTUM/ROC*: at least the series) gives an all-NaN line.
 * ------------------------------------------------------------------------- */
/* :111 */ CTAPI_EXPORT ctStatus_t ct_momentum(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :115 */ CTAPI_EXPORT ctStatus_t ct_roc(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :117 */ CTAPI_EXPORT ctStatus_t ct_rocp(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :119 */ CTAPI_EXPORT ctStatus_t ct_rocr(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :121 */ CTAPI_EXPORT ctStatus_t ct_rocr100(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :422 */ CTAPI_EXPORT ctStatus_t ct_midpoint(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :742 */ CTAPI_EXPORT ctStatus_t ct_linearreg(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :744 */ CTAPI_EXPORT ctStatus_t ct_linearreg_slope(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :747 */ CTAPI_EXPORT ctStatus_t ct_linearreg_intercept(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :750 */ CTAPI_EXPORT ctStatus_t ct_linearreg_angle(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :753 */ CTAPI_EXPORT ctStatus_t ct_tsf(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :403 */ CTAPI_EXPORT ctStatus_t ct_avgprice(const float* host_open, const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :408 */ CTAPI_EXPORT ctStatus_t ct_medprice(const float* host_high, const float* host_low, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :411 */ CTAPI_EXPORT ctStatus_t ct_typprice(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :415 */ CTAPI_EXPORT ctStatus_t ct_wclprice(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :424 */ CTAPI_EXPORT ctStatus_t ct_midprice(const float* host_high, const float* host_low, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :212 */ CTAPI_EXPORT ctStatus_t ct_momentum_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :218 */ CTAPI_EXPORT ctStatus_t ct_roc_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :221 */ CTAPI_EXPORT ctStatus_t ct_rocp_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :224 */ CTAPI_EXPORT ctStatus_t ct_rocr_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :227 */ CTAPI_EXPORT ctStatus_t ct_rocr100_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :992 / :997 */
CTAPI_EXPORT ctStatus_t ct_momentum_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_momentum_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1002 / :1007 */
CTAPI_EXPORT ctStatus_t ct_roc_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_roc_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1011 / :1016 */
CTAPI_EXPORT ctStatus_t ct_rocp_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rocp_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1021 / :1026 */
CTAPI_EXPORT ctStatus_t ct_rocr_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rocr_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1031 / :1036 */
CTAPI_EXPORT ctStatus_t ct_rocr100_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_rocr100_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1511 / :1516 */
CTAPI_EXPORT ctStatus_t ct_midpoint_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_midpoint_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1096 / :1101 */
CTAPI_EXPORT ctStatus_t ct_linearreg_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_linearreg_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1106 / :1112 */
CTAPI_EXPORT ctStatus_t ct_linearreg_slope_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_linearreg_slope_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1118 / :1124 */
CTAPI_EXPORT ctStatus_t ct_linearreg_intercept_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_linearreg_intercept_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1131 / :1137 */
CTAPI_EXPORT ctStatus_t ct_linearreg_angle_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_linearreg_angle_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1143 / :1148 */
CTAPI_EXPORT ctStatus_t ct_tsf_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_tsf_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1421 / :1428 */
CTAPI_EXPORT ctStatus_t ct_avgprice_batch(const float* host_opens, const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_avgprice_batch_d(const float* d_opens, const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1334 / :1339 */
CTAPI_EXPORT ctStatus_t ct_medprice_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_medprice_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1294 / :1300 */
CTAPI_EXPORT ctStatus_t ct_typprice_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_typprice_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1306 / :1312 */
CTAPI_EXPORT ctStatus_t ct_wclprice_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_wclprice_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1721 / :1727 */
CTAPI_EXPORT ctStatus_t ct_midprice_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
CTAPI_EXPORT ctStatus_t ct_midprice_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * Element-wise math transforms and operators (csrc/vecmath.hip): the
 * float function of each element (device-library single precision for the
 * transcendental ones; CEIL, FLOOR, SQRT and the operators exactly rounded).
 * BOP (float, 0 on a zero range), IBS (double, 0.5 on a non-positive
 * range), CHANGE (float x[t] - x[t-period]), AVGDEV (float mean absolute
 * deviation, the reference's order) and MINMAX (NaN-skipping window
 * extrema) in csrc/transforms.hip.  Reference lines cited per declaration.
 * ------------------------------------------------------------------------- */
/* :781 */ CTAPI_EXPORT ctStatus_t ct_acos(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :783 */ CTAPI_EXPORT ctStatus_t ct_asin(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :785 */ CTAPI_EXPORT ctStatus_t ct_atan(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :787 */ CTAPI_EXPORT ctStatus_t ct_ceil(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :789 */ CTAPI_EXPORT ctStatus_t ct_cos(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :791 */ CTAPI_EXPORT ctStatus_t ct_cosh(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :793 */ CTAPI_EXPORT ctStatus_t ct_exp(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :795 */ CTAPI_EXPORT ctStatus_t ct_floor(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :797 */ CTAPI_EXPORT ctStatus_t ct_ln(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :799 */ CTAPI_EXPORT ctStatus_t ct_log10(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :801 */ CTAPI_EXPORT ctStatus_t ct_sin(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :803 */ CTAPI_EXPORT ctStatus_t ct_sinh(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :805 */ CTAPI_EXPORT ctStatus_t ct_sqrt(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :807 */ CTAPI_EXPORT ctStatus_t ct_tan(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :809 */ CTAPI_EXPORT ctStatus_t ct_tanh(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :811 */ CTAPI_EXPORT ctStatus_t ct_add(const float* host_input1, const float* host_input2, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :814 */ CTAPI_EXPORT ctStatus_t ct_sub(const float* host_input1, const float* host_input2, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :817 */ CTAPI_EXPORT ctStatus_t ct_mult(const float* host_input1, const float* host_input2, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :820 */ CTAPI_EXPORT ctStatus_t ct_div(const float* host_input1, const float* host_input2, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :429 */ CTAPI_EXPORT ctStatus_t ct_bop(const float* host_open, const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3732 */ CTAPI_EXPORT ctStatus_t ct_ibs(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :113 */ CTAPI_EXPORT ctStatus_t ct_change(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :153 */ CTAPI_EXPORT ctStatus_t ct_avgdev(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :146 */ CTAPI_EXPORT ctStatus_t ct_minmax(const float* host_input, float* host_min, float* host_max, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :215 */ CTAPI_EXPORT ctStatus_t ct_change_device(const float* device_input, float* device_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1856 */ CTAPI_EXPORT ctStatus_t ct_bop_batch(const float* host_opens, const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1863 */ CTAPI_EXPORT ctStatus_t ct_bop_batch_d(const float* d_opens, const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3736 */ CTAPI_EXPORT ctStatus_t ct_ibs_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3742 */ CTAPI_EXPORT ctStatus_t ct_ibs_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1501 */ CTAPI_EXPORT ctStatus_t ct_change_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1506 */ CTAPI_EXPORT ctStatus_t ct_change_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1521 */ CTAPI_EXPORT ctStatus_t ct_avgdev_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1526 */ CTAPI_EXPORT ctStatus_t ct_avgdev_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1666 */ CTAPI_EXPORT ctStatus_t ct_minmax_batch(const float* host_inputs, float* host_mins, float* host_maxs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1671 */ CTAPI_EXPORT ctStatus_t ct_minmax_batch_d(const float* d_inputs, float* d_mins, float* d_maxs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * TA-Lib 0.8.3 row functions (csrc/rowfn.hip): CUMSUM, PVT, RMA, ZLEMA, EFI,
 * VWMA, ER, VHF, DPO, CHOP, ADR, AO, QSTICK, DONCHIAN, MARKETFI, WAD, MCGD,
 * RVOL, CMF, EMV, FOSC, ERI, PERCENTRANK, SI, ASI_WILDER (TA_ASI), VORTEX
 * and CMOU.  Per row: the leading bars where any input is NaN are skipped,
 * TA-Lib's routine runs in double over the rest, outputs are float with NaN
 * before the first one.  Periods outside TA-Lib's ranges are
 * CT_STATUS_INVALID_ARGUMENT; device calls on stream 0 synchronize.
 * ------------------------------------------------------------------------- */
/* :3491 */ CTAPI_EXPORT ctStatus_t ct_cumsum(const float* host_input, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4072 */ CTAPI_EXPORT ctStatus_t ct_pvt(const float* host_close, const float* host_volume, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4102 */ CTAPI_EXPORT ctStatus_t ct_rma(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4595 */ CTAPI_EXPORT ctStatus_t ct_zlema(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3569 */ CTAPI_EXPORT ctStatus_t ct_efi(const float* host_close, const float* host_volume, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4516 */ CTAPI_EXPORT ctStatus_t ct_vwma(const float* host_input, const float* host_volume, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3604 */ CTAPI_EXPORT ctStatus_t ct_er(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4452 */ CTAPI_EXPORT ctStatus_t ct_vhf(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3557 */ CTAPI_EXPORT ctStatus_t ct_dpo(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3354 */ CTAPI_EXPORT ctStatus_t ct_chop(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3221 */ CTAPI_EXPORT ctStatus_t ct_adr(const float* host_high, const float* host_low, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3250 */ CTAPI_EXPORT ctStatus_t ct_ao(const float* host_high, const float* host_low, float* host_output, int size, int fastPeriod, int slowPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4086 */ CTAPI_EXPORT ctStatus_t ct_qstick(const float* host_open, const float* host_close, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3518 */ CTAPI_EXPORT ctStatus_t ct_donchian(const float* host_high, const float* host_low, float* host_realUpperBand, float* host_realMiddleBand, float* host_realLowerBand, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3921 */ CTAPI_EXPORT ctStatus_t ct_marketfi(const float* host_high, const float* host_low, const float* host_volume, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4532 */ CTAPI_EXPORT ctStatus_t ct_wad(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3954 */ CTAPI_EXPORT ctStatus_t ct_mcgd(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4170 */ CTAPI_EXPORT ctStatus_t ct_rvol(const float* host_volume, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3494 */ CTAPI_EXPORT ctStatus_t ct_cumsum_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3498 */ CTAPI_EXPORT ctStatus_t ct_cumsum_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4076 */ CTAPI_EXPORT ctStatus_t ct_pvt_batch(const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4081 */ CTAPI_EXPORT ctStatus_t ct_pvt_batch_d(const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4105 */ CTAPI_EXPORT ctStatus_t ct_rma_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4110 */ CTAPI_EXPORT ctStatus_t ct_rma_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4598 */ CTAPI_EXPORT ctStatus_t ct_zlema_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4603 */ CTAPI_EXPORT ctStatus_t ct_zlema_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3573 */ CTAPI_EXPORT ctStatus_t ct_efi_batch(const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3579 */ CTAPI_EXPORT ctStatus_t ct_efi_batch_d(const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4520 */ CTAPI_EXPORT ctStatus_t ct_vwma_batch(const float* host_inputs, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4526 */ CTAPI_EXPORT ctStatus_t ct_vwma_batch_d(const float* d_inputs, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3607 */ CTAPI_EXPORT ctStatus_t ct_er_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3612 */ CTAPI_EXPORT ctStatus_t ct_er_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4455 */ CTAPI_EXPORT ctStatus_t ct_vhf_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4460 */ CTAPI_EXPORT ctStatus_t ct_vhf_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3560 */ CTAPI_EXPORT ctStatus_t ct_dpo_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3565 */ CTAPI_EXPORT ctStatus_t ct_dpo_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3358 */ CTAPI_EXPORT ctStatus_t ct_chop_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3365 */ CTAPI_EXPORT ctStatus_t ct_chop_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3224 */ CTAPI_EXPORT ctStatus_t ct_adr_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3230 */ CTAPI_EXPORT ctStatus_t ct_adr_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3254 */ CTAPI_EXPORT ctStatus_t ct_ao_batch(const float* host_highs, const float* host_lows, float* host_outputs, int seriesLen, int numSeries, int fastPeriod, int slowPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3260 */ CTAPI_EXPORT ctStatus_t ct_ao_batch_d(const float* d_highs, const float* d_lows, float* d_outputs, int seriesLen, int numSeries, int fastPeriod, int slowPeriod, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4090 */ CTAPI_EXPORT ctStatus_t ct_qstick_batch(const float* host_opens, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4096 */ CTAPI_EXPORT ctStatus_t ct_qstick_batch_d(const float* d_opens, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3525 */ CTAPI_EXPORT ctStatus_t ct_donchian_batch(const float* host_highs, const float* host_lows, float* host_realUpperBand, float* host_realMiddleBand, float* host_realLowerBand, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3533 */ CTAPI_EXPORT ctStatus_t ct_donchian_batch_d(const float* d_highs, const float* d_lows, float* d_realUpperBand, float* d_realMiddleBand, float* d_realLowerBand, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3926 */ CTAPI_EXPORT ctStatus_t ct_marketfi_batch(const float* host_highs, const float* host_lows, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3932 */ CTAPI_EXPORT ctStatus_t ct_marketfi_batch_d(const float* d_highs, const float* d_lows, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4536 */ CTAPI_EXPORT ctStatus_t ct_wad_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4542 */ CTAPI_EXPORT ctStatus_t ct_wad_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3957 */ CTAPI_EXPORT ctStatus_t ct_mcgd_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3962 */ CTAPI_EXPORT ctStatus_t ct_mcgd_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4173 */ CTAPI_EXPORT ctStatus_t ct_rvol_batch(const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4178 */ CTAPI_EXPORT ctStatus_t ct_rvol_batch_d(const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3413 */ CTAPI_EXPORT ctStatus_t ct_cmf(const float* host_high, const float* host_low, const float* host_close, const float* host_volume, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3585 */ CTAPI_EXPORT ctStatus_t ct_emv(const float* host_high, const float* host_low, const float* host_volume, float* host_output, int size, int period, double volumeDivisor, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3652 */ CTAPI_EXPORT ctStatus_t ct_fosc(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3616 */ CTAPI_EXPORT ctStatus_t ct_eri(const float* host_high, const float* host_low, const float* host_close, float* host_bullPower, float* host_bearPower, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4012 */ CTAPI_EXPORT ctStatus_t ct_percentrank(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4183 */ CTAPI_EXPORT ctStatus_t ct_si(const float* host_open, const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, double limitMove, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3278 */ CTAPI_EXPORT ctStatus_t ct_asi_wilder(const float* host_open, const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, double limitMove, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3418 */ CTAPI_EXPORT ctStatus_t ct_cmf_batch(const float* host_highs, const float* host_lows, const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3426 */ CTAPI_EXPORT ctStatus_t ct_cmf_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3589 */ CTAPI_EXPORT ctStatus_t ct_emv_batch(const float* host_highs, const float* host_lows, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, double volumeDivisor, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3596 */ CTAPI_EXPORT ctStatus_t ct_emv_batch_d(const float* d_highs, const float* d_lows, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, double volumeDivisor, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3655 */ CTAPI_EXPORT ctStatus_t ct_fosc_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3660 */ CTAPI_EXPORT ctStatus_t ct_fosc_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3620 */ CTAPI_EXPORT ctStatus_t ct_eri_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_bullPower, float* host_bearPower, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3628 */ CTAPI_EXPORT ctStatus_t ct_eri_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_bullPower, float* d_bearPower, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4016 */ CTAPI_EXPORT ctStatus_t ct_percentrank_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4021 */ CTAPI_EXPORT ctStatus_t ct_percentrank_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4187 */ CTAPI_EXPORT ctStatus_t ct_si_batch(const float* host_opens, const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, double limitMove, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4195 */ CTAPI_EXPORT ctStatus_t ct_si_batch_d(const float* d_opens, const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, double limitMove, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3285 */ CTAPI_EXPORT ctStatus_t ct_asi_wilder_batch(const float* host_opens, const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, double limitMove, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3293 */ CTAPI_EXPORT ctStatus_t ct_asi_wilder_batch_d(const float* d_opens, const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, double limitMove, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4477 */ CTAPI_EXPORT ctStatus_t ct_vortex(const float* host_high, const float* host_low, const float* host_close, float* host_plusVI, float* host_minusVI, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3434 */ CTAPI_EXPORT ctStatus_t ct_cmou(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4482 */ CTAPI_EXPORT ctStatus_t ct_vortex_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_plusVI, float* host_minusVI, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :4490 */ CTAPI_EXPORT ctStatus_t ct_vortex_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_plusVI, float* d_minusVI, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3437 */ CTAPI_EXPORT ctStatus_t ct_cmou_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3442 */ CTAPI_EXPORT ctStatus_t ct_cmou_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * OBV and AD (csrc/families.hpp FCumT on the look-back scan engine): running
 * sums in double of sign(close change) x volume (OBV; bar 0 the volume) and
 * of the close-location value x volume (AD), rounded to float per bar.
 * ------------------------------------------------------------------------- */
/* :367 */ CTAPI_EXPORT ctStatus_t ct_obv(const float* host_price, const float* host_volume, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :400 */ CTAPI_EXPORT ctStatus_t ct_ad(const float* host_high, const float* host_low, const float* host_close, const float* host_volume, float* host_output, int size, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1435 */ CTAPI_EXPORT ctStatus_t ct_obv_batch(const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1440 */ CTAPI_EXPORT ctStatus_t ct_obv_batch_d(const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1461 */ CTAPI_EXPORT ctStatus_t ct_ad_batch(const float* host_highs, const float* host_lows, const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1468 */ CTAPI_EXPORT ctStatus_t ct_ad_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * ULTOSC (float windows in the reference's order) and MFI (double typical
 * price flows summed per window) in csrc/transforms.hip.
 * ------------------------------------------------------------------------- */
/* :396 */ CTAPI_EXPORT ctStatus_t ct_ultosc(const float* host_high, const float* host_low, const float* host_close, float* host_output, int size, int shortPeriod, int mediumPeriod, int longPeriod, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :360 */ CTAPI_EXPORT ctStatus_t ct_mfi(const float* host_high, const float* host_low, const float* host_close, const float* host_volume, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1318 */ CTAPI_EXPORT ctStatus_t ct_ultosc_batch(const float* host_highs, const float* host_lows, const float* host_closes, float* host_outputs, int seriesLen, int numSeries, int period1, int period2, int period3, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1326 */ CTAPI_EXPORT ctStatus_t ct_ultosc_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, float* d_outputs, int seriesLen, int numSeries, int period1, int period2, int period3, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1445 */ CTAPI_EXPORT ctStatus_t ct_mfi_batch(const float* host_highs, const float* host_lows, const float* host_closes, const float* host_volumes, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :1453 */ CTAPI_EXPORT ctStatus_t ct_mfi_batch_d(const float* d_highs, const float* d_lows, const float* d_closes, const float* d_volumes, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

/* ---------------------------------------------------------------------------
 * Strict rolling statistics (csrc/transforms.hip): ROLLING_MEAN, _MAX, _MIN
 * and _STD over the period bars ending at each bar, NaN when any of them is
 * not finite; a period beyond the series is CT_STATUS_INVALID_ARGUMENT.
 * PIVOT_HIGH / PIVOT_LOW: the pivot's value at the bar that confirms it.
 * SKEW: the population skewness of the window.  TRENDSCORE and
 * MARKET_STRENGTH: the reference's own counting and change-ratio scores.
 * ------------------------------------------------------------------------- */
/* :3148 */ CTAPI_EXPORT ctStatus_t ct_rolling_mean(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3162 */ CTAPI_EXPORT ctStatus_t ct_rolling_max(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3176 */ CTAPI_EXPORT ctStatus_t ct_rolling_min(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3190 */ CTAPI_EXPORT ctStatus_t ct_rolling_std(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3152 */ CTAPI_EXPORT ctStatus_t ct_rolling_mean_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3157 */ CTAPI_EXPORT ctStatus_t ct_rolling_mean_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3166 */ CTAPI_EXPORT ctStatus_t ct_rolling_max_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3171 */ CTAPI_EXPORT ctStatus_t ct_rolling_max_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3180 */ CTAPI_EXPORT ctStatus_t ct_rolling_min_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3185 */ CTAPI_EXPORT ctStatus_t ct_rolling_min_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3194 */ CTAPI_EXPORT ctStatus_t ct_rolling_std_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3199 */ CTAPI_EXPORT ctStatus_t ct_rolling_std_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3107 */ CTAPI_EXPORT ctStatus_t ct_pivot_high(const float* host_input, float* host_output, int size, int strength, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3121 */ CTAPI_EXPORT ctStatus_t ct_pivot_low(const float* host_input, float* host_output, int size, int strength, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3111 */ CTAPI_EXPORT ctStatus_t ct_pivot_high_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int strength, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3116 */ CTAPI_EXPORT ctStatus_t ct_pivot_high_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int strength, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3125 */ CTAPI_EXPORT ctStatus_t ct_pivot_low_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int strength, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3130 */ CTAPI_EXPORT ctStatus_t ct_pivot_low_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int strength, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3081 */ CTAPI_EXPORT ctStatus_t ct_skew(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3084 */ CTAPI_EXPORT ctStatus_t ct_skew_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3089 */ CTAPI_EXPORT ctStatus_t ct_skew_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3054 */ CTAPI_EXPORT ctStatus_t ct_trendscore(const float* host_input, float* host_output, int size, int lookback, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3038 */ CTAPI_EXPORT ctStatus_t ct_market_strength(const float* host_input, float* host_output, int size, int period, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3058 */ CTAPI_EXPORT ctStatus_t ct_trendscore_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int lookback, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3063 */ CTAPI_EXPORT ctStatus_t ct_trendscore_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int lookback, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3042 */ CTAPI_EXPORT ctStatus_t ct_market_strength_batch(const float* host_inputs, float* host_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);
/* :3048 */ CTAPI_EXPORT ctStatus_t ct_market_strength_batch_d(const float* d_inputs, float* d_outputs, int seriesLen, int numSeries, int period, int maxConcurrency, hipStream_t stream CTAPI_STREAM_DEFAULT);

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_TACUDA_H */
