// Device-pointer entry points of the indicator families, shared by the C ABI
// translation units (series, fused batch, streaming, autograd).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/tacuda.h"

namespace tca {
ctStatus_t dev_sma(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_sum(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_wma(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_ema(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_dema(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_tema(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_t3(const float* x, float* y, int len, int rows, int p, float v, hipStream_t s);
ctStatus_t dev_kama(const float* x, float* y, int len, int rows, int p, int fast, int slow,
                    hipStream_t s);
ctStatus_t dev_rsi(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_macd(const float* x, float* m, float* sg, float* h, int len, int rows, int fast,
                    int slow, int sig, int maType, hipStream_t s);
ctStatus_t dev_bbands(const float* x, float* u, float* m, float* l, int len, int rows, int p,
                      float up, float dn, hipStream_t s);
ctStatus_t dev_var(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_stddev(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_max(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_min(const float* x, float* y, int len, int rows, int p, hipStream_t s);
ctStatus_t dev_atr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                   int p, hipStream_t s);
ctStatus_t dev_natr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                    int p, hipStream_t s);
ctStatus_t dev_obv(const float* c, const float* v, float* y, int len, int rows, hipStream_t s);
ctStatus_t dev_ad(const float* h, const float* l, const float* c, const float* v, float* y,
                  int len, int rows, hipStream_t s);
ctStatus_t dev_trange(const float* h, const float* l, const float* c, float* y, int len, int rows,
                      hipStream_t s);
ctStatus_t dev_stoch(const float* h, const float* l, const float* c, float* k, float* d, int len,
                     int rows, int kp, int dp, hipStream_t s);
ctStatus_t dev_willr(const float* h, const float* l, const float* c, float* y, int len, int rows,
                     int p, hipStream_t s);
ctStatus_t dev_cci(const float* h, const float* l, const float* c, float* y, int len, int rows,
                   int p, hipStream_t s);
ctStatus_t dev_aroon(const float* h, const float* l, float* dn, float* up, int len, int rows,
                     int p, hipStream_t s);
ctStatus_t dev_aroonosc(const float* h, const float* l, float* y, int len, int rows, int p,
                        hipStream_t s);
ctStatus_t dev_plus_dm(const float* h, const float* l, const float* c, float* y, int len,
                       int rows, int p, hipStream_t s);
ctStatus_t dev_minus_dm(const float* h, const float* l, const float* c, float* y, int len,
                        int rows, int p, hipStream_t s);
ctStatus_t dev_plus_di(const float* h, const float* l, const float* c, float* y, int len,
                       int rows, int p, hipStream_t s);
ctStatus_t dev_minus_di(const float* h, const float* l, const float* c, float* y, int len,
                        int rows, int p, hipStream_t s);
ctStatus_t dev_dx(const float* h, const float* l, const float* c, float* y, int len, int rows,
                  int p, hipStream_t s);
ctStatus_t dev_adx(const float* h, const float* l, const float* c, float* y, int len, int rows,
                   int p, hipStream_t s);
ctStatus_t dev_trima(const float* x, float* y, int len, int rows, int p, hipStream_t s);
// TA-Lib MACDEXT: a moving-average type (TA_MAType numbering) per line.
ctStatus_t dev_macdext(const float* x, float* m, float* sg, float* h, int len, int rows, int fast,
                       int ft, int slow, int st, int sig, int sgt, hipStream_t s);
// The moving average `type` (ctMaType_t) of the families above.
ctStatus_t dev_ma(const float* x, float* y, int len, int rows, int p, int type, hipStream_t s);
}  // namespace tca
