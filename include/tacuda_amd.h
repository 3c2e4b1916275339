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

#ifdef __cplusplus
}
#endif
#endif /* TACUDA_AMD_EXT_H */
