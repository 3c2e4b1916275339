// Element-wise vector math: TA-Lib's math transforms (ACOS, ASIN, ATAN, CEIL,
// COS, COSH, EXP, FLOOR, LN, LOG10, SIN, SINH, SQRT, TAN, TANH) and math
// operators (ADD, SUB, MULT, DIV) in host and device forms.  Reference entry
// points: include/tacuda.h:779-822, include/tacuda_device.h:399-438;
// behaviour: src/api/api_vector_math.cu (float in, float out, the float
// function of the element; no look-back; size <= 0 or a null pointer is an
// invalid argument).
//
// HBM-bound streaming: each work-item moves 16-byte vectors (four floats)
// when every pointer is 16-byte aligned, a grid of at most 8 waves' worth of
// blocks per CU strides over the rest.  The transcendental functions are
// the device library's single-precision ones (a few ulp from the correctly
// rounded value, as are CUDA's); CEIL, FLOOR, SQRT and the four operators
// are exactly rounded, so they agree with any IEEE implementation bit for
// bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/tacuda.h"
#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"

namespace tca {
namespace {

constexpr int kVB = 256;  // threads per block

enum UOp {
  kAcos, kAsin, kAtan, kCeil, kCos, kCosh, kExp, kFloor, kLn, kLog10, kSin, kSinh, kSqrt, kTan,
  kTanh
};
enum BOp { kAdd, kSub, kMult, kDiv };

template <int OP>
__device__ __forceinline__ float uop(float x) {
  if constexpr (OP == kAcos) return acosf(x);
  else if constexpr (OP == kAsin) return asinf(x);
  else if constexpr (OP == kAtan) return atanf(x);
  else if constexpr (OP == kCeil) return ceilf(x);
  else if constexpr (OP == kCos) return cosf(x);
  else if constexpr (OP == kCosh) return coshf(x);
  else if constexpr (OP == kExp) return expf(x);
  else if constexpr (OP == kFloor) return floorf(x);
  else if constexpr (OP == kLn) return logf(x);
  else if constexpr (OP == kLog10) return log10f(x);
  else if constexpr (OP == kSin) return sinf(x);
  else if constexpr (OP == kSinh) return sinhf(x);
  else if constexpr (OP == kSqrt) return sqrtf(x);
  else if constexpr (OP == kTan) return tanf(x);
  else return tanhf(x);
}

template <int OP>
__device__ __forceinline__ float bop(float a, float b) {
  if constexpr (OP == kAdd) return a + b;
  else if constexpr (OP == kSub) return a - b;
  else if constexpr (OP == kMult) return a * b;
  else return a / b;
}

// VEC: all pointers 16-byte aligned; the float4 part covers n & ~3, the
// scalar tail the rest.
template <int OP, bool VEC>
__global__ __launch_bounds__(kVB) void k_unary(const float* __restrict__ x, float* __restrict__ y,
                                               long long n) {
  const long long stride = (long long)gridDim.x * kVB;
  long long i = (long long)blockIdx.x * kVB + threadIdx.x;
  long long tail = 0;
  if constexpr (VEC) {
    const long long n4 = n >> 2;
    const float4* x4 = reinterpret_cast<const float4*>(x);
    float4* y4 = reinterpret_cast<float4*>(y);
    for (long long j = i; j < n4; j += stride) {
      const float4 v = x4[j];
      y4[j] = make_float4(uop<OP>(v.x), uop<OP>(v.y), uop<OP>(v.z), uop<OP>(v.w));
    }
    tail = n4 << 2;
  }
  for (long long j = tail + i; j < n; j += stride) y[j] = uop<OP>(x[j]);
}

template <int OP, bool VEC>
__global__ __launch_bounds__(kVB) void k_binary(const float* __restrict__ a,
                                                const float* __restrict__ b,
                                                float* __restrict__ y, long long n) {
  const long long stride = (long long)gridDim.x * kVB;
  long long i = (long long)blockIdx.x * kVB + threadIdx.x;
  long long tail = 0;
  if constexpr (VEC) {
    const long long n4 = n >> 2;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float4* y4 = reinterpret_cast<float4*>(y);
    for (long long j = i; j < n4; j += stride) {
      const float4 u = a4[j], v = b4[j];
      y4[j] = make_float4(bop<OP>(u.x, v.x), bop<OP>(u.y, v.y), bop<OP>(u.z, v.z),
                          bop<OP>(u.w, v.w));
    }
    tail = n4 << 2;
  }
  for (long long j = tail + i; j < n; j += stride) y[j] = bop<OP>(a[j], b[j]);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// One work-item per float4, at most 2048 blocks of four waves (eight blocks
// per CU); the grid-stride loop covers longer arrays.
dim3 vgrid(long long n) {
  const long long b = ((n + 3) / 4 + kVB - 1) / kVB;
  return dim3((unsigned)(b < 2048 ? (b > 0 ? b : 1) : 2048));
}

ctStatus_t vm_args(const char* fn, const void* a, const void* b, const void* y, int n) {
  if (!a || !b || !y || n <= 0)
    return fail(CT_STATUS_INVALID_ARGUMENT, "%s: null pointer or non-positive size", fn);
  return CT_STATUS_SUCCESS;
}

template <int OP>
ctStatus_t run_unary(const char* fn, const float* x, float* y, int n, hipStream_t s) {
  ctStatus_t r = vm_args(fn, x, x, y, n);
  if (r != CT_STATUS_SUCCESS) return r;
  if (aligned16(x) && aligned16(y))
    hipLaunchKernelGGL((k_unary<OP, true>), vgrid(n), dim3(kVB), 0, s, x, y, (long long)n);
  else
    hipLaunchKernelGGL((k_unary<OP, false>), vgrid(4ll * n), dim3(kVB), 0, s, x, y, (long long)n);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

template <int OP>
ctStatus_t run_binary(const char* fn, const float* a, const float* b, float* y, int n,
                      hipStream_t s) {
  ctStatus_t r = vm_args(fn, a, b, y, n);
  if (r != CT_STATUS_SUCCESS) return r;
  if (aligned16(a) && aligned16(b) && aligned16(y))
    hipLaunchKernelGGL((k_binary<OP, true>), vgrid(n), dim3(kVB), 0, s, a, b, y, (long long)n);
  else
    hipLaunchKernelGGL((k_binary<OP, false>), vgrid(4ll * n), dim3(kVB), 0, s, a, b, y,
                       (long long)n);
  TCA_HIP(hipGetLastError());
  return CT_STATUS_SUCCESS;
}

#define VM_RET(x)                               \
  do {                                          \
    const ctStatus_t r_ = (x);                  \
    if (r_ != CT_STATUS_SUCCESS) return r_;     \
  } while (0)

// Host forms: the reference's pooled device buffers and one synchronization
// (api_vector_math.cu run_unary_host / run_binary_host), here hostio's
// pooled staging.
template <int OP>
ctStatus_t host_unary(const char* fn, const float* hx, float* hy, int n, hipStream_t s) {
  VM_RET(vm_args(fn, hx, hx, hy, n));
  HostRun hr(s);
  const float* x;
  float* y;
  VM_RET(hr.in(hx, (size_t)n, &x));
  VM_RET(hr.out(hy, (size_t)n, &y));
  VM_RET(run_unary<OP>(fn, x, y, n, s));
  return hr.finish();
}

template <int OP>
ctStatus_t host_binary(const char* fn, const float* ha, const float* hb, float* hy, int n,
                       hipStream_t s) {
  VM_RET(vm_args(fn, ha, hb, hy, n));
  HostRun hr(s);
  const float *a, *b;
  float* y;
  VM_RET(hr.in(ha, (size_t)n, &a));
  VM_RET(hr.in(hb, (size_t)n, &b));
  VM_RET(hr.out(hy, (size_t)n, &y));
  VM_RET(run_binary<OP>(fn, a, b, y, n, s));
  return hr.finish();
}

}  // namespace
}  // namespace tca

using namespace tca;

#define TCA_UNARY(NAME, OP)                                                                   \
  extern "C" {                                                                                \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_d(const float* x, float* y, int n, hipStream_t s) {     \
    return guarded([&] { return run_unary<OP>("ct_" #NAME "_d", x, y, n, s); });              \
  }                                                                                           \
  CTAPI_EXPORT ctStatus_t ct_##NAME(const float* hx, float* hy, int n, hipStream_t s) {       \
    return guarded([&] { return host_unary<OP>("ct_" #NAME, hx, hy, n, s); });                \
  }                                                                                           \
  }

#define TCA_BINARY(NAME, OP)                                                                   \
  extern "C" {                                                                                 \
  CTAPI_EXPORT ctStatus_t ct_##NAME##_d(const float* a, const float* b, float* y, int n,       \
                                        hipStream_t s) {                                       \
    return guarded([&] { return run_binary<OP>("ct_" #NAME "_d", a, b, y, n, s); });           \
  }                                                                                            \
  CTAPI_EXPORT ctStatus_t ct_##NAME(const float* ha, const float* hb, float* hy, int n,        \
                                    hipStream_t s) {                                           \
    return guarded([&] { return host_binary<OP>("ct_" #NAME, ha, hb, hy, n, s); });            \
  }                                                                                            \
  }

TCA_UNARY(acos, kAcos)
TCA_UNARY(asin, kAsin)
TCA_UNARY(atan, kAtan)
TCA_UNARY(ceil, kCeil)
TCA_UNARY(cos, kCos)
TCA_UNARY(cosh, kCosh)
TCA_UNARY(exp, kExp)
TCA_UNARY(floor, kFloor)
TCA_UNARY(ln, kLn)
TCA_UNARY(log10, kLog10)
TCA_UNARY(sin, kSin)
TCA_UNARY(sinh, kSinh)
TCA_UNARY(sqrt, kSqrt)
TCA_UNARY(tan, kTan)
TCA_UNARY(tanh, kTanh)
TCA_BINARY(add, kAdd)
TCA_BINARY(sub, kSub)
TCA_BINARY(mult, kMult)
TCA_BINARY(div, kDiv)
