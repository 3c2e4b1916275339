// ctx_pack_batch_noalloc_d: the fused multi-indicator pass of
// ctx_pack_batch_d for callers that cannot allocate device memory (code
// under hipGraph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its look-back scratch is a file-scope __device__ array of
// fixed size instead of the per-stream workspace cache: every call on a GPU
// shares that GPU's array, so calls that overlap in time on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/tacuda_amd.h"
#include "host.hpp"

namespace tca {
namespace {

// Three parts (the pass's close and high/low/close recurrence groups, and
// the records of its one-launch form), each a WsHeader plus
// kNoAllocPartBytes of look-back slots and carries.
constexpr size_t kNoAllocPartBytes = size_t(8) << 20;
constexpr size_t kPartStride = (sizeof(WsHeader) + kNoAllocPartBytes + 255) / 256 * 256;
static_assert(kPartStride % 256 == 0, "parts keep 256-byte alignment");

static __device__ __align__(256) unsigned char scratch[kWsParts * kPartStride];

// Zeroes the array and opens epoch 1 in each part's header (look-back
// slots publish the epoch of their launch, so zeroed slots are unpublished).
__global__ void k_noalloc_init(unsigned char* base) {
  const size_t words = kWsParts * kPartStride / 16;
  uint4* w = reinterpret_cast<uint4*>(base);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    w[i] = make_uint4(0u, 0u, 0u, 0u);
}
__global__ void k_noalloc_epoch(unsigned char* base) {
  if (threadIdx.x < kWsParts)
    reinterpret_cast<WsHeader*>(base + threadIdx.x * kPartStride)->epoch = 1u;
}

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The array's device address on the current device, set up once per device:
// hipGetSymbolAddress (the array's name is only a host-side handle), then
// the init kernels on the caller's stream.
ctStatus_t scratch_for(hipStream_t s, unsigned char** out) {
  int dev = 0;
  TCA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices)
    return fail(CT_STATUS_UNSUPPORTED, "pack_noalloc: device ordinal %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    unsigned char* p = nullptr;
    TCA_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&p), HIP_SYMBOL(scratch)));
    hipLaunchKernelGGL(k_noalloc_init, dim3(1024), dim3(256), 0, s, p);
    TCA_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_noalloc_epoch, dim3(1), dim3(64), 0, s, p);
    TCA_HIP(hipGetLastError());
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s != nullptr) TCA_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) TCA_HIP(hipStreamSynchronize(s));
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return CT_STATUS_SUCCESS;
}

}  // namespace
}  // namespace tca

using namespace tca;

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_pack_batch_noalloc_d(const float* h, const float* l, const float* c,
                                                 int seriesLen, int numSeries,
                                                 const ctxPackParams* p, const ctxPackOutputs* o,
                                                 hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    unsigned char* base = nullptr;
    ctStatus_t st = scratch_for(s, &base);
    if (st != CT_STATUS_SUCCESS) return st;
    WsFixed fx{};
    for (int k = 0; k < kWsParts; ++k) fx.part[k] = reinterpret_cast<char*>(base + k * kPartStride);
    fx.bytes = kNoAllocPartBytes;
    WsFixedScope scope(&fx);
    return pack_batch_body(h, l, c, seriesLen, numSeries, p, o, s);
  });
}

}  // extern "C"
