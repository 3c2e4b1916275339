// Status strings, thread-local errors, version, device memory helpers and the
// per-stream workspace cache.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/tacuda_device.h"
#include "host.hpp"
#include "hostio.hpp"
#include "ws_registry.hpp"

namespace tca {

namespace {
thread_local std::string g_error;

std::mutex g_ws_mu;
WsRegistry g_ws;
WsRegistry g_tmp;  // whole-row scratch, keyed like the workspaces
}  // namespace

thread_local const Ragged* t_ragged = nullptr;

bool ragged_geometry(const Ragged* r, int seg, Geometry* g, int* maxSegsPerRow) {
  for (int i = 0; i < 3; ++i) {
    if (kRaggedSegSizes[i] != seg) continue;
    g->offsets = r->dOff;
    g->segStart = r->dSegStart[i];
    g->ld = 0;
    g->len = r->maxLen;
    g->rows = r->rows;
    g->segsPerRow = 0;
    g->totalSegs = r->totalSegs[i];
    *maxSegsPerRow = r->maxSegs[i];
    return true;
  }
  return false;
}

void clear_error() { g_error.clear(); }

ctStatus_t fail(ctStatus_t st, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return st;
}

ctStatus_t hip_fail(hipError_t e, const char* what) {
  (void)hipGetLastError();
  ctStatus_t st = CT_STATUS_DEVICE_ERROR;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) st = CT_STATUS_ALLOC_FAILED;
  if (e == hipErrorInvalidValue || e == hipErrorInvalidDevicePointer) st = CT_STATUS_INVALID_ARGUMENT;
  if (e == hipErrorLaunchFailure || e == hipErrorInvalidDeviceFunction) st = CT_STATUS_KERNEL_FAILED;
  if (e == hipErrorStreamCaptureUnsupported || e == hipErrorStreamCaptureInvalidated)
    st = CT_STATUS_UNSUPPORTED;
  return fail(st, "%s: %s", what, hipGetErrorString(e));
}

thread_local const WsFixed* t_ws_fixed = nullptr;

ctStatus_t workspace_for(hipStream_t stream, size_t treeBytes, Ws* out, int part) {
  if (const WsFixed* f = t_ws_fixed) {
    if (part < 0 || part >= kWsParts) return fail(CT_STATUS_INVALID_ARGUMENT, "workspace part");
    if (treeBytes > f->bytes)
      return fail(CT_STATUS_INVALID_ARGUMENT,
                  "this call needs %zu bytes of look-back scratch; the allocation-free entry "
                  "point holds %zu per part",
                  treeBytes, f->bytes);
    out->hdr = reinterpret_cast<WsHeader*>(f->part[part]);
    out->area = reinterpret_cast<double*>(f->part[part] + sizeof(WsHeader));
    return CT_STATUS_SUCCESS;
  }
  int dev = 0;
  TCA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  const WsRegistry::Key key{dev, stream, part};
  void* base = g_ws.find(key, treeBytes);
  if (base == nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr) TCA_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(CT_STATUS_UNSUPPORTED,
                  "look-back workspace of %zu bytes is not reserved for this stream; run the "
                  "call once eagerly (or call ctx_workspace_reserve) before capturing",
                  treeBytes);
    // The buffer this replaces stays allocated (retired): a graph captured
    // on this stream may still hold its address.
    const size_t want = g_ws.grow_size(key, treeBytes);
    const size_t bytes = sizeof(WsHeader) + want;
    TCA_HIP(hipMalloc(&base, bytes));
    hipError_t e = hipMemsetAsync(base, 0, bytes, stream);
    WsHeader h{};
    h.epoch = 1;
    if (e == hipSuccess)
      e = hipMemcpyAsync(base, &h, sizeof h, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
      (void)hipFree(base);
      return hip_fail(e, "workspace init");
    }
    g_ws.install(key, base, want);
  }
  out->hdr = static_cast<WsHeader*>(base);
  out->area = reinterpret_cast<double*>(static_cast<char*>(base) + sizeof(WsHeader));
  return CT_STATUS_SUCCESS;
}

ctStatus_t temp_for(hipStream_t stream, size_t bytes, int slot, void** out) {
  *out = nullptr;
  if (slot < 0 || slot >= kTmpSlots) return fail(CT_STATUS_INVALID_ARGUMENT, "temp slot");
  if (t_ws_fixed)
    return fail(CT_STATUS_UNSUPPORTED, "this call needs row scratch the allocation-free entry "
                                       "point does not provide");
  int dev = 0;
  TCA_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_ws_mu);
  const WsRegistry::Key key{dev, stream, slot};
  void* base = g_tmp.find(key, bytes);
  if (base == nullptr) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (stream != nullptr) TCA_HIP(hipStreamIsCapturing(stream, &cs));
    if (cs != hipStreamCaptureStatusNone)
      return fail(CT_STATUS_UNSUPPORTED,
                  "row scratch of %zu bytes is not reserved for this stream; run the call once "
                  "eagerly before capturing",
                  bytes);
    const size_t want = g_tmp.grow_size(key, bytes);
    TCA_HIP(hipMalloc(&base, want));
    g_tmp.install(key, base, want);
  }
  *out = base;
  return CT_STATUS_SUCCESS;
}

void workspace_release(hipStream_t stream) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return;
  std::vector<WsRegistry::Buf> bufs;
  {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    bufs = g_ws.release_stream(dev, stream);
    for (const auto& b : g_tmp.release_stream(dev, stream)) bufs.push_back(b);
  }
  if (bufs.empty()) return;
  (void)hipStreamSynchronize(stream);
  for (const auto& b : bufs) (void)hipFree(b.base);
}

void workspace_release_all() {
  std::vector<WsRegistry::Buf> bufs;
  {
    std::lock_guard<std::mutex> lk(g_ws_mu);
    bufs = g_ws.release_all();
    for (const auto& b : g_tmp.release_all()) bufs.push_back(b);
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (const auto& b : bufs) {
    (void)hipSetDevice(b.device);
    (void)hipDeviceSynchronize();
    (void)hipFree(b.base);
  }
  (void)hipSetDevice(cur);
}

size_t workspace_allocations() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  return g_ws.allocations() + g_tmp.allocations();
}

}  // namespace tca

using namespace tca;

extern "C" {

CTAPI_EXPORT const char* ct_status_string(ctStatus_t s) {
  switch (s) {
    case CT_STATUS_SUCCESS: return "success";
    case CT_STATUS_ALLOC_FAILED: return "allocation failed";
    case CT_STATUS_COPY_FAILED: return "copy failed";
    case CT_STATUS_KERNEL_FAILED: return "kernel failed";
    case CT_STATUS_INVALID_ARGUMENT: return "invalid argument";
    case CT_STATUS_DEVICE_ERROR: return "device error";
    case CT_STATUS_UNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

CTAPI_EXPORT const char* ct_get_last_error(void) { return g_error.c_str(); }

CTAPI_EXPORT const char* ct_version(void) { return CT_VERSION_STRING; }

CTAPI_EXPORT float* ct_device_alloc(int count) {
  clear_error();
  if (count <= 0) {
    fail(CT_STATUS_INVALID_ARGUMENT, "ct_device_alloc: count must be positive");
    return nullptr;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (size_t)count * sizeof(float));
  if (e != hipSuccess) {
    hip_fail(e, "hipMalloc");
    return nullptr;
  }
  return static_cast<float*>(p);
}

CTAPI_EXPORT void* ct_device_alloc_ex(int count, ctDataType_t t) {
  clear_error();
  size_t es = t == CT_DTYPE_F16 ? 2 : t == CT_DTYPE_F32 ? 4 : t == CT_DTYPE_F64 ? 8 : 0;
  if (count <= 0 || es == 0) {
    fail(CT_STATUS_INVALID_ARGUMENT, "ct_device_alloc_ex: bad count or data type");
    return nullptr;
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, (size_t)count * es);
  if (e != hipSuccess) {
    hip_fail(e, "hipMalloc");
    return nullptr;
  }
  return p;
}

CTAPI_EXPORT void ct_device_free(float* p) {
  if (p) (void)hipFree(p);
}
CTAPI_EXPORT void ct_device_free_ex(void* p) {
  if (p) (void)hipFree(p);
}

CTAPI_EXPORT ctStatus_t ct_device_upload(float* d, const float* h, int count, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!d || !h || count <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ct_device_upload: bad arguments");
    hipError_t e = hipMemcpyAsync(d, h, (size_t)count * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) {
      hip_fail(e, "hipMemcpyAsync");
      return CT_STATUS_COPY_FAILED;
    }
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ct_device_download(float* h, const float* d, int count, hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    if (!d || !h || count <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ct_device_download: bad arguments");
    hipError_t e = hipMemcpyAsync(h, d, (size_t)count * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      hip_fail(e, "hipMemcpyAsync");
      return CT_STATUS_COPY_FAILED;
    }
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ct_device_sync(hipStream_t s) {
  return guarded([&]() -> ctStatus_t {
    TCA_HIP(hipStreamSynchronize(s));
    return CT_STATUS_SUCCESS;
  });
}

// -------------------------------------------------------- kernel timer --
namespace tca {
namespace {
struct KRec {
  const char* name;
  hipEvent_t a, b;
};
std::atomic<bool> g_kt_on{false};
std::mutex g_kt_mu;
std::vector<KRec> g_kt;

bool kt_capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &st) != hipSuccess || st != hipStreamCaptureStatusNone;
}
}  // namespace

KTime::KTime(hipStream_t stream, const char* name) : s_(stream), name_(name) {
  // the allocation-free entry point creates no events either
  if (!g_kt_on.load(std::memory_order_relaxed) || t_ws_fixed || kt_capturing(s_)) return;
  if (hipEventCreate(&a_) != hipSuccess) {
    a_ = nullptr;
    return;
  }
  if (hipEventRecord(a_, s_) != hipSuccess) {
    (void)hipEventDestroy(a_);
    a_ = nullptr;
  }
}

KTime::~KTime() {
  if (!a_) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) != hipSuccess || hipEventRecord(b, s_) != hipSuccess) {
    if (b) (void)hipEventDestroy(b);
    (void)hipEventDestroy(a_);
    return;
  }
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt.push_back(KRec{name_, a_, b});
}
}  // namespace tca

CTAPI_EXPORT ctStatus_t ctx_kernel_timing(int enable) {
  tca::g_kt_on.store(enable != 0);
  return CT_STATUS_SUCCESS;
}

CTAPI_EXPORT ctStatus_t ctx_kernel_times(char* names, int namesLen, float* ms, int maxN,
                                         int* count) {
  return guarded([&]() -> ctStatus_t {
    if (!count || maxN < 0 || (maxN > 0 && (!ms || !names || namesLen <= 0)))
      return fail(CT_STATUS_INVALID_ARGUMENT, "kernel_times: bad arguments");
    std::vector<tca::KRec> recs;
    {
      std::lock_guard<std::mutex> lk(tca::g_kt_mu);
      recs.swap(tca::g_kt);
    }
    int n = 0;
    size_t used = 0;
    ctStatus_t st = CT_STATUS_SUCCESS;
    for (const auto& r : recs) {
      float t = 0.0f;
      hipError_t e = hipEventSynchronize(r.b);
      if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
      if (e != hipSuccess && st == CT_STATUS_SUCCESS) st = hip_fail(e, "kernel_times");
      const size_t len = std::strlen(r.name) + 1;
      if (st == CT_STATUS_SUCCESS && n < maxN && used + len <= (size_t)namesLen) {
        std::memcpy(names + used, r.name, len - 1);
        names[used + len - 1] = '\n';
        used += len;
        ms[n++] = t;
      }
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    if (maxN > 0) names[used > 0 ? used - 1 : 0] = '\0';
    *count = n;
    return st;
  });
}

// The reference's buffer-pool controls (tacuda.h:824-827) cover this
// build's two caches: the host calls' device buffers and pinned staging, and
// the per-stream look-back workspaces.
CTAPI_EXPORT void ct_device_buffer_pool_cleanup(void) {
  workspace_release_all();
  pool_cleanup();
}
CTAPI_EXPORT size_t ct_device_buffer_pool_allocation_count(void) {
  return workspace_allocations() + pool_allocations();
}
CTAPI_EXPORT size_t ct_device_buffer_pool_set_max_cache_bytes(size_t bytes) {
  return pool_set_max_cache(bytes);
}

}  // extern "C"
