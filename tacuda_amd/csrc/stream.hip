// Streaming API: single-series streams on the host, batch (portfolio)
// streams on the host or on the GPU with the states resident in HBM
// (reference src/api_stream.cpp:38-91, src/api_stream_batch.cu:89-330 for
// the surface; the state machines are in stream_states.hpp).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/tacuda_stream.h"
#include "host.hpp"
#include "stream_states.hpp"
#ifdef TCA_STREAM_FAULT_INJECTION
#include <atomic>

#include "stream_faults.h"
#endif

namespace tca {
namespace {

// Controlled failures for the state-transition tests: only the
// fault-injection build (libtacuda_amd_faults.so) has the hook; elsewhere
// FAULT() is a constant false.
#ifdef TCA_STREAM_FAULT_INJECTION
std::atomic<int> g_next_fault{0};
bool consume_fault(int site) {
  int expected = site;
  return g_next_fault.compare_exchange_strong(expected, 0);
}
#define FAULT(site) consume_fault(site)
#else
#define FAULT(site) false
#endif

using st::State;

struct Single {
  State s;
  std::vector<float> ring;
};

struct Batch {
  State proto;
  int rows = 0;
  int ringSize = 0;
  int exec = CT_STREAM_EXEC_AUTO;
  long long bars = 0;
  std::vector<State> hs;
  std::vector<float> hring;
  bool onDevice = false;      // the device copy is the current one
  State* ds = nullptr;
  float* dring = nullptr;
  hipStream_t stream = nullptr;    // private stream of host-array GPU appends
  hipStream_t lastUser = nullptr;  // stream of the last device-pointer append
  bool userPending = false;
  float* pin = nullptr;  // pinned staging, 7 planes of rows*n floats
  size_t pinFloats = 0;
  float* dstage = nullptr;
  size_t dstageFloats = 0;
  // A GPU append that failed from its kernel launch on may have advanced the
  // device state past what the caller received: the handle then refuses
  // work until reset (reference api_stream_batch.cu:187-189, 219-231).
  bool failed = false;
};

ctStatus_t refuse_failed() {
  return fail(CT_STATUS_DEVICE_ERROR,
              "stream batch: an earlier GPU append failed after its launch and left the state "
              "undefined; reset the handle");
}
ctStatus_t mark_failed(Batch* b, ctStatus_t st, const char* what) {
  b->failed = true;
  return fail(st, "stream batch: %s", what);
}

// Opaque tokens: increasing integers, never reused while loaded.
std::mutex g_reg_mu;
uintptr_t g_next_token = 0x10;
std::map<uintptr_t, std::unique_ptr<Single>> g_single;
std::map<uintptr_t, std::unique_ptr<Batch>> g_batch;

uintptr_t new_token() {
  g_next_token += 0x10;
  return g_next_token;
}

Single* find_single(const void* h) {
  std::lock_guard<std::mutex> lk(g_reg_mu);
  auto it = g_single.find(reinterpret_cast<uintptr_t>(h));
  return it == g_single.end() ? nullptr : it->second.get();
}
Batch* find_batch(const void* h) {
  std::lock_guard<std::mutex> lk(g_reg_mu);
  auto it = g_batch.find(reinterpret_cast<uintptr_t>(h));
  return it == g_batch.end() ? nullptr : it->second.get();
}

State make_state(int kind, int p1, int p2, int p3, float up, float dn) {
  State s;
  std::memset(&s, 0, sizeof s);
  s.kind = kind;
  s.p1 = p1;
  s.p2 = p2;
  s.p3 = p3;
  s.up = up;
  s.dn = dn;
  st::reset(s);
  return s;
}

// Parameter checks of the create calls (reference api_stream.cpp): periods
// >= 1 (MACD periods swapped when reversed, as ct_macd does).
ctStatus_t make_checked(int kind, int p1, int p2, int p3, float up, float dn, State* out) {
  switch (kind) {
    case st::kObv: case st::kAd:
      break;
    case st::kMacd:
      if (p1 <= 0 || p2 <= 0 || p3 <= 0)
        return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid MACD periods");
      if (p2 < p1) std::swap(p1, p2);
      break;
    default:
      if (p1 <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid period");
      if (p1 > 100000) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: period above 100000");
  }
  *out = make_state(kind, p1, p2, p3, up, dn);
  return CT_STATUS_SUCCESS;
}

// ------------------------------------------------------------ single --
ctStatus_t single_create(int kind, int p1, int p2, int p3, float up, float dn, ctStream_t** out) {
  if (!out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null output handle");
  *out = nullptr;
  State s;
  ctStatus_t r = make_checked(kind, p1, p2, p3, up, dn, &s);
  if (r != CT_STATUS_SUCCESS) return r;
  auto obj = std::make_unique<Single>();
  obj->s = s;
  obj->ring.assign(st::ring_size(s), 0.0f);
  std::lock_guard<std::mutex> lk(g_reg_mu);
  const uintptr_t t = new_token();
  g_single[t] = std::move(obj);
  *out = reinterpret_cast<ctStream_t*>(t);
  return CT_STATUS_SUCCESS;
}

ctStatus_t single_append(ctStream_t* h, int kind, const float* hi, const float* lo,
                         const float* cl, const float* vo, int n, float* o0, float* o1,
                         float* o2) {
  Single* s = find_single(h);
  if (!s) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid or destroyed handle");
  if (s->s.kind != kind) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: handle is another family");
  if (n < 0) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: negative bar count");
  if (n == 0) return CT_STATUS_SUCCESS;
  if (!o0 || (st::outputs(s->s) == 3 && (!o1 || !o2)))
    return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null output");
  for (int i = 0; i < n; ++i) {
    st::Bar b{hi ? hi[i] : 0.0f, lo ? lo[i] : 0.0f, cl ? cl[i] : 0.0f, vo ? vo[i] : 0.0f};
    float out[3];
    st::step(s->s, s->ring.data(), b, out);
    o0[i] = out[0];
    if (o1) o1[i] = out[1];
    if (o2) o2[i] = out[2];
  }
  return CT_STATUS_SUCCESS;
}

// ------------------------------------------------------------- batch --
__global__ __launch_bounds__(256) void k_stream_batch(State* S, float* rings, int ringSize,
                                                      int rows, const float* h, const float* l,
                                                      const float* c, const float* v, int n,
                                                      float* o0, float* o1, float* o2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= rows) return;
  State x = S[s];
  float* ring = rings ? rings + (size_t)s * ringSize : nullptr;
  for (int i = 0; i < n; ++i) {
    const size_t idx = (size_t)s * n + i;
    st::Bar b{h ? h[idx] : 0.0f, l ? l[idx] : 0.0f, c ? c[idx] : 0.0f, v ? v[idx] : 0.0f};
    float out[3];
    st::step(x, ring, b, out);
    o0[idx] = out[0];
    if (o1) o1[idx] = out[1];
    if (o2) o2[idx] = out[2];
  }
  S[s] = x;
}

ctStatus_t wait_user(Batch* b) {
  if (b->userPending) {
    TCA_HIP(hipStreamSynchronize(b->lastUser));
    b->userPending = false;
  }
  return CT_STATUS_SUCCESS;
}

// Failures here leave the host copy current and untouched: the caller may
// retry.
ctStatus_t to_device(Batch* b) {
  if (b->onDevice) return CT_STATUS_SUCCESS;
  if (!b->ds) {
    TCA_HIP(hipMalloc(&b->ds, sizeof(State) * b->rows));
    if (b->ringSize) TCA_HIP(hipMalloc(&b->dring, sizeof(float) * b->rows * (size_t)b->ringSize));
  }
  if (!b->stream) TCA_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  if (FAULT(CT_STREAM_BATCH_FAULT_STATE_UPLOAD))
    return fail(CT_STATUS_COPY_FAILED, "stream batch: state upload failed");
  TCA_HIP(hipMemcpy(b->ds, b->hs.data(), sizeof(State) * b->rows, hipMemcpyHostToDevice));
  if (b->ringSize)
    TCA_HIP(hipMemcpy(b->dring, b->hring.data(), sizeof(float) * b->hring.size(),
                      hipMemcpyHostToDevice));
  b->onDevice = true;
  return CT_STATUS_SUCCESS;
}

// A failed state download leaves neither copy trustworthy (the host copy may
// be partly overwritten): the handle refuses work until reset.
ctStatus_t to_host(Batch* b) {
  if (!b->onDevice) return CT_STATUS_SUCCESS;
  ctStatus_t r = wait_user(b);
  if (r != CT_STATUS_SUCCESS) return r;
  hipError_t e = FAULT(CT_STREAM_BATCH_FAULT_STATE_DOWNLOAD)
                     ? hipErrorUnknown
                     : hipMemcpy(b->hs.data(), b->ds, sizeof(State) * b->rows,
                                 hipMemcpyDeviceToHost);
  if (e == hipSuccess && b->ringSize)
    e = hipMemcpy(b->hring.data(), b->dring, sizeof(float) * b->hring.size(),
                  hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return mark_failed(b, CT_STATUS_COPY_FAILED, "state download failed");
  }
  b->onDevice = false;
  return CT_STATUS_SUCCESS;
}

bool use_gpu(const Batch* b) {
  if (b->exec == CT_STREAM_EXEC_GPU) return true;
  if (b->exec == CT_STREAM_EXEC_CPU) return false;
  return b->rows >= CT_STREAM_AUTO_GPU_SERIES;
}

void batch_free(Batch* b) {
  if (b->userPending) (void)hipStreamSynchronize(b->lastUser);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->ds) (void)hipFree(b->ds);
  if (b->dring) (void)hipFree(b->dring);
  if (b->pin) (void)hipHostFree(b->pin);
  if (b->dstage) (void)hipFree(b->dstage);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

ctStatus_t batch_create(int kind, int p1, int p2, int p3, float up, float dn, int rows,
                        ctStreamBatch_t** out) {
  if (!out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null output handle");
  *out = nullptr;
  if (rows <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: numSeries must be positive");
  State s;
  ctStatus_t r = make_checked(kind, p1, p2, p3, up, dn, &s);
  if (r != CT_STATUS_SUCCESS) return r;
  auto obj = std::make_unique<Batch>();
  obj->proto = s;
  obj->rows = rows;
  obj->ringSize = st::ring_size(s);
  obj->hs.assign(rows, s);
  obj->hring.assign((size_t)rows * obj->ringSize, 0.0f);
  std::lock_guard<std::mutex> lk(g_reg_mu);
  const uintptr_t t = new_token();
  g_batch[t] = std::move(obj);
  *out = reinterpret_cast<ctStreamBatch_t*>(t);
  return CT_STATUS_SUCCESS;
}

// Host-array append: CPU loop, or pinned staging + one kernel + one sync.
ctStatus_t batch_append(ctStreamBatch_t* h, int kind, const float* hi, const float* lo,
                        const float* cl, const float* vo, int n, float* o0, float* o1,
                        float* o2) {
  Batch* b = find_batch(h);
  if (!b) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid or destroyed handle");
  if (b->proto.kind != kind)
    return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: handle is another family");
  if (n < 0) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: negative bar count");
  if (n == 0) return CT_STATUS_SUCCESS;
  const int nout = st::outputs(b->proto);
  if (!o0 || (nout == 3 && (!o1 || !o2)))
    return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null output");
  const size_t tot = (size_t)b->rows * n;
  if (b->failed) return refuse_failed();
  if (!use_gpu(b)) {
    ctStatus_t r = to_host(b);
    if (r != CT_STATUS_SUCCESS) return r;
    for (int s = 0; s < b->rows; ++s) {
      State& x = b->hs[s];
      float* ring = b->ringSize ? b->hring.data() + (size_t)s * b->ringSize : nullptr;
      for (int i = 0; i < n; ++i) {
        const size_t idx = (size_t)s * n + i;
        st::Bar bar{hi ? hi[idx] : 0.0f, lo ? lo[idx] : 0.0f, cl ? cl[idx] : 0.0f,
                    vo ? vo[idx] : 0.0f};
        float out[3];
        st::step(x, ring, bar, out);
        o0[idx] = out[0];
        if (nout == 3) {
          o1[idx] = out[1];
          o2[idx] = out[2];
        }
      }
    }
    b->bars += n;
    return CT_STATUS_SUCCESS;
  }
  ctStatus_t r = wait_user(b);
  if (r != CT_STATUS_SUCCESS) return r;
  r = to_device(b);
  if (r != CT_STATUS_SUCCESS) return r;
  // planes: 4 inputs then 3 outputs
  const size_t need = 7 * tot;
  if (FAULT(CT_STREAM_BATCH_FAULT_IO_ALLOC))
    return fail(CT_STATUS_ALLOC_FAILED, "stream batch: staging allocation failed");
  if (b->pinFloats < need) {
    if (b->pin) TCA_HIP(hipHostFree(b->pin));
    b->pin = nullptr;
    TCA_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->pin), need * sizeof(float)));
    b->pinFloats = need;
  }
  if (b->dstageFloats < need) {
    if (b->dstage) TCA_HIP(hipFree(b->dstage));
    b->dstage = nullptr;
    TCA_HIP(hipMalloc(&b->dstage, need * sizeof(float)));
    b->dstageFloats = need;
  }
  const float* ins[4] = {hi, lo, cl, vo};
  const float* dins[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < 4; ++q) {
    if (!ins[q]) continue;
    std::memcpy(b->pin + q * tot, ins[q], tot * sizeof(float));
    if (FAULT(CT_STREAM_BATCH_FAULT_INPUT_UPLOAD)) {
      // queued uploads of other inputs only touch staging
      (void)hipStreamSynchronize(b->stream);
      return fail(CT_STATUS_COPY_FAILED, "stream batch: input upload failed");
    }
    TCA_HIP(hipMemcpyAsync(b->dstage + q * tot, b->pin + q * tot, tot * sizeof(float),
                           hipMemcpyHostToDevice, b->stream));
    dins[q] = b->dstage + q * tot;
  }
  float* douts[3] = {b->dstage + 4 * tot, nout == 3 ? b->dstage + 5 * tot : nullptr,
                     nout == 3 ? b->dstage + 6 * tot : nullptr};
  const int blocks = (b->rows + 255) / 256;
  // From the launch on, the device state may advance past what the caller
  // receives: any failure marks the handle.
  hipLaunchKernelGGL(k_stream_batch, dim3(blocks), dim3(256), 0, b->stream, b->ds, b->dring,
                     b->ringSize, b->rows, dins[0], dins[1], dins[2], dins[3], n, douts[0],
                     douts[1], douts[2]);
  if (hipGetLastError() != hipSuccess)
    return mark_failed(b, CT_STATUS_KERNEL_FAILED, "kernel launch failed");
  if (hipStreamSynchronize(b->stream) != hipSuccess || FAULT(CT_STREAM_BATCH_FAULT_KERNEL)) {
    (void)hipGetLastError();
    return mark_failed(b, CT_STATUS_KERNEL_FAILED, "kernel execution failed");
  }
  if (FAULT(CT_STREAM_BATCH_FAULT_OUTPUT_DOWNLOAD) ||
      hipMemcpyAsync(b->pin + 4 * tot, douts[0], nout * tot * sizeof(float),
                     hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    (void)hipGetLastError();
    return mark_failed(b, CT_STATUS_COPY_FAILED, "output download failed");
  }
  float* outs[3] = {o0, o1, o2};
  for (int q = 0; q < nout; ++q)
    std::memcpy(outs[q], b->pin + (4 + q) * tot, tot * sizeof(float));
  b->bars += n;
  return CT_STATUS_SUCCESS;
}

// Inputs each family reads: bit 0 high, 1 low, 2 close, 3 volume.
unsigned needs(int kind) {
  switch (kind) {
    case st::kAtr: case st::kNatr: return 0x7;
    case st::kObv: return 0xc;
    case st::kAd: return 0xf;
    default: return 0x4;
  }
}

}  // namespace
}  // namespace tca

using namespace tca;

#define S_CREATE1(fam, KIND)                                                        \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_create(int period, ctStream_t** out) {  \
    return guarded([&] { return single_create(KIND, period, 0, 0, 0, 0, out); });  \
  }                                                                                 \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_batch_create(int period, int rows,      \
                                                         ctStreamBatch_t** out) {   \
    return guarded([&] { return batch_create(KIND, period, 0, 0, 0, 0, rows, out); }); \
  }

extern "C" {

S_CREATE1(sma, st::kSma)
S_CREATE1(ema, st::kEma)
S_CREATE1(wma, st::kWma)
S_CREATE1(rsi, st::kRsi)
S_CREATE1(atr, st::kAtr)
S_CREATE1(natr, st::kNatr)

CTAPI_EXPORT ctStatus_t ct_obv_stream_create(ctStream_t** out) {
  return guarded([&] { return single_create(st::kObv, 0, 0, 0, 0, 0, out); });
}
CTAPI_EXPORT ctStatus_t ct_ad_stream_create(ctStream_t** out) {
  return guarded([&] { return single_create(st::kAd, 0, 0, 0, 0, 0, out); });
}
CTAPI_EXPORT ctStatus_t ct_macd_stream_create(int f, int s, int g, ctStream_t** out) {
  return guarded([&] { return single_create(st::kMacd, f, s, g, 0, 0, out); });
}
CTAPI_EXPORT ctStatus_t ct_bbands_stream_create(int p, float up, float dn, ctStream_t** out) {
  return guarded([&] { return single_create(st::kBbands, p, 0, 0, up, dn, out); });
}
CTAPI_EXPORT ctStatus_t ct_obv_stream_batch_create(int rows, ctStreamBatch_t** out) {
  return guarded([&] { return batch_create(st::kObv, 0, 0, 0, 0, 0, rows, out); });
}
CTAPI_EXPORT ctStatus_t ct_ad_stream_batch_create(int rows, ctStreamBatch_t** out) {
  return guarded([&] { return batch_create(st::kAd, 0, 0, 0, 0, 0, rows, out); });
}
CTAPI_EXPORT ctStatus_t ct_macd_stream_batch_create(int f, int s, int g, int rows,
                                                    ctStreamBatch_t** out) {
  return guarded([&] { return batch_create(st::kMacd, f, s, g, 0, 0, rows, out); });
}
CTAPI_EXPORT ctStatus_t ct_bbands_stream_batch_create(int p, float up, float dn, int rows,
                                                      ctStreamBatch_t** out) {
  return guarded([&] { return batch_create(st::kBbands, p, 0, 0, up, dn, rows, out); });
}

// ------------------------------------------------------- single appends --
#define S_APPEND_C(fam, KIND)                                                              \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_append(ctStream_t* h, const float* c, int n,   \
                                                   float* out) {                           \
    return guarded([&]() -> ctStatus_t {                                                   \
      if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");       \
      return single_append(h, KIND, nullptr, nullptr, c, nullptr, n, out, nullptr, nullptr); \
    });                                                                                    \
  }
S_APPEND_C(sma, st::kSma)
S_APPEND_C(ema, st::kEma)
S_APPEND_C(wma, st::kWma)
S_APPEND_C(rsi, st::kRsi)
#define S_APPEND_HLC(fam, KIND)                                                            \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_append(ctStream_t* st_, const float* h,        \
                                                   const float* l, const float* c, int n,  \
                                                   float* out) {                           \
    return guarded([&]() -> ctStatus_t {                                                   \
      if (n > 0 && (!h || !l || !c))                                                       \
        return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");                     \
      return single_append(st_, KIND, h, l, c, nullptr, n, out, nullptr, nullptr);         \
    });                                                                                    \
  }
S_APPEND_HLC(atr, st::kAtr)
S_APPEND_HLC(natr, st::kNatr)

CTAPI_EXPORT ctStatus_t ct_obv_stream_append(ctStream_t* h, const float* c, const float* v, int n,
                                             float* out) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && (!c || !v)) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");
    return single_append(h, st::kObv, nullptr, nullptr, c, v, n, out, nullptr, nullptr);
  });
}
CTAPI_EXPORT ctStatus_t ct_ad_stream_append(ctStream_t* st_, const float* h, const float* l,
                                            const float* c, const float* v, int n, float* out) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && (!h || !l || !c || !v))
      return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");
    return single_append(st_, st::kAd, h, l, c, v, n, out, nullptr, nullptr);
  });
}
CTAPI_EXPORT ctStatus_t ct_macd_stream_append(ctStream_t* h, const float* c, int n, float* m,
                                              float* s, float* hist) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");
    return single_append(h, st::kMacd, nullptr, nullptr, c, nullptr, n, m, s, hist);
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands_stream_append(ctStream_t* h, const float* c, int n, float* u,
                                                float* m, float* l) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: null input");
    return single_append(h, st::kBbands, nullptr, nullptr, c, nullptr, n, u, m, l);
  });
}

// -------------------------------------------------------- batch appends --
#define B_APPEND_C(fam, KIND)                                                                \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_batch_append(ctStreamBatch_t* b, const float* c, \
                                                         int n, float* out) {                \
    return guarded([&]() -> ctStatus_t {                                                     \
      if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");   \
      return batch_append(b, KIND, nullptr, nullptr, c, nullptr, n, out, nullptr, nullptr);  \
    });                                                                                      \
  }
B_APPEND_C(sma, st::kSma)
B_APPEND_C(ema, st::kEma)
B_APPEND_C(wma, st::kWma)
B_APPEND_C(rsi, st::kRsi)
#define B_APPEND_HLC(fam, KIND)                                                              \
  CTAPI_EXPORT ctStatus_t ct_##fam##_stream_batch_append(ctStreamBatch_t* b, const float* h, \
                                                         const float* l, const float* c,     \
                                                         int n, float* out) {                \
    return guarded([&]() -> ctStatus_t {                                                     \
      if (n > 0 && (!h || !l || !c))                                                         \
        return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");                 \
      return batch_append(b, KIND, h, l, c, nullptr, n, out, nullptr, nullptr);              \
    });                                                                                      \
  }
B_APPEND_HLC(atr, st::kAtr)
B_APPEND_HLC(natr, st::kNatr)

CTAPI_EXPORT ctStatus_t ct_obv_stream_batch_append(ctStreamBatch_t* b, const float* c,
                                                   const float* v, int n, float* out) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && (!c || !v)) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");
    return batch_append(b, st::kObv, nullptr, nullptr, c, v, n, out, nullptr, nullptr);
  });
}
CTAPI_EXPORT ctStatus_t ct_ad_stream_batch_append(ctStreamBatch_t* b, const float* h,
                                                  const float* l, const float* c, const float* v,
                                                  int n, float* out) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && (!h || !l || !c || !v))
      return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");
    return batch_append(b, st::kAd, h, l, c, v, n, out, nullptr, nullptr);
  });
}
CTAPI_EXPORT ctStatus_t ct_macd_stream_batch_append(ctStreamBatch_t* b, const float* c, int n,
                                                    float* m, float* s, float* hist) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");
    return batch_append(b, st::kMacd, nullptr, nullptr, c, nullptr, n, m, s, hist);
  });
}
CTAPI_EXPORT ctStatus_t ct_bbands_stream_batch_append(ctStreamBatch_t* b, const float* c, int n,
                                                      float* u, float* m, float* l) {
  return guarded([&]() -> ctStatus_t {
    if (n > 0 && !c) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");
    return batch_append(b, st::kBbands, nullptr, nullptr, c, nullptr, n, u, m, l);
  });
}

CTAPI_EXPORT ctStatus_t ctx_stream_batch_append_d(ctStreamBatch_t* h, const float* dh,
                                                  const float* dl, const float* dc,
                                                  const float* dv, int n, float* o0, float* o1,
                                                  float* o2, hipStream_t stream) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid or destroyed handle");
    if (n < 0) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: negative bar count");
    if (n == 0) return CT_STATUS_SUCCESS;
    const unsigned need = needs(b->proto.kind);
    const float* in[4] = {dh, dl, dc, dv};
    for (int q = 0; q < 4; ++q)
      if ((need >> q & 1) && !in[q])
        return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null input");
    const int nout = st::outputs(b->proto);
    if (!o0 || (nout == 3 && (!o1 || !o2)))
      return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: null output");
    if (b->failed) return refuse_failed();
    if (b->userPending && b->lastUser != stream) {
      ctStatus_t r = wait_user(b);
      if (r != CT_STATUS_SUCCESS) return r;
    }
    b->exec = CT_STREAM_EXEC_GPU;
    ctStatus_t r = to_device(b);
    if (r != CT_STATUS_SUCCESS) return r;
    const int blocks = (b->rows + 255) / 256;
    hipLaunchKernelGGL(k_stream_batch, dim3(blocks), dim3(256), 0, stream, b->ds, b->dring,
                       b->ringSize, b->rows, (need & 1) ? dh : nullptr, (need & 2) ? dl : nullptr,
                       (need & 4) ? dc : nullptr, (need & 8) ? dv : nullptr, n, o0,
                       nout == 3 ? o1 : nullptr, nout == 3 ? o2 : nullptr);
    if (hipGetLastError() != hipSuccess)
      return mark_failed(b, CT_STATUS_KERNEL_FAILED, "kernel launch failed");
    b->lastUser = stream;
    b->userPending = true;
    b->bars += n;
    return CT_STATUS_SUCCESS;
  });
}

// -------------------------------------------------- lifetime / queries --
CTAPI_EXPORT ctStatus_t ct_stream_destroy(ctStream_t* h) {
  return guarded([&]() -> ctStatus_t {
    if (!h) return CT_STATUS_SUCCESS;
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_single.find(reinterpret_cast<uintptr_t>(h));
    if (it == g_single.end())
      return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid or destroyed handle");
    g_single.erase(it);
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_reset(ctStream_t* h) {
  return guarded([&]() -> ctStatus_t {
    Single* s = find_single(h);
    if (!s) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid or destroyed handle");
    st::reset(s->s);
    std::fill(s->ring.begin(), s->ring.end(), 0.0f);
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_lookback(const ctStream_t* h, int* out) {
  return guarded([&]() -> ctStatus_t {
    Single* s = find_single(h);
    if (!s || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid handle or output");
    *out = st::lookback(s->s);
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_bars(const ctStream_t* h, long long* out) {
  return guarded([&]() -> ctStatus_t {
    Single* s = find_single(h);
    if (!s || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid handle or output");
    *out = s->s.n;
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_outputs(const ctStream_t* h, int* out) {
  return guarded([&]() -> ctStatus_t {
    Single* s = find_single(h);
    if (!s || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream: invalid handle or output");
    *out = st::outputs(s->s);
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ct_stream_batch_destroy(ctStreamBatch_t* h) {
  return guarded([&]() -> ctStatus_t {
    if (!h) return CT_STATUS_SUCCESS;
    std::unique_ptr<Batch> obj;
    {
      std::lock_guard<std::mutex> lk(g_reg_mu);
      auto it = g_batch.find(reinterpret_cast<uintptr_t>(h));
      if (it == g_batch.end())
        return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid or destroyed handle");
      obj = std::move(it->second);
      g_batch.erase(it);
    }
    batch_free(obj.get());
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_reset(ctStreamBatch_t* h) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid or destroyed handle");
    ctStatus_t r = wait_user(b);
    if (r != CT_STATUS_SUCCESS) return r;
    if (b->stream) TCA_HIP(hipStreamSynchronize(b->stream));
    std::fill(b->hs.begin(), b->hs.end(), b->proto);
    std::fill(b->hring.begin(), b->hring.end(), 0.0f);
    b->onDevice = false;
    b->failed = false;
    b->bars = 0;
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_lookback(const ctStreamBatch_t* h, int* out) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid handle or output");
    *out = st::lookback(b->proto);
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_series(const ctStreamBatch_t* h, int* out) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid handle or output");
    *out = b->rows;
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_outputs(const ctStreamBatch_t* h, int* out) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid handle or output");
    *out = st::outputs(b->proto);
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_bars(const ctStreamBatch_t* h, long long* out) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid handle or output");
    *out = b->bars;
    return CT_STATUS_SUCCESS;
  });
}
CTAPI_EXPORT ctStatus_t ct_stream_batch_set_exec(ctStreamBatch_t* h, ctStreamExec_t exec) {
  return guarded([&]() -> ctStatus_t {
    Batch* b = find_batch(h);
    if (!b) return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: invalid or destroyed handle");
    if (exec != CT_STREAM_EXEC_AUTO && exec != CT_STREAM_EXEC_CPU && exec != CT_STREAM_EXEC_GPU)
      return fail(CT_STATUS_INVALID_ARGUMENT, "stream batch: unknown execution mode");
    // the states move with the next append (reference
    // api_stream_batch.cu:464-482)
    b->exec = exec;
    return CT_STATUS_SUCCESS;
  });
}

#ifdef TCA_STREAM_FAULT_INJECTION
CTAPI_EXPORT void ct_test_stream_batch_fail_next(ctStreamBatchFault_t site) {
  g_next_fault.store(static_cast<int>(site));
}
#endif

}  // extern "C"
