// Device buffer pool, pinned staging and the HostRun helper of the host
// entry points.
#include "hostio.hpp"

#include <atomic>
#include <cstring>
#include <iterator>
#include <limits>
#include <map>
#include <mutex>
#include <unordered_map>

#include "host.hpp"

namespace tca {

namespace {

// ------------------------------------------------------------ device pool --
struct DevPool {
  std::mutex mu;
  std::map<std::pair<int, size_t>, std::vector<void*>> free;  // (device, class)
  std::unordered_map<void*, std::pair<int, size_t>> live;     // every pool block
  size_t cached = 0;
  size_t maxCached = std::numeric_limits<size_t>::max();
  size_t allocs = 0;
};
DevPool& dpool() {
  static DevPool* p = new DevPool;  // outlives static destructors
  return *p;
}

// Power-of-two classes from 64 KiB to 256 MiB, then 64 MiB multiples.
size_t size_class(size_t bytes) {
  const size_t lo = size_t(64) << 10, big = size_t(256) << 20, step = size_t(64) << 20;
  if (bytes <= lo) return lo;
  if (bytes > big) return (bytes + step - 1) / step * step;
  size_t c = lo;
  while (c < bytes) c <<= 1;
  return c;
}

// Frees cached blocks until at most `limit` bytes stay cached (mu held).
void evict_locked(DevPool& P, size_t limit) {
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto it = P.free.begin(); it != P.free.end() && P.cached > limit;) {
    auto& v = it->second;
    while (!v.empty() && P.cached > limit) {
      void* b = v.back();
      v.pop_back();
      (void)hipSetDevice(it->first.first);
      (void)hipFree(b);
      P.live.erase(b);
      P.cached -= it->first.second;
    }
    it = v.empty() ? P.free.erase(it) : std::next(it);
  }
  (void)hipSetDevice(cur);
}

// ---------------------------------------------------------- pinned staging --
constexpr size_t kChunk = size_t(4) << 20;     // bytes per staging chunk
constexpr size_t kStageMin = size_t(1) << 20;  // staged from 1 MiB (mode 0)

struct PinPair {
  void* buf = nullptr;  // 2 * kChunk pinned bytes
  hipEvent_t ev[2] = {nullptr, nullptr};
};
std::mutex g_pin_mu;
std::vector<PinPair>* g_pin_free = new std::vector<PinPair>;
std::unordered_map<void*, PinPair>* g_pin_live = new std::unordered_map<void*, PinPair>;

std::atomic<int> g_mode{0};

}  // namespace

ctStatus_t pool_acquire(size_t bytes, void** out) {
  *out = nullptr;
  int dev = 0;
  TCA_HIP(hipGetDevice(&dev));
  const size_t cls = size_class(bytes == 0 ? 1 : bytes);
  DevPool& P = dpool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    auto it = P.free.find({dev, cls});
    if (it != P.free.end() && !it->second.empty()) {
      *out = it->second.back();
      it->second.pop_back();
      P.cached -= cls;
      return CT_STATUS_SUCCESS;
    }
  }
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cls);
  if (e != hipSuccess) {
    // drop the cache and try once more before giving up
    {
      std::lock_guard<std::mutex> lk(P.mu);
      evict_locked(P, 0);
    }
    (void)hipGetLastError();
    e = hipMalloc(&p, cls);
    if (e != hipSuccess) return hip_fail(e, "device buffer pool");
  }
  std::lock_guard<std::mutex> lk(P.mu);
  P.live[p] = {dev, cls};
  ++P.allocs;
  *out = p;
  return CT_STATUS_SUCCESS;
}

void pool_release(void* p) {
  if (!p) return;
  DevPool& P = dpool();
  std::lock_guard<std::mutex> lk(P.mu);
  auto it = P.live.find(p);
  if (it == P.live.end()) return;
  const auto key = it->second;
  if (P.cached + key.second > P.maxCached) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(key.first);
    (void)hipFree(p);
    (void)hipSetDevice(cur);
    P.live.erase(it);
    return;
  }
  P.free[key].push_back(p);
  P.cached += key.second;
}

void pool_cleanup() {
  DevPool& P = dpool();
  {
    std::lock_guard<std::mutex> lk(P.mu);
    evict_locked(P, 0);
  }
  std::lock_guard<std::mutex> lk(g_pin_mu);
  for (PinPair& q : *g_pin_free) {
    (void)hipHostFree(q.buf);
    for (hipEvent_t e : q.ev)
      if (e) (void)hipEventDestroy(e);
  }
  g_pin_free->clear();
}

size_t pool_allocations() {
  DevPool& P = dpool();
  std::lock_guard<std::mutex> lk(P.mu);
  return P.allocs;
}

size_t pool_set_max_cache(size_t bytes) {
  DevPool& P = dpool();
  std::lock_guard<std::mutex> lk(P.mu);
  const size_t old = P.maxCached;
  P.maxCached = bytes;
  if (P.cached > bytes) evict_locked(P, bytes);
  return old;
}

int host_staging_mode() { return g_mode.load(std::memory_order_relaxed); }
void set_host_staging_mode(int mode) { g_mode.store(mode, std::memory_order_relaxed); }

// ----------------------------------------------------------------- HostRun --
HostRun::~HostRun() {
  if (!done_) (void)hipStreamSynchronize(s_);
  for (void* d : dev_) pool_release(d);
  if (pin_) {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    auto it = g_pin_live->find(pin_);
    if (it != g_pin_live->end()) {
      g_pin_free->push_back(it->second);
      g_pin_live->erase(it);
    }
  }
}

ctStatus_t HostRun::stage_slot(int k, char** buf) {
  if (!pin_) {
    PinPair q;
    {
      std::lock_guard<std::mutex> lk(g_pin_mu);
      if (!g_pin_free->empty()) {
        q = g_pin_free->back();
        g_pin_free->pop_back();
      }
    }
    if (!q.buf) {
      TCA_HIP(hipHostMalloc(&q.buf, 2 * kChunk, hipHostMallocDefault));
      for (hipEvent_t& e : q.ev) {
        hipError_t r = hipEventCreateWithFlags(&e, hipEventDisableTiming);
        if (r != hipSuccess) {
          (void)hipHostFree(q.buf);
          return hip_fail(r, "staging event");
        }
      }
    }
    {
      std::lock_guard<std::mutex> lk(g_pin_mu);
      (*g_pin_live)[q.buf] = q;
    }
    pin_ = q.buf;
    ev_[0] = q.ev[0];
    ev_[1] = q.ev[1];
  }
  const int slot = k & 1;
  // the slot's previous transfer must be over before it is overwritten
  if (evUsed_[slot]) TCA_HIP(hipEventSynchronize(ev_[slot]));
  *buf = static_cast<char*>(pin_) + slot * kChunk;
  return CT_STATUS_SUCCESS;
}

ctStatus_t HostRun::staged_upload(void* dev, const void* host, size_t bytes) {
  for (size_t off = 0; off < bytes; off += kChunk) {
    const size_t len = bytes - off < kChunk ? bytes - off : kChunk;
    const int k = chunk_++;
    char* buf = nullptr;
    ctStatus_t r = stage_slot(k, &buf);
    if (r != CT_STATUS_SUCCESS) return r;
    std::memcpy(buf, static_cast<const char*>(host) + off, len);
    hipError_t e = hipMemcpyAsync(static_cast<char*>(dev) + off, buf, len,
                                  hipMemcpyHostToDevice, s_);
    if (e == hipSuccess) e = hipEventRecord(ev_[k & 1], s_);
    if (e != hipSuccess) {
      hip_fail(e, "host upload");
      return CT_STATUS_COPY_FAILED;
    }
    evUsed_[k & 1] = true;
  }
  return CT_STATUS_SUCCESS;
}

ctStatus_t HostRun::staged_download(void* host, const void* dev, size_t bytes) {
  // chunk i is copied into its slot while chunk i-1 is copied out of the
  // other one
  bool pend = false;
  int pk = 0;
  size_t poff = 0, plen = 0;
  auto drain = [&]() -> ctStatus_t {
    TCA_HIP(hipEventSynchronize(ev_[pk & 1]));
    std::memcpy(static_cast<char*>(host) + poff, static_cast<char*>(pin_) + (pk & 1) * kChunk,
                plen);
    pend = false;
    return CT_STATUS_SUCCESS;
  };
  for (size_t off = 0; off < bytes; off += kChunk) {
    const size_t len = bytes - off < kChunk ? bytes - off : kChunk;
    const int k = chunk_++;
    char* buf = nullptr;
    ctStatus_t r = stage_slot(k, &buf);
    if (r != CT_STATUS_SUCCESS) return r;
    hipError_t e = hipMemcpyAsync(buf, static_cast<const char*>(dev) + off, len,
                                  hipMemcpyDeviceToHost, s_);
    if (e == hipSuccess) e = hipEventRecord(ev_[k & 1], s_);
    if (e != hipSuccess) {
      hip_fail(e, "host download");
      return CT_STATUS_COPY_FAILED;
    }
    evUsed_[k & 1] = true;
    if (pend) {
      r = drain();
      if (r != CT_STATUS_SUCCESS) return r;
    }
    pend = true;
    pk = k;
    poff = off;
    plen = len;
  }
  return pend ? drain() : CT_STATUS_SUCCESS;
}

static bool staged(size_t bytes) {
  const int m = host_staging_mode();
  return m == 2 || (m == 0 && bytes >= kStageMin);
}

ctStatus_t HostRun::in(const void* host, size_t bytes, const void** dev) {
  *dev = nullptr;
  if (!host) return fail(CT_STATUS_INVALID_ARGUMENT, "host call: null input pointer");
  void* d = nullptr;
  ctStatus_t r = pool_acquire(bytes, &d);
  if (r != CT_STATUS_SUCCESS) return r;
  dev_.push_back(d);
  if (staged(bytes)) {
    r = staged_upload(d, host, bytes);
    if (r != CT_STATUS_SUCCESS) return r;
  } else {
    hipError_t e = hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, s_);
    if (e != hipSuccess) {
      hip_fail(e, "host upload");
      return CT_STATUS_COPY_FAILED;
    }
  }
  *dev = d;
  return CT_STATUS_SUCCESS;
}

ctStatus_t HostRun::out(void* host, size_t bytes, void** dev) {
  *dev = nullptr;
  if (!host) return fail(CT_STATUS_INVALID_ARGUMENT, "host call: null output pointer");
  void* d = nullptr;
  ctStatus_t r = pool_acquire(bytes, &d);
  if (r != CT_STATUS_SUCCESS) return r;
  dev_.push_back(d);
  outs_.push_back(Out{host, d, bytes});
  *dev = d;
  return CT_STATUS_SUCCESS;
}

ctStatus_t HostRun::finish() {
  for (const Out& o : outs_) {
    if (staged(o.bytes)) {
      ctStatus_t r = staged_download(o.host, o.dev, o.bytes);
      if (r != CT_STATUS_SUCCESS) return r;
    } else {
      hipError_t e = hipMemcpyAsync(o.host, o.dev, o.bytes, hipMemcpyDeviceToHost, s_);
      if (e != hipSuccess) {
        hip_fail(e, "host download");
        return CT_STATUS_COPY_FAILED;
      }
    }
  }
  TCA_HIP(hipStreamSynchronize(s_));
  done_ = true;
  return CT_STATUS_SUCCESS;
}

}  // namespace tca
