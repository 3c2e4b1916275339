// Host entry points (ct_<name>, ct_<name>_batch, ct_cdl_<name>): device
// buffers come from a process-wide pool and host arrays travel through
// pinned staging buffers, so a host call allocates nothing once the pool is
// warm (reference api.cpp:230-287 with DeviceBufferPool.h:19-88; the
// reference page-locks host matrices of 1 MB or more, api_batch.cu:56,
// 186-208 -- here arrays of 1 MB or more are copied through a pair of
// pinned chunk buffers whose transfers overlap the host-side copies, and
// smaller ones are copied directly).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/tacuda.h"

namespace tca {

// Device buffer pool: blocks in power-of-two size classes per device, kept
// after use up to a cache limit (ct_device_buffer_pool_set_max_cache_bytes)
// and freed by ct_device_buffer_pool_cleanup.
ctStatus_t pool_acquire(size_t bytes, void** out);
void pool_release(void* p);
void pool_cleanup();
size_t pool_allocations();
size_t pool_set_max_cache(size_t bytes);

// How host arrays reach the device: 0 = by size (pinned staging from 1 MB,
// direct copies below), 1 = always direct (pageable) copies, 2 = always
// pinned staging.
int host_staging_mode();
void set_host_staging_mode(int mode);

// One host call: `in` acquires a device buffer and starts the upload of a
// host array, `out` acquires a device buffer for a result and remembers its
// host destination, `finish` downloads every result in order and waits for
// the stream.  Buffers go back to the pool when the HostRun ends (after
// waiting for the stream if `finish` did not complete).
class HostRun {
 public:
  explicit HostRun(hipStream_t s) : s_(s) {}
  ~HostRun();
  HostRun(const HostRun&) = delete;
  HostRun& operator=(const HostRun&) = delete;

  ctStatus_t in(const void* host, size_t bytes, const void** dev);
  ctStatus_t out(void* host, size_t bytes, void** dev);
  ctStatus_t finish();

  template <class T>
  ctStatus_t in(const T* host, size_t n, const T** dev) {
    const void* d = nullptr;
    ctStatus_t r = in(static_cast<const void*>(host), n * sizeof(T), &d);
    *dev = static_cast<const T*>(d);
    return r;
  }
  template <class T>
  ctStatus_t out(T* host, size_t n, T** dev) {
    void* d = nullptr;
    ctStatus_t r = out(static_cast<void*>(host), n * sizeof(T), &d);
    *dev = static_cast<T*>(d);
    return r;
  }
  hipStream_t stream() const { return s_; }

 private:
  struct Out {
    void* host;
    const void* dev;
    size_t bytes;
  };
  ctStatus_t staged_upload(void* dev, const void* host, size_t bytes);
  ctStatus_t staged_download(void* host, const void* dev, size_t bytes);
  ctStatus_t stage_slot(int k, char** buf);

  hipStream_t s_;
  std::vector<void*> dev_;
  std::vector<Out> outs_;
  void* pin_ = nullptr;  // two chunk buffers, from the pinned pool
  hipEvent_t ev_[2] = {nullptr, nullptr};
  bool evUsed_[2] = {false, false};
  int chunk_ = 0;  // chunks issued through the staging pair so far
  bool done_ = false;
};

}  // namespace tca
