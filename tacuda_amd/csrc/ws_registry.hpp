// Bookkeeping of the look-back workspaces, kept free of HIP so that its
// rules can be unit-tested on the host.
//
// One live buffer per (device, stream, part).  A graph captured on a stream
// keeps the address of the buffer that stream had at capture time, so
// growing a stream's buffer never frees the old one: the old buffer is
// retired, stays allocated, and is handed back for freeing only when that
// stream's workspaces are released (a plan's private stream when the plan is
// destroyed) or when every workspace is released (pool cleanup).  Growth is
// geometric, so a stream retires O(log(size)) buffers at most.
#pragma once
#include <cstddef>
#include <map>
#include <tuple>
#include <vector>

namespace tca {

class WsRegistry {
 public:
  struct Key {
    int device;
    const void* stream;
    int part;
    bool operator<(const Key& o) const {
      return std::tie(device, stream, part) < std::tie(o.device, o.stream, o.part);
    }
  };
  struct Buf {
    void* base = nullptr;
    size_t bytes = 0;  // usable bytes (after the header)
    int device = -1;
  };

  static constexpr size_t kMinBytes = size_t(1) << 20;

  // The live buffer of `k` when it holds `need` bytes, else nullptr.
  void* find(const Key& k, size_t need) const {
    auto it = live_.find(k);
    if (it == live_.end() || it->second.base == nullptr || it->second.bytes < need) return nullptr;
    return it->second.base;
  }

  // Bytes to allocate for `need` on `k`: at least kMinBytes and at least
  // twice the buffer it replaces.
  size_t grow_size(const Key& k, size_t need) const {
    size_t want = need < kMinBytes ? kMinBytes : need;
    auto it = live_.find(k);
    if (it != live_.end() && it->second.base && want < 2 * it->second.bytes)
      want = 2 * it->second.bytes;
    return want;
  }

  // Makes `base` the live buffer of `k`; the previous one is retired, not
  // freed.
  void install(const Key& k, void* base, size_t bytes) {
    Buf& b = live_[k];
    if (b.base) retired_.push_back(Entry{k, b});
    b = Buf{base, bytes, k.device};
    ++allocs_;
  }

  // Removes every buffer (live and retired) of (device, stream) and returns
  // them for freeing; the caller frees them once the stream is idle.
  std::vector<Buf> release_stream(int device, const void* stream) {
    std::vector<Buf> out;
    for (auto it = live_.begin(); it != live_.end();) {
      if (it->first.device == device && it->first.stream == stream) {
        if (it->second.base) out.push_back(it->second);
        it = live_.erase(it);
      } else {
        ++it;
      }
    }
    for (auto it = retired_.begin(); it != retired_.end();) {
      if (it->key.device == device && it->key.stream == stream) {
        out.push_back(it->buf);
        it = retired_.erase(it);
      } else {
        ++it;
      }
    }
    return out;
  }

  // Removes and returns every buffer.
  std::vector<Buf> release_all() {
    std::vector<Buf> out;
    for (auto& kv : live_)
      if (kv.second.base) out.push_back(kv.second);
    for (auto& e : retired_) out.push_back(e.buf);
    live_.clear();
    retired_.clear();
    return out;
  }

  size_t allocations() const { return allocs_; }
  size_t retired() const { return retired_.size(); }

 private:
  struct Entry {
    Key key;
    Buf buf;
  };
  std::map<Key, Buf> live_;
  std::vector<Entry> retired_;
  size_t allocs_ = 0;
};

}  // namespace tca
