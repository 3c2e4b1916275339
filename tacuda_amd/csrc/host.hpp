// Host-side plumbing shared by the C ABI translation units: thread-local
// error state (reference src/utils/ErrorState.cpp:6-60), argument checks,
// the per-stream look-back workspace cache and launch helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <exception>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/tacuda_amd.h"
#include "engine.hpp"

namespace tca {

// ------------------------------------------------------------- errors --
void clear_error();
ctStatus_t fail(ctStatus_t st, const char* fmt, ...);
ctStatus_t hip_fail(hipError_t e, const char* what);

#define TCA_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return ::tca::hip_fail(e_, #call); \
  } while (0)

// Runs `body` (returning ctStatus_t) with the error cleared on entry and no
// C++ exception escaping the ABI (reference api.cpp:217-228).
template <class F>
ctStatus_t guarded(F&& body) {
  clear_error();
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(CT_STATUS_ALLOC_FAILED, "host allocation failed");
  } catch (const std::exception& ex) {
    return fail(CT_STATUS_DEVICE_ERROR, "%s", ex.what());
  } catch (...) {
    return fail(CT_STATUS_DEVICE_ERROR, "unknown error");
  }
}

// ---------------------------------------------------------- workspace --
// Look-back tree area plus the ticket/epoch header, one per HIP stream and
// part so that launches ordered on one stream share it safely; one launch
// whose blocks run two independent look-back scans gives each its own part,
// and the fused pass's one-launch form keeps its tagged records in a part of
// their own (kWsParts per stream).  Growing it is refused while the stream is being
// captured; growth outside capture retires the old buffer instead of freeing
// it (graphs captured earlier keep its address), and retired buffers are
// freed when the stream's workspaces are released or at pool cleanup.
constexpr int kWsParts = 3;
ctStatus_t workspace_for(hipStream_t stream, size_t treeBytes, Ws* out, int part = 0);
// While a WsFixedScope is alive on a thread, workspace_for hands out these
// caller-provided buffers (each `bytes` of area after a WsHeader) instead of
// the per-stream cache, and refuses requests larger than `bytes`.
struct WsFixed {
  char* part[kWsParts];
  size_t bytes;
};
extern thread_local const WsFixed* t_ws_fixed;
struct WsFixedScope {
  explicit WsFixedScope(const WsFixed* f) { t_ws_fixed = f; }
  ~WsFixedScope() { t_ws_fixed = nullptr; }
};
void workspace_release_all();
// Per-stream temporary device buffers (slot 0..kTmpSlots-1) for kernels that
// need whole-row scratch; growth follows the workspace rules (refused under
// capture, old buffers retired until the stream's release or pool cleanup).
constexpr int kTmpSlots = 2;
ctStatus_t temp_for(hipStream_t stream, size_t bytes, int slot, void** out);
void workspace_release(hipStream_t stream);
size_t workspace_allocations();

// -------------------------------------------------------- kernel timer --
// While ctx_kernel_timing is on, a KTime scope records a pair of HIP events
// around the launches it encloses on `stream` (none while the stream is
// being captured).
class KTime {
 public:
  KTime(hipStream_t stream, const char* name);
  ~KTime();
  KTime(const KTime&) = delete;
  KTime& operator=(const KTime&) = delete;

 private:
  hipStream_t s_;
  const char* name_;
  hipEvent_t a_ = nullptr;
};

// ------------------------------------------------------------ geometry --
inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

inline Geometry dense_geometry(int len, int rows, int seg) {
  Geometry g{};
  g.offsets = nullptr;
  g.segStart = nullptr;
  g.ld = len;
  g.len = len;
  g.rows = rows;
  g.segsPerRow = div_up(len, seg);
  g.totalSegs = g.segsPerRow * rows;
  return g;
}

// ------------------------------------------------------------- ragged --
// Layout of a ragged batch: series r occupies [off[r], off[r+1]) of the
// concatenated arrays.  Built once on the host; keeps the device copy of the
// offsets and, per segment size, the first global segment of every row.
constexpr int kRaggedSegSizes[3] = {256, 512, 1024};
struct Ragged {
  int rows = 0;
  int maxLen = 0;
  long long total = 0;
  long long* dOff = nullptr;             // device, rows + 1
  int* dSegStart[3] = {nullptr, nullptr, nullptr};  // device, rows + 1
  int totalSegs[3] = {0, 0, 0};
  int maxSegs[3] = {0, 0, 0};
};

// Set for the duration of a ragged entry point: the launchers build ragged
// geometry from it instead of the dense (len, rows) shape.
extern thread_local const Ragged* t_ragged;
struct RaggedScope {
  explicit RaggedScope(const Ragged* r) { t_ragged = r; }
  ~RaggedScope() { t_ragged = nullptr; }
};

// Geometry of the active ragged layout for a segment size; false when the
// size has no table (never for the tile sizes the kernels use).
bool ragged_geometry(const Ragged* r, int seg, Geometry* g, int* maxSegsPerRow);
const Ragged* ragged_of(struct ctxRagged_st* layout);

// ------------------------------------------------------- fused scan --
// The fused long-series scan (EMA, KAMA, T3, SMA) without the stream-0
// synchronization.  scan_serves: whether the family calls of those
// indicators route a dense batch to it -- at least kScanMinRows rows of at
// least two 1024-bar segments, kScanMinSegments segments in all (batches
// that fill the GPU; with a few long rows its per-row look-back chains leave
// the GPU short of independent work), and a window look-back within its
// halo.
constexpr long long kScanMinSegments = 512;
constexpr int kScanMinRows = 16;
ctStatus_t scan_run(const float* close, int seriesLen, int numSeries, const ctxScanParams* params,
                    const ctxScanOutputs* outputs, hipStream_t stream);
bool scan_serves(int len, int rows, int windowLookback);

// ------------------------------------------------------ fused pass --
// The family calls of EMA, RSI, MACD and ATR on rows of at least 8192 bars run
// the fused pass (the pack's outputs for that family, the same
// oracle-following arithmetic): its one-launch form wherever that serves (at
// most 1024 rows, 16384 512-bar segments and 1,050,624 bars per row), its
// two passes for everything larger.  Both were measured at least as fast as
// the per-family look-back kernels on every shape they take (one row of 8192
// to 2.2 million bars, up to 256 rows of 16384 and 8 of 2^20; within 2% on
// one ATR row of 1.1-1.5 million bars, where the look-back chains are
// shortest relative to the passes' fixed launches).  *served says whether
// the pass ran; when not, nothing was launched and the caller takes its own
// path.
ctStatus_t pipe_family(const ctxPackParams& params, const ctxPackOutputs& outputs,
                       const float* h, const float* l, const float* c, int len, int rows,
                       hipStream_t stream, bool* served);
// The family calls of SMA, WMA, BBANDS, STOCH, WILLR and CCI (dense rows of
// any length, window look-backs within the pack's 1024 bars) run the fused
// pass's window kernel: one wave per 256-bar column tile, every window
// evaluated directly in the oracle's order, so the outputs are those of the
// family kernels.  *served as for pipe_family.
ctStatus_t win_family(const ctxPackParams& params, const ctxPackOutputs& outputs,
                      const float* h, const float* l, const float* c, int len, int rows,
                      hipStream_t stream, bool* served);

// The body of ctx_pack_batch_d (validation, the fused pass, the stream-0
// wait) without the ABI guard.
ctStatus_t pack_batch_body(const float* h, const float* l, const float* c, int seriesLen,
                           int numSeries, const ctxPackParams* p, const ctxPackOutputs* o,
                           hipStream_t s);

// Synchronize when the caller used the legacy default stream: batch calls
// on stream 0 return once the batch is done (reference api_batch.cu:377-386).
inline ctStatus_t finish_batch(hipStream_t stream) {
  if (stream == nullptr) TCA_HIP(hipStreamSynchronize(nullptr));
  return CT_STATUS_SUCCESS;
}

}  // namespace tca
