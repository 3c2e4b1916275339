// Ragged batch layouts (no counterpart in the reference ABI): thousands of
// series of uneven lengths, concatenated, run in one launch per family.
#include <hip/hip_runtime.h>

#include <climits>
#include <vector>

#include "../../include/tacuda_amd.h"
#include "host.hpp"

struct ctxRagged_st {
  tca::Ragged r;
};

using namespace tca;

namespace {

void release(Ragged& r) {
  if (r.dOff) (void)hipFree(r.dOff);
  for (int i = 0; i < 3; ++i)
    if (r.dSegStart[i]) (void)hipFree(r.dSegStart[i]);
  r = Ragged{};
}

}  // namespace

extern "C" {

CTAPI_EXPORT ctStatus_t ctx_ragged_create(const long long* offsets, int numSeries,
                                          ctxRagged* out) {
  return guarded([&]() -> ctStatus_t {
    if (!offsets || !out) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: null pointer");
    *out = nullptr;
    if (numSeries <= 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: numSeries must be positive");
    if (offsets[0] < 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: negative offset");
    int maxLen = 0;
    for (int r = 0; r < numSeries; ++r) {
      const long long len = offsets[r + 1] - offsets[r];
      if (len < 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: offsets must not decrease");
      if (len > INT_MAX) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: series too long");
      if (len > maxLen) maxLen = (int)len;
    }
    if (maxLen == 0) return fail(CT_STATUS_INVALID_ARGUMENT, "ragged: every series is empty");
    auto* h = new ctxRagged_st;
    Ragged& R = h->r;
    R.rows = numSeries;
    R.maxLen = maxLen;
    R.total = offsets[numSeries] - offsets[0];
    auto bail = [&](ctStatus_t st) {
      release(R);
      delete h;
      return st;
    };
    hipError_t e = hipMalloc(&R.dOff, sizeof(long long) * (numSeries + 1));
    if (e == hipSuccess)
      e = hipMemcpy(R.dOff, offsets, sizeof(long long) * (numSeries + 1), hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "ragged: offsets upload"));
    std::vector<int> ss(numSeries + 1);
    for (int i = 0; i < 3; ++i) {
      const int seg = kRaggedSegSizes[i];
      long long acc = 0;
      int mx = 0;
      for (int r = 0; r < numSeries; ++r) {
        ss[r] = (int)acc;
        const long long len = offsets[r + 1] - offsets[r];
        const int n = (int)((len + seg - 1) / seg);
        if (n > mx) mx = n;
        acc += n;
        if (acc > INT_MAX) return bail(fail(CT_STATUS_INVALID_ARGUMENT, "ragged: too many segments"));
      }
      ss[numSeries] = (int)acc;
      R.totalSegs[i] = (int)acc;
      R.maxSegs[i] = mx;
      e = hipMalloc(&R.dSegStart[i], sizeof(int) * (numSeries + 1));
      if (e == hipSuccess)
        e = hipMemcpy(R.dSegStart[i], ss.data(), sizeof(int) * (numSeries + 1),
                      hipMemcpyHostToDevice);
      if (e != hipSuccess) return bail(hip_fail(e, "ragged: segment table upload"));
    }
    *out = h;
    return CT_STATUS_SUCCESS;
  });
}

CTAPI_EXPORT ctStatus_t ctx_ragged_destroy(ctxRagged layout) {
  return guarded([&]() -> ctStatus_t {
    if (!layout) return CT_STATUS_SUCCESS;
    (void)hipDeviceSynchronize();
    release(layout->r);
    delete layout;
    return CT_STATUS_SUCCESS;
  });
}

}  // extern "C"

namespace tca {
const Ragged* ragged_of(ctxRagged layout) { return layout ? &layout->r : nullptr; }
}  // namespace tca
