/* Test-only hook of the fault-injection build (libtacuda_amd_faults.so, not
 * part of the public ABI): GPU failures cannot be provoked on demand, so the
 * batch-stream error paths are driven through named sites.  Site numbers
 * follow the reference's include/utils/StreamBatchFaults.h:20-31. */
#ifndef TACUDA_AMD_STREAM_FAULTS_H
#define TACUDA_AMD_STREAM_FAULTS_H

#include "../../include/tacuda_stream.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ctStreamBatchFault {
  CT_STREAM_BATCH_FAULT_NONE = 0,
  CT_STREAM_BATCH_FAULT_STATE_UPLOAD = 1,
  CT_STREAM_BATCH_FAULT_IO_ALLOC = 2,
  CT_STREAM_BATCH_FAULT_INPUT_UPLOAD = 3,
  CT_STREAM_BATCH_FAULT_KERNEL = 4,
  CT_STREAM_BATCH_FAULT_OUTPUT_DOWNLOAD = 5,
  CT_STREAM_BATCH_FAULT_STATE_DOWNLOAD = 6
} ctStreamBatchFault_t;

/* The next operation reaching `site` reports failure, once. */
CTAPI_EXPORT void ct_test_stream_batch_fail_next(ctStreamBatchFault_t site);

#ifdef __cplusplus
}
#endif
#endif
