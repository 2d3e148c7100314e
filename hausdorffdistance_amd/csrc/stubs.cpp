// Temporary: variants not yet built return HD_EUNSUPPORTED.
#include "hd_internal.h"


extern "C" {
int hd_batch_directed_device(hd_context*, int, const float*, const uint64_t*,
                             const float*, const uint64_t*, uint32_t, uint32_t,
                             hd_directed_result*) {
  return hd::Fail(HD_EUNSUPPORTED, "batch not built yet");
}
}
