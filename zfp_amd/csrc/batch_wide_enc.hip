// batch_wide_enc.hip -- instantiation of the batched codec's ENC pass for 4D
// and reversible jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_ENC, true);
} // namespace zfphip
