// batch_wide_len.hip -- instantiation of the batched codec's LEN pass for 4D
// and reversible jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_LEN, true);
} // namespace zfphip
