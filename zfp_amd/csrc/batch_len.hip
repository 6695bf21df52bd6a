// batch_len.hip -- instantiation of the batched codec's LEN pass for 1-3D
// lossy jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_LEN, false);
} // namespace zfphip
