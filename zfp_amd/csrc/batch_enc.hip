// batch_enc.hip -- instantiation of the batched codec's ENC pass for 1-3D
// lossy jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_ENC, false);
} // namespace zfphip
