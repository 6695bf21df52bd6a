// batch_dec.hip -- instantiation of the batched codec's DEC pass for 1-3D
// lossy jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_DEC, false);
} // namespace zfphip
