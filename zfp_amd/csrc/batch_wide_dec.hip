// batch_wide_dec.hip -- instantiation of the batched codec's DEC pass for 4D
// and reversible jobs.
#include "batch_kernels.h"

namespace zfphip {
template ZH_BATCH_PASS(BK_DEC, true);
} // namespace zfphip
