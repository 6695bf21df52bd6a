// batch_len.hip -- instantiation of the batched codec's LEN pass.
#include "batch_kernels.h"

namespace zfphip {
template int launch_batch_pass<BK_LEN>(unsigned, unsigned, hipStream_t, const BJob*, const uint2*, u32, u32, u64*,
                                   const u64*);
} // namespace zfphip
