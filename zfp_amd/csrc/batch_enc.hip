// batch_enc.hip -- instantiation of the batched codec's ENC pass.
#include "batch_kernels.h"

namespace zfphip {
template int launch_batch_pass<BK_ENC>(unsigned, unsigned, hipStream_t, const BJob*, const uint2*, u32, u32, u64*,
                                   const u64*);
} // namespace zfphip
