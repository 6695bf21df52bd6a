// Instantiates the 3D kernel family (see codec_kernels.h).
#include "codec_kernels.h"

namespace zfphip {
template int dispatch_encode_d<3>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
template int dispatch_decode_d<3>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);
} // namespace zfphip
