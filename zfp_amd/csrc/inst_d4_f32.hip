// Instantiates the 4D kernels of one scalar type (float); the 4D family is
// split by type so that its units compile in parallel.
#include "codec_kernels.h"

namespace zfphip {
template int launch_encode<float, 4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
template int launch_decode<float, 4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*,
                                    hipStream_t);
} // namespace zfphip
