// Instantiates the 4D dispatch (see codec_kernels.h); the kernels of each
// scalar type are instantiated in inst_d4_<type>.hip.
#include "codec_kernels.h"

namespace zfphip {
extern template int launch_encode<i32, 4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
extern template int launch_decode<i32, 4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*,
                                    hipStream_t);
extern template int launch_encode<i64, 4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
extern template int launch_decode<i64, 4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*,
                                    hipStream_t);
extern template int launch_encode<float, 4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
extern template int launch_decode<float, 4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*,
                                    hipStream_t);
extern template int launch_encode<double, 4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
extern template int launch_decode<double, 4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*,
                                    hipStream_t);

template int dispatch_encode_d<4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                    const u64*, const Workspace*);
template int dispatch_decode_d<4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);
} // namespace zfphip
