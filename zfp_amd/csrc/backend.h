// backend.h -- host-side declarations shared by the launchers
// (codec_kernels.h, instantiated per dimensionality) and the C ABI
// (zfp_hip.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "codec_block.h"
#include "zfp_hip.h"

namespace zfphip {

// a field: type, dimensionality, geometry and block count
struct Desc {
  int type;
  int dims;
  Geom g;
  u32 nblocks;
  size_t esize;
};

// Records a failed HIP call (message available from zfp_hip_last_error),
// clears HIP's error state and returns ZFP_HIP_ERUNTIME.
int record_hip_error(hipError_t e, const char* what);
// Records a library-level failure reason; returns `code`.
int record_error(int code, const char* what);

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return record_hip_error(e_, #x); } while (0)

// fixed-rate stream: every block exactly maxbits long (reversible streams,
// minexp below ZFP_MIN_EXP, never are)
inline bool fixed_rate(const Params& z) { return z.minbits == z.maxbits && z.minexp >= -1074; }

// Whether the block offsets compression writes are where decompression
// reads the blocks.  They are not for reversible streams with minbits > 1:
// an all-zero float block is written as one bit and read as minbits bits
// (src/template/revencodef.c:64-68, revdecodef.c:24-59), so serial
// decompression places the following blocks elsewhere, and parity means
// following it (the sequential parse does).
inline bool index_matches_decoder(const Params& z) { return !(z.minexp < -1074 && z.minbits > 1); }

// the library lock: every entry point that touches shared state holds it
extern std::mutex g_mutex;
// field descriptor from zfp-style extents and strides (0 = default)
int make_desc(int type, const size_t* n, const ptrdiff_t* s, Desc& d);
// 0 when the current device is a usable gfx950 (lock held)
int device_check();
// whether p is device (or managed) memory
bool is_device_ptr(const void* p);

// zero words [w0, w1) of a stream, keeping the low `keep` bits of word w0
__global__ void k_prepare(u64* out, u64 w0, u64 w1, u32 keep);
// the same for words [offset/64, (offset + *bits + 63)/64), *bits on the device
__global__ void k_prepare_dev(u64* out, u64 offset, const u64* bits);
// exclusive prefix sum of n u64 values (in and out may not alias)
int exclusive_scan(const u64* in, u64* out, u64 n, hipStream_t st);
// the same with a caller-owned temporary of scan_tmp_words(n) words; it
// allocates nothing, so it may be captured into a HIP graph
u64 scan_tmp_words(u64 n);
int exclusive_scan_tmp(const u64* in, u64* out, u64 n, u64* tmp, hipStream_t st);

// grow-only device buffer
struct Scratch {
  void* p = nullptr;
  size_t n = 0;
  int get(size_t bytes, void** out)
  {
    if (bytes > n) {
      if (p)
        (void)hipFree(p);
      p = nullptr;
      n = 0;
      if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return record_error(ZFP_HIP_ENOMEM, "device scratch allocation failed");
      }
      n = bytes;
    }
    *out = p;
    return ZFP_HIP_OK;
  }
  void release()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// scratch of the variable-rate passes (lengths + offsets, scan tiles), the
// index-check flag of variable-rate decoding, and the staged bits of
// variable-rate coding (encode_once) and the chain rounds of index-less
// decoding; all entry points run under one library lock, so one set suffices
extern Scratch g_lens, g_scan_tmp, g_flag, g_stage;

// internal status of dispatch_decode (never returned by the C ABI): a
// block's length disagrees with the block index the caller supplied, so the
// index does not describe this stream (the field may hold partial values)
constexpr int ZFP_HIP_ESTALE = 100;

// device buffer for a stream's block index of at least `bytes`: the spare
// buffer of the index registry when it is large enough, else a new
// allocation.  The caller owns it: it goes to the registry or back through
// index_release.
int index_buffer(u64 bytes, u64** p);
void index_release(u64* p, u64 bytes);

// launch the codec for one field; defined in codec_kernels.h and instantiated
// for D = 1..4 in inst_d<D>.hip.  On variable-rate encodes, *boffs (when
// non-null) receives the nblocks+1 block offsets in a buffer from
// index_buffer, which the caller owns.
// Device scratch of a variable-rate encode that the caller provides instead
// of the library's grow-only buffers: block lengths and offsets of nblocks + 1
// words each, and scan_tmp_words(nblocks + 1) words of scan tile sums.  With
// it the encode allocates nothing and records no block index.
struct Workspace {
  u64* lens;
  u64* offs;
  u64* tmp;
};

template <int D>
int dispatch_encode_d(const Desc& d, const void* data, const Params& z, u64* out, u64 offset,
                      u64* bits, hipStream_t st, u64** boffs, const u64* at = nullptr,
                      const Workspace* ws = nullptr);
template <int D>
int dispatch_decode_d(const Desc& d, void* data, const Params& z, const u64* in, u64 nwords,
                      u64 offset, const u64* boffs, u64* bits, hipStream_t st);

extern template int dispatch_encode_d<1>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                           const u64*, const Workspace*);
extern template int dispatch_encode_d<2>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                           const u64*, const Workspace*);
extern template int dispatch_encode_d<3>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                           const u64*, const Workspace*);
extern template int dispatch_encode_d<4>(const Desc&, const void*, const Params&, u64*, u64, u64*, hipStream_t, u64**,
                                           const u64*, const Workspace*);
extern template int dispatch_decode_d<1>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);
extern template int dispatch_decode_d<2>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);
extern template int dispatch_decode_d<3>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);
extern template int dispatch_decode_d<4>(const Desc&, void*, const Params&, const u64*, u64, u64, const u64*, u64*, hipStream_t);

} // namespace zfphip
