// noalloc.hip -- zfp_hip_encode_noalloc: zfp_hip_encode for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).  The scratch of variable-rate
// encoding -- block lengths, block offsets and the scan's tile sums -- lives
// in one file-scope device array instead of the library's grow-only buffers,
// and no block index is recorded (that would allocate).  Fixed-rate encoding
// uses no scratch at all and is the same as zfp_hip_encode.
//
// The array is `scratch` below; host code never uses its name as an address
// (in host code the name is only a stand-in): the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "backend.h"
#include "zfp_hip.h"

namespace zfphip {

namespace {

// 80 MiB: block lengths and offsets of up to ZFP_HIP_NOALLOC_MAX_BLOCKS
// blocks (16 bytes each) plus the scan's tile sums (8 bytes per 2048 blocks)
constexpr u64 SCRATCH_WORDS = (80ull << 20) / sizeof(u64);

__device__ __align__(256) u64 scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
u64* g_scratch_ptr[MAX_DEVICES] = {};

// the device address of `scratch` on the current device (lock held)
int scratch_address(u64** p)
{
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES)
    return record_error(ZFP_HIP_ENODEV, "device number beyond the scratch table");
  if (!g_scratch_ptr[dev]) {
    void* a = nullptr;
    HIPCHK(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (u64*)a;
  }
  *p = g_scratch_ptr[dev];
  return ZFP_HIP_OK;
}

} // namespace

int dispatch_encode(const Desc& d, const void* data, const Params& z, u64* out, u64 offset, u64* bits,
                    hipStream_t st, u64** boffs, const u64* at, const Workspace* ws);

} // namespace zfphip

using namespace zfphip;

extern "C" int zfp_hip_encode_noalloc(int type, const size_t n[4], const ptrdiff_t s[4], const void* d_data,
                                      uint minbits, uint maxbits, uint maxprec, int minexp, uint64_t* d_words,
                                      uint64_t offset, uint64_t* bits_out, void* hip_stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  int rc = device_check();
  if (rc)
    return rc;
  Desc d;
  rc = make_desc(type, n, s, d);
  if (rc)
    return rc;
  if (minbits > maxbits || !maxprec || maxprec > 64 || !d_data || !d_words)
    return record_error(ZFP_HIP_EINVAL, "invalid codec parameters or null pointer");
  if (d.nblocks > ZFP_HIP_NOALLOC_MAX_BLOCKS)
    return record_error(ZFP_HIP_EINVAL, "more blocks than the allocation-free encode's scratch holds");
  Params z{minbits, maxbits, maxprec, minexp};
  u64* base = nullptr;
  if ((rc = scratch_address(&base)))
    return rc;
  const u64 nb1 = (u64)d.nblocks + 1;
  Workspace ws{base, base + nb1, base + 2 * nb1};
  static_assert(2 * ((u64)ZFP_HIP_NOALLOC_MAX_BLOCKS + 1) + ((u64)ZFP_HIP_NOALLOC_MAX_BLOCKS + 1 + 2047) / 2048 <=
                    SCRATCH_WORDS, "scratch holds the largest field admitted");
  u64 bits = 0;
  rc = dispatch_encode(d, d_data, z, d_words, offset, &bits, (hipStream_t)hip_stream, nullptr, nullptr, &ws);
  if (rc)
    return rc;
  if (bits_out)
    *bits_out = bits;
  return ZFP_HIP_OK;
}
