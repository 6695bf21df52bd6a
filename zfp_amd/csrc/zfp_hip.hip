// zfp_hip.hip -- C ABI of the MI355X zfp block codec (include/zfp_hip.h):
// argument checking, host/device memory handling, stream-state bookkeeping
// and the per-stream block-index registry used to decode variable-rate
// streams in parallel.  The kernels live in codec_kernels.h and helpers.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "backend.h"
#include "zfp_hip.h"

namespace zfphip {

Scratch g_lens, g_scan_tmp, g_flag, g_stage;

namespace {
thread_local char g_last_error[256] = "";
}

int record_error(int code, const char* what)
{
  snprintf(g_last_error, sizeof(g_last_error), "%s", what);
  return code;
}

int record_hip_error(hipError_t e, const char* what)
{
  (void)hipGetLastError();
  snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
  return ZFP_HIP_ERUNTIME;
}

std::mutex g_mutex;

namespace {
Scratch g_field, g_stream;
} // namespace

int make_desc(int type, const size_t* n, const ptrdiff_t* s, Desc& d)
{
  memset(&d, 0, sizeof(d));
  d.type = type;
  switch (type) {
    case zfp_type_int32: case zfp_type_float: d.esize = 4; break;
    case zfp_type_int64: case zfp_type_double: d.esize = 8; break;
    default: return record_error(ZFP_HIP_EINVAL, "unsupported scalar type");
  }
  // dimensionality as zfp_field_dimensionality (src/zfp.c:211-214)
  d.dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  if (!d.dims)
    return record_error(ZFP_HIP_EINVAL, "field has no extent");
  // default strides (src/zfp.c:260-278)
  i64 def[4];
  def[0] = 1;
  def[1] = (i64)n[0];
  def[2] = (i64)(n[0] * (n[1] ? n[1] : 1));
  def[3] = (i64)(n[0] * (n[1] ? n[1] : 1) * (n[2] ? n[2] : 1));
  u64 nb = 1;
  for (int i = 0; i < 4; i++) {
    if (i < d.dims) {
      if (n[i] > 0xffffffffull)
        return record_error(ZFP_HIP_EINVAL, "extent above 2^32-1");
      d.g.n[i] = (u32)n[i];
      d.g.s[i] = (s && s[i]) ? (i64)s[i] : def[i];
      d.g.nb[i] = (u32)((n[i] + 3) / 4);
      nb *= d.g.nb[i];
    }
    else {
      d.g.n[i] = 1;
      d.g.s[i] = 0;
      d.g.nb[i] = 1;
    }
  }
  if (nb > 0xffffffffull)
    return record_error(ZFP_HIP_EINVAL, "more than 2^32-1 blocks");
  d.g.dims = (u32)d.dims;
  // lane order: axes by increasing |stride| (stable), unused axes last
  for (int i = 0; i < 4; i++)
    d.g.lax[i] = (u32)i;
  for (int i = 1; i < d.dims; i++)
    for (int j = i; j > 0; j--) {
      i64 a = d.g.s[d.g.lax[j]], b = d.g.s[d.g.lax[j - 1]];
      if ((a < 0 ? -a : a) < (b < 0 ? -b : b)) {
        u32 t = d.g.lax[j]; d.g.lax[j] = d.g.lax[j - 1]; d.g.lax[j - 1] = t;
      }
    }
  for (int i = 0; i < 4; i++)
    d.g.lnb[i] = d.g.nb[d.g.lax[i]];
  d.nblocks = (u32)nb;
  return ZFP_HIP_OK;
}

int dispatch_encode(const Desc& d, const void* data, const Params& z, u64* out, u64 offset, u64* bits,
                    hipStream_t st, u64** boffs, const u64* at = nullptr, const Workspace* ws = nullptr)
{
  switch (d.dims) {
    case 1: return dispatch_encode_d<1>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case 2: return dispatch_encode_d<2>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case 3: return dispatch_encode_d<3>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case 4: return dispatch_encode_d<4>(d, data, z, out, offset, bits, st, boffs, at, ws);
  }
  return ZFP_HIP_EINVAL;
}

int dispatch_decode(const Desc& d, void* data, const Params& z, const u64* in, u64 nwords, u64 offset,
                    const u64* boffs, u64* bits, hipStream_t st)
{
  switch (d.dims) {
    case 1: return dispatch_decode_d<1>(d, data, z, in, nwords, offset, boffs, bits, st);
    case 2: return dispatch_decode_d<2>(d, data, z, in, nwords, offset, boffs, bits, st);
    case 3: return dispatch_decode_d<3>(d, data, z, in, nwords, offset, boffs, bits, st);
    case 4: return dispatch_decode_d<4>(d, data, z, in, nwords, offset, boffs, bits, st);
  }
  return ZFP_HIP_EINVAL;
}

namespace {

// Block index of a variable-rate stream: offsets[b] relative to `start`
// (nblocks + 1 entries, on the device).  Keyed by the stream buffer's begin
// pointer; an entry is tried for a decode only at the same start offset,
// block count and codec parameters (type/dims/extents too unless imported).
// The key says nothing about the buffer's contents -- another stream may
// since have been written there -- so every decode checks each block's
// length against the entry and drops the entry on a mismatch.
struct Index {
  u64 start = 0;
  u32 nblocks = 0;
  int type = 0, dims = 0; // type 0: imported, matches any field with nblocks blocks
  size_t n[4] = {0, 0, 0, 0};
  Params z{};
  bool any_params = false;
  bool host = false;      // keyed by a host buffer (freed and reused by its owner, unseen here)
  u64* d_offsets = nullptr;
  u64 stamp = 0;          // last use, for eviction
};
std::map<const void*, Index> g_index;
u64 g_index_clock = 0;

// Bounds of the registry: entries beyond them are evicted, least recently
// used first (a later decode of an evicted stream parses it instead).
// Host stream buffers come and go with their owner, so few of them are kept.
constexpr size_t INDEX_MAX_ENTRIES = 4096;
constexpr size_t INDEX_MAX_HOST = 64;
constexpr u64 INDEX_MAX_BYTES = 1ull << 31;

// one index buffer that left the registry, kept for the next variable-rate
// encode: compressing again and again into one stream buffer then allocates
// no index after its second call
u64* g_spare = nullptr;
u64 g_spare_bytes = 0;

u64 index_bytes(const Index& ix) { return ((u64)ix.nblocks + 1) * sizeof(u64); }

void index_drop(const void* key)
{
  auto it = g_index.find(key);
  if (it == g_index.end())
    return;
  if (it->second.d_offsets)
    index_release(it->second.d_offsets, index_bytes(it->second));
  g_index.erase(it);
}

void index_put(const void* key, const Index& ix)
{
  index_drop(key);
  Index e = ix;
  e.host = !is_device_ptr(key);
  e.stamp = ++g_index_clock;
  g_index[key] = e;
  for (;;) {
    u64 bytes = 0;
    size_t hosts = 0;
    auto old = g_index.end(), old_host = g_index.end();
    for (auto it = g_index.begin(); it != g_index.end(); ++it) {
      bytes += index_bytes(it->second);
      if (it->first == key)
        continue;
      if (old == g_index.end() || it->second.stamp < old->second.stamp)
        old = it;
      if (it->second.host) {
        hosts++;
        if (old_host == g_index.end() || it->second.stamp < old_host->second.stamp)
          old_host = it;
      }
    }
    hosts += e.host ? 1 : 0;
    if (hosts > INDEX_MAX_HOST && old_host != g_index.end())
      index_drop(old_host->first);
    else if ((g_index.size() > INDEX_MAX_ENTRIES || bytes > INDEX_MAX_BYTES) && old != g_index.end())
      index_drop(old->first);
    else
      break;
  }
}

const Index* index_find(const void* key, u64 start, const Desc& d, const size_t* n, const Params& z)
{
  auto it = g_index.find(key);
  if (it == g_index.end())
    return nullptr;
  Index& ix = it->second;
  if (ix.start != start || ix.nblocks != d.nblocks)
    return nullptr;
  if (ix.type) {
    if (ix.type != d.type || ix.dims != d.dims)
      return nullptr;
    for (int i = 0; i < 4; i++)
      if (ix.n[i] != n[i])
        return nullptr;
  }
  if (!ix.any_params && (ix.z.minbits != z.minbits || ix.z.maxbits != z.maxbits ||
                         ix.z.maxprec != z.maxprec || ix.z.minexp != z.minexp))
    return nullptr;
  ix.stamp = ++g_index_clock;
  return &ix;
}

// decode with the registry's index for `key` when one applies; an index
// that disagrees with the stream is dropped and the stream parsed instead
int decode_indexed(const void* key, const Desc& d, void* data, const size_t* n, const Params& z, const u64* in,
                   u64 nwords, u64 offset, u64 start, u64* bits, hipStream_t st)
{
  const Index* ix = fixed_rate(z) || !index_matches_decoder(z) ? nullptr : index_find(key, start, d, n, z);
  int rc = dispatch_decode(d, data, z, in, nwords, offset, ix ? ix->d_offsets : nullptr, bits, st);
  if (rc == ZFP_HIP_ESTALE) {
    index_drop(key);
    rc = dispatch_decode(d, data, z, in, nwords, offset, nullptr, bits, st);
  }
  return rc;
}

int g_dev_state = -1; // -1 unknown, 0 usable, else an error code

} // namespace

int index_buffer(u64 bytes, u64** p)
{
  if (g_spare && g_spare_bytes >= bytes) {
    *p = g_spare;
    g_spare = nullptr;
    g_spare_bytes = 0;
    return ZFP_HIP_OK;
  }
  if (hipMalloc((void**)p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return ZFP_HIP_ENOMEM;
  }
  return ZFP_HIP_OK;
}

// keeps the larger of the released buffer and the spare
void index_release(u64* p, u64 bytes)
{
  if (g_spare && g_spare_bytes >= bytes) {
    (void)hipFree(p);
    return;
  }
  if (g_spare)
    (void)hipFree(g_spare);
  g_spare = p;
  g_spare_bytes = bytes;
}

// called with the library lock held; also clears HIP error state left by
// other users of the runtime in this thread, so that the checks after our
// own calls see only our errors
int device_check()
{
  (void)hipGetLastError();
  if (g_dev_state >= 0) {
    if (g_dev_state)
      record_error(g_dev_state, "no usable gfx950 device");
    return g_dev_state;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1) {
    (void)hipGetLastError();
    g_dev_state = ZFP_HIP_ENODEV;
    return record_error(g_dev_state, "no HIP device");
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    g_dev_state = ZFP_HIP_ENODEV;
    return record_error(g_dev_state, "cannot query the HIP device");
  }
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_dev_state = ZFP_HIP_ENODEV;
    return record_error(g_dev_state, "the current HIP device is not gfx950");
  }
  return g_dev_state = ZFP_HIP_OK;
}

bool is_device_ptr(const void* p)
{
  if (!p)
    return false;
  hipPointerAttribute_t a;
  memset(&a, 0, sizeof(a));
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

namespace {

// lowest and highest element offsets of a field from element 0 (src/zfp.c:18-39)
void field_span(const Desc& d, i64* lo, i64* hi)
{
  i64 a = 0, b = 0;
  for (int i = 0; i < d.dims; i++) {
    i64 e = d.g.s[i] * (i64)(d.g.n[i] - 1);
    if (e < 0)
      a += e;
    else
      b += e;
  }
  *lo = a;
  *hi = b;
}

// worst-case bits of one block (the per-block term of zfp_stream_maximum_size,
// src/zfp.c:712-742)
u64 max_block_bits(const Desc& d, const Params& z)
{
  u32 vals = 1u << (2 * d.dims);
  u32 p = d.esize == 4 ? 32 : 64;
  // block headers: 1 + EBITS, or for reversible blocks 2 + EBITS + PBITS
  // (floats) and PBITS (integers)
  const bool rev = z.minexp < ZFP_MIN_EXP;
  u32 hdr = d.type == zfp_type_float ? (rev ? 15 : 9) : d.type == zfp_type_double ? (rev ? 19 : 12)
          : rev ? (d.esize == 4 ? 5 : 6) : 0;
  u64 mb = hdr + vals - 1 + (u64)vals * (z.maxprec < p ? z.maxprec : p);
  if (mb > z.maxbits)
    mb = z.maxbits;
  if (mb < z.minbits)
    mb = z.minbits;
  return mb;
}

// the reference's bit stream state (include/zfp/bitstream.inl:133-143)
struct StreamState {
  size_t bits;
  u64 buffer;
  u64* ptr;
  u64* begin;
  u64* end;
};

const char* params_problem(const zfp_stream* zs)
{
  if (zs->minbits > zs->maxbits || !zs->maxprec || zs->maxprec > 64)
    return "invalid compression parameters";
  return nullptr;
}

size_t decline(const char* why)
{
  record_error(ZFP_HIP_EINVAL, why);
  return 0;
}

size_t hip_failed(hipError_t e, const char* what)
{
  record_hip_error(e, what);
  return 0;
}

} // namespace
} // namespace zfphip

using namespace zfphip;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int zfp_hip_device_check(void)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  return device_check();
}

extern "C" const char* zfp_hip_version(void)
{
  return "zfp-amd 0.1 (zfp codec 5, gfx950)";
}

extern "C" const char* zfp_hip_last_error(void)
{
  return g_last_error;
}

extern "C" int zfp_hip_encode(int type, const size_t n[4], const ptrdiff_t s[4], const void* d_data,
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
  Params z{minbits, maxbits, maxprec, minexp};
  u64 bits = 0;
  u64* boffs = nullptr;
  rc = dispatch_encode(d, d_data, z, d_words, offset, &bits, (hipStream_t)hip_stream,
                       fixed_rate(z) || !index_matches_decoder(z) ? nullptr : &boffs);
  if (rc)
    return rc;
  if (boffs) {
    Index ix;
    ix.start = offset;
    ix.nblocks = d.nblocks;
    ix.type = type;
    ix.dims = d.dims;
    for (int i = 0; i < 4; i++)
      ix.n[i] = n[i];
    ix.z = z;
    ix.d_offsets = boffs;
    index_put(d_words, ix);
  }
  if (bits_out)
    *bits_out = bits;
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_decode(int type, const size_t n[4], const ptrdiff_t s[4], void* d_data,
                              uint minbits, uint maxbits, uint maxprec, int minexp,
                              const uint64_t* d_words, uint64_t nwords, uint64_t offset, uint64_t* bits_out,
                              void* hip_stream)
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
  if (nwords <= (offset >> 6))
    return record_error(ZFP_HIP_EINVAL, "the stream ends before its first block");
  Params z{minbits, maxbits, maxprec, minexp};
  // reads stop at the caller's nwords words, or at the end of the
  // worst-case stream when that comes first; a stream cut short decodes its
  // missing bits as zeros
  nwords = std::min<u64>(nwords, (offset + (u64)d.nblocks * max_block_bits(d, z) + 63) >> 6);
  u64 bits = 0;
  rc = decode_indexed(d_words, d, d_data, n, z, d_words, nwords, offset, offset, &bits, (hipStream_t)hip_stream);
  if (rc)
    return rc;
  if (bits_out)
    *bits_out = bits;
  return ZFP_HIP_OK;
}

// ---------------------------------------------------------------------------
// block codec for the compressed arrays: count whole blocks stored one after
// another (block k at 4^dims * k values, x fastest).  As a field they are
// an array of extents 4 x ... x 4*count with default strides, whose stream
// order is the order of the blocks.
// ---------------------------------------------------------------------------
namespace {

// slots -> bit offsets (slot * maxbits); flag[0] is set when a block would
// end past the stream's nwords words
__global__ void k_slot_offsets(const u64* __restrict__ slots, u64 n, u64 maxbits, u64 offset, u64 nwords,
                               u64* __restrict__ at, u64* __restrict__ flag)
{
  const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n)
    return;
  const u64 s = slots[i];
  const u64 cap = nwords >= (1ull << 58) ? ~0ull : nwords * 64; // the stream's bits
  // the slots that fit whole after `offset`
  const u64 fit = cap >= offset ? (cap - offset) / maxbits : 0;
  const bool ok = s < fit;
  at[i] = ok ? s * maxbits : 0;
  if (!ok)
    atomicOr((unsigned long long*)flag, 1ull);
}

int block_field(unsigned dims, size_t count, size_t* n)
{
  if (dims < 1 || dims > 4 || count == 0 || count > (1ull << 29))
    return record_error(ZFP_HIP_EINVAL, "block count or dimensionality out of range");
  for (int i = 0; i < 4; i++)
    n[i] = 0;
  for (unsigned i = 0; i + 1 < dims; i++)
    n[i] = 4;
  n[dims - 1] = 4 * count;
  return ZFP_HIP_OK;
}

// device bit offsets of the slots, checked against the stream (lock held)
int slot_offsets(const u64* d_slots, size_t count, const Params& z, u64 offset, u64 nwords, hipStream_t st,
                 const u64** at)
{
  if (!fixed_rate(z) || (z.maxbits & 63) || (offset & 63))
    return record_error(ZFP_HIP_EINVAL, "blocks at chosen slots need fixed rate with word-aligned blocks");
  void* p;
  int rc = g_lens.get((count + 1) * sizeof(u64), &p);
  if (rc)
    return rc;
  u64* flag = (u64*)p;
  u64* off = flag + 1;
  HIPCHK(hipMemsetAsync(flag, 0, sizeof(u64), st));
  hipLaunchKernelGGL(k_slot_offsets, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, d_slots, (u64)count,
                     (u64)z.maxbits, offset, nwords, off, flag);
  HIPCHK(hipGetLastError());
  u64 bad = 0;
  HIPCHK(hipMemcpyAsync(&bad, flag, sizeof(u64), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (bad)
    return record_error(ZFP_HIP_EINVAL, "a slot lies past the end of the stream");
  *at = off;
  return ZFP_HIP_OK;
}

} // namespace

extern "C" int zfp_hip_encode_blocks(int type, unsigned dims, size_t count, const void* d_blocks,
                                     const uint64_t* d_slots, uint minbits, uint maxbits, uint maxprec, int minexp,
                                     uint64_t* d_words, uint64_t nwords, uint64_t offset, uint64_t* bits_out,
                                     void* hip_stream)
{
  size_t n[4];
  const ptrdiff_t s[4] = {0, 0, 0, 0};
  int rc = block_field(dims, count, n);
  if (rc)
    return rc;
  if (!d_slots)
    return zfp_hip_encode(type, n, s, d_blocks, minbits, maxbits, maxprec, minexp, d_words, offset, bits_out,
                          hip_stream);
  std::lock_guard<std::mutex> lock(g_mutex);
  if ((rc = device_check()))
    return rc;
  Desc d;
  if ((rc = make_desc(type, n, s, d)))
    return rc;
  if (minbits > maxbits || !maxprec || maxprec > 64 || !d_blocks || !d_words)
    return record_error(ZFP_HIP_EINVAL, "invalid codec parameters or null pointer");
  Params z{minbits, maxbits, maxprec, minexp};
  const hipStream_t st = (hipStream_t)hip_stream;
  const u64* at = nullptr;
  if ((rc = slot_offsets(d_slots, count, z, offset, nwords, st, &at)))
    return rc;
  u64 bits = 0;
  if ((rc = dispatch_encode(d, d_blocks, z, d_words, offset, &bits, st, nullptr, at)))
    return rc;
  if (bits_out)
    *bits_out = bits;
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_decode_blocks(int type, unsigned dims, size_t count, void* d_blocks,
                                     const uint64_t* d_slots, uint minbits, uint maxbits, uint maxprec, int minexp,
                                     const uint64_t* d_words, uint64_t nwords, uint64_t offset, uint64_t* bits_out,
                                     void* hip_stream)
{
  size_t n[4];
  const ptrdiff_t s[4] = {0, 0, 0, 0};
  int rc = block_field(dims, count, n);
  if (rc)
    return rc;
  if (!d_slots)
    return zfp_hip_decode(type, n, s, d_blocks, minbits, maxbits, maxprec, minexp, d_words, nwords, offset,
                          bits_out, hip_stream);
  std::lock_guard<std::mutex> lock(g_mutex);
  if ((rc = device_check()))
    return rc;
  Desc d;
  if ((rc = make_desc(type, n, s, d)))
    return rc;
  if (minbits > maxbits || !maxprec || maxprec > 64 || !d_blocks || !d_words)
    return record_error(ZFP_HIP_EINVAL, "invalid codec parameters or null pointer");
  Params z{minbits, maxbits, maxprec, minexp};
  const hipStream_t st = (hipStream_t)hip_stream;
  const u64* at = nullptr;
  if ((rc = slot_offsets(d_slots, count, z, offset, nwords, st, &at)))
    return rc;
  u64 bits = 0;
  if ((rc = dispatch_decode(d, d_blocks, z, d_words, nwords, offset, at, &bits, st)))
    return rc;
  if (bits_out)
    *bits_out = bits;
  return ZFP_HIP_OK;
}

// ---------------------------------------------------------------------------
// The block codec for arrays held in host memory (zfp::codec::hip3 and the
// other codecs of include/zfp/hipcodec.hpp): blocks and the stream stay on
// the host; the blocks go to the device as one contiguous run, are coded
// there as a compact fixed-rate stream (block k at k * maxbits), and each
// block's words are moved between that stream and its slot of the caller's.
// ---------------------------------------------------------------------------
namespace {

std::mutex g_host_blocks_mutex;
Scratch g_host_blocks, g_host_words;

int host_blocks_check(int type, unsigned dims, size_t count, const void* blocks, const uint64_t* slots,
                      uint minbits, uint maxbits, int minexp, const void* words, uint64_t nwords, uint64_t offset)
{
  if (type < 1 || type > 4 || dims < 1 || dims > 4 || !count || count > (1ull << 29) || !blocks || !slots || !words)
    return record_error(ZFP_HIP_EINVAL, "invalid block batch");
  if (!fixed_rate(Params{minbits, maxbits, 64, minexp}) || !maxbits || (maxbits & 63) || (offset & 63))
    return record_error(ZFP_HIP_EINVAL, "host block batches need fixed rate with word-aligned blocks");
  const u64 nw = maxbits / 64, w0 = offset / 64;
  for (size_t k = 0; k < count; k++)
    if (slots[k] > (nwords - std::min<u64>(nwords, w0)) / nw || w0 + (slots[k] + 1) * nw > nwords)
      return record_error(ZFP_HIP_EINVAL, "a slot lies past the end of the stream");
  return ZFP_HIP_OK;
}

} // namespace

extern "C" int zfp_hip_encode_blocks_host(int type, unsigned dims, size_t count, const void* blocks,
                                          const uint64_t* slots, uint minbits, uint maxbits, uint maxprec,
                                          int minexp, uint64_t* words, uint64_t nwords, uint64_t offset)
{
  int rc = host_blocks_check(type, dims, count, blocks, slots, minbits, maxbits, minexp, words, nwords, offset);
  if (rc)
    return rc;
  std::lock_guard<std::mutex> lock(g_host_blocks_mutex);
  const size_t esize = type == 1 || type == 3 ? 4 : 8;
  const size_t bbytes = count * ((size_t)1 << (2 * dims)) * esize;
  const u64 nw = maxbits / 64;
  void *db, *dw;
  if ((rc = g_host_blocks.get(bbytes, &db)) || (rc = g_host_words.get(count * nw * 8, &dw)))
    return rc;
  HIPCHK(hipMemcpy(db, blocks, bbytes, hipMemcpyHostToDevice));
  u64 bits = 0;
  if ((rc = zfp_hip_encode_blocks(type, dims, count, db, nullptr, minbits, maxbits, maxprec, minexp, (u64*)dw,
                                  count * nw, 0, &bits, nullptr)))
    return rc;
  std::vector<u64> compact(count * nw);
  HIPCHK(hipMemcpy(compact.data(), dw, count * nw * 8, hipMemcpyDeviceToHost));
  for (size_t k = 0; k < count; k++)
    std::memcpy(words + offset / 64 + slots[k] * nw, compact.data() + k * nw, nw * 8);
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_decode_blocks_host(int type, unsigned dims, size_t count, void* blocks, const uint64_t* slots,
                                          uint minbits, uint maxbits, uint maxprec, int minexp,
                                          const uint64_t* words, uint64_t nwords, uint64_t offset)
{
  int rc = host_blocks_check(type, dims, count, blocks, slots, minbits, maxbits, minexp, words, nwords, offset);
  if (rc)
    return rc;
  std::lock_guard<std::mutex> lock(g_host_blocks_mutex);
  const size_t esize = type == 1 || type == 3 ? 4 : 8;
  const size_t bbytes = count * ((size_t)1 << (2 * dims)) * esize;
  const u64 nw = maxbits / 64;
  void *db, *dw;
  if ((rc = g_host_blocks.get(bbytes, &db)) || (rc = g_host_words.get(count * nw * 8, &dw)))
    return rc;
  std::vector<u64> compact(count * nw);
  for (size_t k = 0; k < count; k++)
    std::memcpy(compact.data() + k * nw, words + offset / 64 + slots[k] * nw, nw * 8);
  HIPCHK(hipMemcpy(dw, compact.data(), count * nw * 8, hipMemcpyHostToDevice));
  u64 bits = 0;
  if ((rc = zfp_hip_decode_blocks(type, dims, count, db, nullptr, minbits, maxbits, maxprec, minexp, (u64*)dw,
                                  count * nw, 0, &bits, nullptr)))
    return rc;
  HIPCHK(hipMemcpy(blocks, db, bbytes, hipMemcpyDeviceToHost));
  return ZFP_HIP_OK;
}

extern "C" size_t zfp_hip_compress(zfp_stream* zs, const zfp_field* field)
{
  if (!zs || !zs->stream || !field)
    return decline("null stream or field");
  if (const char* why = params_problem(zs))
    return decline(why);
  size_t n[4] = {field->nx, field->ny, field->nz, field->nw};
  ptrdiff_t s[4] = {field->sx, field->sy, field->sz, field->sw};
  std::lock_guard<std::mutex> lock(g_mutex);
  if (device_check())
    return 0;
  Desc d;
  if (make_desc(field->type, n, s, d))
    return 0;
  if (!field->data)
    return decline("field has no data pointer");
  StreamState* bs = (StreamState*)zs->stream;
  // write position (include/zfp/bitstream.inl:323-327): whole words already
  // in memory plus `lead` bits still buffered in the stream state
  const u64 word0 = (u64)(bs->ptr - bs->begin);
  const u32 lead = (u32)bs->bits;
  Params z{zs->minbits, zs->maxbits, zs->maxprec, zs->minexp};
  hipStream_t st = nullptr;
  hipError_t e;

  // field on the device (host fields are staged by their memory span)
  const void* d_data = field->data;
  if (!is_device_ptr(field->data)) {
    i64 lo, hi;
    field_span(d, &lo, &hi);
    size_t bytes = (size_t)(hi - lo + 1) * d.esize;
    void* p;
    if (g_field.get(bytes, &p))
      return 0;
    const char* hbegin = (const char*)field->data + lo * (i64)d.esize;
    if ((e = hipMemcpy(p, hbegin, bytes, hipMemcpyHostToDevice)) != hipSuccess)
      return hip_failed(e, "copy of the field to the device");
    d_data = (const char*)p - lo * (i64)d.esize;
  }

  // output: words from the current one, whose low `lead` bits are the
  // buffered ones
  const bool dev_stream = is_device_ptr(bs->begin);
  const u64 cap_words = ((u64)lead + (u64)d.nblocks * max_block_bits(d, z) + 63) / 64 + 1;
  u64* d_out;
  if (dev_stream) {
    d_out = bs->ptr;
  }
  else {
    void* p;
    if (g_stream.get(cap_words * sizeof(u64), &p))
      return 0;
    d_out = (u64*)p;
  }
  if ((e = hipMemcpy(d_out, &bs->buffer, sizeof(u64), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_failed(e, "store of the buffered stream bits");
  u64 bits = 0;
  u64* boffs = nullptr;
  const bool fixed = fixed_rate(z);
  if (dispatch_encode(d, d_data, z, d_out, lead, &bits, st, fixed || !index_matches_decoder(z) ? nullptr : &boffs))
    return 0;
  const u64 words = (lead + bits + 63) / 64;
  if (!dev_stream && (e = hipMemcpy(bs->ptr, d_out, words * sizeof(u64), hipMemcpyDeviceToHost)) != hipSuccess) {
    if (boffs)
      index_release(boffs, ((u64)d.nblocks + 1) * sizeof(u64));
    return hip_failed(e, "copy of the stream to the host");
  }
  if (boffs) {
    Index ix;
    ix.start = word0 * 64 + lead;
    ix.nblocks = d.nblocks;
    ix.type = field->type;
    ix.dims = d.dims;
    for (int i = 0; i < 4; i++)
      ix.n[i] = n[i];
    ix.z = z;
    ix.d_offsets = boffs;
    index_put(bs->begin, ix);
  }
  // leave the stream as serial compression plus stream_flush leaves it
  bs->ptr += words;
  bs->bits = 0;
  bs->buffer = 0;
  g_last_error[0] = 0;
  return (size_t)(bs->ptr - bs->begin) * sizeof(u64);
}

extern "C" size_t zfp_hip_decompress(zfp_stream* zs, zfp_field* field)
{
  if (!zs || !zs->stream || !field)
    return decline("null stream or field");
  if (const char* why = params_problem(zs))
    return decline(why);
  size_t n[4] = {field->nx, field->ny, field->nz, field->nw};
  ptrdiff_t s[4] = {field->sx, field->sy, field->sz, field->sw};
  std::lock_guard<std::mutex> lock(g_mutex);
  if (device_check())
    return 0;
  Desc d;
  if (make_desc(field->type, n, s, d))
    return 0;
  if (!field->data)
    return decline("field has no data pointer");
  StreamState* bs = (StreamState*)zs->stream;
  // read position (include/zfp/bitstream.inl:316-320)
  const u64 rpos = (u64)(bs->ptr - bs->begin) * 64 - bs->bits;
  const u64 word0 = rpos >> 6;
  const u32 lead = (u32)(rpos & 63);
  Params z{zs->minbits, zs->maxbits, zs->maxprec, zs->minexp};
  hipStream_t st = nullptr;
  hipError_t e;

  // readable words: up to the end of the stream buffer, or of the stream
  u64 avail = bs->end > bs->begin + word0 ? (u64)(bs->end - bs->begin) - word0 : 0;
  u64 need = std::min<u64>(avail, (lead + (u64)d.nblocks * max_block_bits(d, z) + 63) / 64);
  const u64* d_in;
  if (is_device_ptr(bs->begin)) {
    d_in = bs->begin + word0;
  }
  else {
    void* p;
    if (g_stream.get((need ? need : 1) * sizeof(u64), &p))
      return 0;
    if (need && (e = hipMemcpy(p, bs->begin + word0, need * sizeof(u64), hipMemcpyHostToDevice)) != hipSuccess)
      return hip_failed(e, "copy of the stream to the device");
    d_in = (const u64*)p;
  }

  void* d_data = field->data;
  char* hbegin = nullptr;
  size_t span_bytes = 0;
  void* dbuf = nullptr;
  if (!is_device_ptr(field->data)) {
    i64 lo, hi;
    field_span(d, &lo, &hi);
    span_bytes = (size_t)(hi - lo + 1) * d.esize;
    if (g_field.get(span_bytes, &dbuf))
      return 0;
    hbegin = (char*)field->data + lo * (i64)d.esize;
    // a strided host layout with gaps is copied in first so the gaps survive
    // the copy back
    u64 values = (u64)d.g.n[0] * d.g.n[1] * d.g.n[2] * d.g.n[3];
    if (values * d.esize != span_bytes &&
        (e = hipMemcpy(dbuf, hbegin, span_bytes, hipMemcpyHostToDevice)) != hipSuccess)
      return hip_failed(e, "copy of the field to the device");
    d_data = (char*)dbuf - lo * (i64)d.esize;
  }
  u64 bits = 0;
  if (decode_indexed(bs->begin, d, d_data, n, z, d_in, need, lead, rpos, &bits, st))
    return 0;
  if (hbegin && (e = hipMemcpy(hbegin, dbuf, span_bytes, hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_failed(e, "copy of the field to the host");
  // leave the stream as serial decompression plus stream_align leaves it
  const u64 end = rpos + bits;
  bs->ptr = bs->begin + (end + 63) / 64;
  bs->bits = 0;
  bs->buffer = 0;
  g_last_error[0] = 0;
  return (size_t)(bs->ptr - bs->begin) * sizeof(u64);
}

extern "C" size_t zfp_hip_index_size(const void* stream_begin)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  auto it = g_index.find(stream_begin);
  return it == g_index.end() ? 0 : (size_t)it->second.nblocks + 1;
}

extern "C" int zfp_hip_index_export(const void* stream_begin, uint64_t* offsets, size_t count)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  auto it = g_index.find(stream_begin);
  if (it == g_index.end())
    return record_error(ZFP_HIP_ENOINDEX, "no block index for this stream");
  const Index& ix = it->second;
  if (!offsets || count < (size_t)ix.nblocks + 1)
    return record_error(ZFP_HIP_EINVAL, "offset array too small");
  std::vector<u64> tmp((size_t)ix.nblocks + 1);
  HIPCHK(hipMemcpy(tmp.data(), ix.d_offsets, tmp.size() * sizeof(u64), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); i++)
    offsets[i] = ix.start + tmp[i];
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_index_import(const void* stream_begin, const uint64_t* offsets, size_t count)
{
  if (!stream_begin || !offsets || count < 2 || count - 1 > 0xffffffffull)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  for (size_t i = 1; i < count; i++)
    if (offsets[i] < offsets[i - 1])
      return record_error(ZFP_HIP_EINVAL, "index offsets decrease");
  std::lock_guard<std::mutex> lock(g_mutex);
  int rc = device_check();
  if (rc)
    return rc;
  std::vector<u64> rel(count);
  for (size_t i = 0; i < count; i++)
    rel[i] = offsets[i] - offsets[0];
  u64* dev = nullptr;
  HIPCHK(hipMalloc(&dev, count * sizeof(u64)));
  hipError_t e = hipMemcpy(dev, rel.data(), count * sizeof(u64), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    return record_hip_error(e, "copy of the index to the device");
  }
  Index ix;
  ix.start = offsets[0];
  ix.nblocks = (u32)(count - 1);
  ix.type = 0; // matches any field with this many blocks
  ix.any_params = true;
  ix.d_offsets = dev;
  index_put(stream_begin, ix);
  return ZFP_HIP_OK;
}

// ---------------------------------------------------------------------------
// the reference's hybrid4 block index (include/zfp/index.hpp:152-310): per
// chunk of 4 blocks a 32-bit base (offset of the chunk's first block >> 12)
// and four 16-bit offsets from base << 12; the last chunk is padded with
// zero-length blocks; blocks of at most ZFP_MAX_BITS bits, offsets below
// 2^44.  Offsets count from the first block.
// ---------------------------------------------------------------------------
namespace {
struct Hybrid4Record {
  uint32_t hi;
  uint16_t lo[4];
};
static_assert(sizeof(Hybrid4Record) == 12, "hybrid4 record layout");
} // namespace

extern "C" int zfp_hip_hybrid4_from_offsets(const uint64_t* offsets, size_t nblocks, void* records)
{
  if (!offsets || !records || !nblocks)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  const u64 base0 = offsets[0];
  Hybrid4Record* r = (Hybrid4Record*)records;
  for (size_t b = 0; b < nblocks; b++)
    if (offsets[b + 1] < offsets[b] || offsets[b + 1] - offsets[b] > ZFP_MAX_BITS)
      return record_error(ZFP_HIP_EINVAL, "block size out of range for a hybrid4 index");
  for (size_t c = 0; c < (nblocks + 3) / 4; c++) {
    const u64 ptr = offsets[4 * c] - base0;
    if (ptr >> 44)
      return record_error(ZFP_HIP_EINVAL, "block offset too large for a hybrid4 index");
    r[c].hi = (uint32_t)(ptr >> 12);
    const u64 base = (u64)r[c].hi << 12;
    for (int k = 0; k < 4; k++) {
      const size_t b = 4 * c + k;
      r[c].lo[k] = (uint16_t)(offsets[b < nblocks ? b : nblocks] - base0 - base);
    }
  }
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_offsets_from_hybrid4(const void* records, size_t nblocks, uint64_t end_bits,
                                            uint64_t* offsets)
{
  if (!offsets || !records || !nblocks)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  const Hybrid4Record* r = (const Hybrid4Record*)records;
  for (size_t b = 0; b < nblocks; b++)
    offsets[b] = ((u64)r[b / 4].hi << 12) + r[b / 4].lo[b % 4];
  offsets[nblocks] = end_bits;
  for (size_t b = 0; b < nblocks; b++)
    if (offsets[b + 1] < offsets[b])
      return record_error(ZFP_HIP_EINVAL, "hybrid4 offsets decrease");
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_index_export_hybrid4(const void* stream_begin, void* records, size_t nrecords,
                                            uint64_t* start, uint64_t* end_bits)
{
  const size_t count = zfp_hip_index_size(stream_begin);
  if (!count)
    return record_error(ZFP_HIP_ENOINDEX, "no block index for this stream");
  if (!records || nrecords < (count - 1 + 3) / 4)
    return record_error(ZFP_HIP_EINVAL, "record array too small");
  std::vector<uint64_t> off(count);
  int rc = zfp_hip_index_export(stream_begin, off.data(), count);
  if (rc)
    return rc;
  if ((rc = zfp_hip_hybrid4_from_offsets(off.data(), count - 1, records)))
    return rc;
  if (start)
    *start = off[0];
  if (end_bits)
    *end_bits = off[count - 1] - off[0];
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_index_import_hybrid4(const void* stream_begin, uint64_t start, const void* records,
                                            size_t nblocks, uint64_t end_bits)
{
  if (!nblocks || nblocks > 0xffffffffull)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  std::vector<uint64_t> off(nblocks + 1);
  int rc = zfp_hip_offsets_from_hybrid4(records, nblocks, end_bits, off.data());
  if (rc)
    return rc;
  for (auto& o : off)
    o += start;
  return zfp_hip_index_import(stream_begin, off.data(), off.size());
}

// ---------------------------------------------------------------------------
// the reference's hybrid8<dims> block index (include/zfp/index.hpp:312-...):
// per chunk of 8 blocks two words.  A block size is split into its low 8
// bits and hb = 2 (dims - 1) high bits; so is the chunk's first offset.
// Word 0 holds the offset's high part above 7 hb-bit fields that hold the
// high bits of sizes 0..6; word 1 holds the offset's low 8 bits above seven
// 8-bit fields with the low bits of sizes 0..6.  Block k's offset is the
// chunk's offset plus sizes 0..k-1, i.e. the sums of the top k+1 fields of
// each word.  Sizes must be below 2^(8 + hb), the chunk offset's high part
// must fit its field.  Offsets count from the first block.
// ---------------------------------------------------------------------------
extern "C" int zfp_hip_hybrid8_from_offsets(unsigned dims, const uint64_t* offsets, size_t nblocks, uint64_t* words)
{
  if (!offsets || !words || !nblocks || dims < 1 || dims > 4)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  const unsigned hb = 2 * (dims - 1);
  const u64 b0 = offsets[0];
  for (size_t b = 0; b < nblocks; b++)
    if (offsets[b + 1] < offsets[b] || ((offsets[b + 1] - offsets[b]) >> (hb + 8)))
      return record_error(ZFP_HIP_EINVAL, "block size out of range for a hybrid8 index");
  for (size_t c = 0; c < (nblocks + 7) / 8; c++) {
    const u64 ptr = offsets[8 * c] - b0;
    const u64 h = ptr >> 8;
    u64 hi = h << (7 * hb);
    if ((hi >> (7 * hb)) != h)
      return record_error(ZFP_HIP_EINVAL, "block offset too large for a hybrid8 index");
    u64 lo = (ptr & 0xff) << 56;
    for (unsigned k = 0; k < 7; k++) {
      const size_t b = 8 * c + k;
      const u64 size = b < nblocks ? offsets[b + 1] - offsets[b] : 0;
      hi += (size >> 8) << ((6 - k) * hb);
      lo += (size & 0xff) << ((6 - k) * 8);
    }
    words[2 * c] = hi;
    words[2 * c + 1] = lo;
  }
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_offsets_from_hybrid8(unsigned dims, const uint64_t* words, size_t nblocks, uint64_t end_bits,
                                            uint64_t* offsets)
{
  if (!offsets || !words || !nblocks || dims < 1 || dims > 4)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  const unsigned hb = 2 * (dims - 1);
  const u64 hmask = hb ? (1ull << hb) - 1 : 0;
  for (size_t b = 0; b < nblocks; b++) {
    const size_t c = b / 8;
    const unsigned k = (unsigned)(b % 8);
    const u64 hi = words[2 * c], lo = words[2 * c + 1];
    const u64 base = hb ? hi >> (8 * hb) : hi; // the offset's high part above field 7
    u64 hs = 0, ls = 0;
    for (unsigned f = 7 - k; f <= 7; f++) {
      hs += hb ? (hi >> (f * hb)) & hmask : 0;
      ls += (lo >> (8 * f)) & 0xff;
    }
    offsets[b] = (((base << hb) + hs) << 8) + ls;
  }
  offsets[nblocks] = end_bits;
  for (size_t b = 0; b < nblocks; b++)
    if (offsets[b + 1] < offsets[b])
      return record_error(ZFP_HIP_EINVAL, "hybrid8 offsets decrease");
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_index_export_hybrid8(const void* stream_begin, unsigned dims, uint64_t* words,
                                            size_t nwords, uint64_t* start, uint64_t* end_bits)
{
  const size_t count = zfp_hip_index_size(stream_begin);
  if (!count)
    return record_error(ZFP_HIP_ENOINDEX, "no block index for this stream");
  if (!words || nwords < 2 * ((count - 1 + 7) / 8))
    return record_error(ZFP_HIP_EINVAL, "word array too small");
  std::vector<uint64_t> off(count);
  int rc = zfp_hip_index_export(stream_begin, off.data(), count);
  if (rc)
    return rc;
  if ((rc = zfp_hip_hybrid8_from_offsets(dims, off.data(), count - 1, words)))
    return rc;
  if (start)
    *start = off[0];
  if (end_bits)
    *end_bits = off[count - 1] - off[0];
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_index_import_hybrid8(const void* stream_begin, unsigned dims, uint64_t start,
                                            const uint64_t* words, size_t nblocks, uint64_t end_bits)
{
  if (!nblocks || nblocks > 0xffffffffull)
    return record_error(ZFP_HIP_EINVAL, "invalid index");
  std::vector<uint64_t> off(nblocks + 1);
  int rc = zfp_hip_offsets_from_hybrid8(dims, words, nblocks, end_bits, off.data());
  if (rc)
    return rc;
  for (auto& o : off)
    o += start;
  return zfp_hip_index_import(stream_begin, off.data(), off.size());
}

extern "C" void zfp_hip_index_clear(void)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  for (auto& kv : g_index)
    if (kv.second.d_offsets)
      (void)hipFree(kv.second.d_offsets);
  g_index.clear();
  if (g_spare)
    (void)hipFree(g_spare);
  g_spare = nullptr;
  g_spare_bytes = 0;
  // the codec's grow-only scratch (lengths, scan tiles, flags, staged bits)
  for (Scratch* sc : {&g_lens, &g_scan_tmp, &g_flag, &g_stage})
    sc->release();
}
