// helpers.hip -- small kernels shared by every kernel family: zeroing of the
// stream words that blocks OR into, and the exclusive prefix sum that turns
// per-block lengths into stream offsets.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_scan.hpp>

#include "backend.h"

namespace zfphip {

// zero words [w0, w1) of `out`, keeping the low `keep` bits of word w0
__global__ void k_prepare(u64* out, u64 w0, u64 w1, u32 keep)
{
  u64 i = w0 + (u64)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w1)
    return;
  if (i == w0 && keep)
    out[i] &= lowmask(keep);
  else
    out[i] = 0;
}

// the same for the words that `*bits` bits from bit `offset` on cover, with
// the bit count read on the device (grid-stride loop)
__global__ void k_prepare_dev(u64* out, u64 offset, const u64* bits)
{
  const u64 w0 = offset >> 6, w1 = (offset + *bits + 63) >> 6;
  const u32 keep = (u32)(offset & 63);
  for (u64 i = w0 + (u64)blockIdx.x * blockDim.x + threadIdx.x; i < w1; i += (u64)gridDim.x * blockDim.x) {
    if (i == w0 && keep)
      out[i] &= lowmask(keep);
    else
      out[i] = 0;
  }
}


// ---------------------------------------------------------------------------
// exclusive prefix sum of u64 block lengths (three passes: per-tile scan,
// serial scan of tile totals in one workgroup, add tile prefixes)
// ---------------------------------------------------------------------------
constexpr int SCAN_BS = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_BS * SCAN_ITEMS;

__device__ __forceinline__ u64 wg_exclusive_scan(u64 v, u64* sh, u64* total)
{
  // inclusive Hillis-Steele scan over the workgroup in LDS
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < SCAN_BS; d <<= 1) {
    u64 a = t >= d ? sh[t - d] : 0ull;
    __syncthreads();
    sh[t] += a;
    __syncthreads();
  }
  u64 inc = sh[t];
  *total = sh[SCAN_BS - 1];
  __syncthreads();
  return inc - v;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_tiles(const u64* in, u64* out, u64 n, u64* tile_sums)
{
  __shared__ u64 sh[SCAN_BS];
  u64 base = (u64)blockIdx.x * SCAN_TILE + (u64)threadIdx.x * SCAN_ITEMS;
  u64 v[SCAN_ITEMS];
  u64 sum = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    v[i] = base + i < n ? in[base + i] : 0ull;
    sum += v[i];
  }
  u64 total;
  u64 run = wg_exclusive_scan(sum, sh, &total);
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    if (base + i < n)
      out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0)
    tile_sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_BS) k_scan_sums(u64* sums, u64 ntiles)
{
  __shared__ u64 sh[SCAN_BS];
  u64 carry = 0;
  for (u64 c = 0; c < ntiles; c += SCAN_BS) {
    u64 i = c + threadIdx.x;
    u64 v = i < ntiles ? sums[i] : 0ull;
    u64 total;
    u64 ex = wg_exclusive_scan(v, sh, &total);
    if (i < ntiles)
      sums[i] = carry + ex;
    carry += total;
  }
}

// every value of a tile gets the same prefix, so lanes take consecutive
// words (coalesced) rather than the scan's per-thread runs
__global__ void __launch_bounds__(SCAN_BS) k_scan_add(u64* out, u64 n, const u64* tile_sums)
{
  u64 base = (u64)blockIdx.x * SCAN_TILE + threadIdx.x;
  u64 add = tile_sums[blockIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++)
    if (base + (u64)i * SCAN_BS < n)
      out[base + (u64)i * SCAN_BS] += add;
}


u64 scan_tmp_words(u64 n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// one pass over the data (rocPRIM's decoupled look-back scan) with its
// temporary storage from the grow-only scratch; exclusive_scan_tmp below is
// the allocation-free variant the batch graphs capture.  Tiles of 256 x 40
// values: on 4.2M lengths (a 2D 8191 x 8193 field) 23.5 us on MI355X against
// 44 us with rocPRIM's default tiles, whose longer look-back chain dominates.
using ScanTiles = rocprim::scan_config<256, 40, rocprim::block_load_method::block_load_transpose,
                                       rocprim::block_store_method::block_store_transpose,
                                       rocprim::block_scan_algorithm::using_warp_scan>;

int exclusive_scan(const u64* in, u64* out, u64 n, hipStream_t st)
{
  size_t bytes = 0;
  HIPCHK(rocprim::exclusive_scan<ScanTiles>(nullptr, bytes, in, out, (u64)0, (size_t)n, rocprim::plus<u64>(), st));
  void* tp;
  int rc = g_scan_tmp.get(bytes ? bytes : 8, &tp);
  if (rc)
    return rc;
  HIPCHK(rocprim::exclusive_scan<ScanTiles>(tp, bytes, in, out, (u64)0, (size_t)n, rocprim::plus<u64>(), st));
  return ZFP_HIP_OK;
}

int exclusive_scan_tmp(const u64* in, u64* out, u64 n, u64* sums, hipStream_t st)
{
  const u64 ntiles = scan_tmp_words(n);
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(SCAN_BS), 0, st, in, out, n, sums);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_BS), 0, st, sums, ntiles);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ntiles), dim3(SCAN_BS), 0, st, out, n, (const u64*)sums);
  HIPCHK(hipGetLastError());
  return ZFP_HIP_OK;
}


} // namespace zfphip
