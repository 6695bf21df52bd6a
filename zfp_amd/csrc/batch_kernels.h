// batch_kernels.h -- device side of the batched codec: many fields of mixed
// type (int32, int64, float, double), dimensionality (1-4), size and mode
// (reversible included) go through persistent launches, one per pass and
// family.
//
// Work is a list of tasks (job, first lane); each wave takes tasks in turn
// (task = wave id, + number of waves, ...) and runs up to blocks_per_task
// blocks of that job, 64 at a time, one block per lane, with the same
// per-block codec as the single-field kernels.  The job's kind selects the
// codec instantiation; it is uniform across the wave, so the switch costs a
// scalar branch.
//
// Families: 1-3D lossy jobs (12 kinds) run in one kernel; 4D and reversible
// jobs (20 kinds: 4D lossy per type, reversible per type and
// dimensionality) in a second.  The second family's blocks keep large
// lane-private arrays, whose registers and scratch would otherwise be the
// first family's too, since a kernel is allocated for its largest case.
//
// Passes: LEN (block lengths of variable-rate jobs), then an exclusive scan
// of those lengths, ENC (every job writes its blocks: fixed rate at
// b * maxbits, variable rate at its scanned offset) and, for decompression,
// DEC.  ENC and DEC stage the blocks of fixed-rate jobs with word-aligned
// blocks in stream order through LDS, one slot per lane, as the single-field
// kernels do: the wave's 64 blocks are one run of stream words, which ENC
// codes into the slots (the straight-line slot coder) and then stores with
// coalesced stores, and DEC copies in with coalesced loads and decodes from
// the slots by bit index.
#pragma once

#include <hip/hip_runtime.h>

#include "backend.h"
#include "codec_block.h"
#include "codec_device.h"
#include "codec_kernels.h"

namespace zfphip {

enum { BK_LEN = 0, BK_ENC = 1, BK_DEC = 2 };
enum { BM_FIXED = 0, BM_VAR = 1 };

// one field of a batch, as the device sees it
struct BJob {
  Geom g;
  Params z;
  void* data;    // field, element 0
  u64* words;    // stream words
  u64 offset;    // bit offset of the field's first block in `words`
  u64* index;    // variable rate: nblocks + 1 block offsets relative to `offset`
  u64 vbase;     // slot of the first block in the batch's length/offset arrays
  u64 nwords;    // readable words of `words` (decompression bound)
  u32 nblocks;
  u32 kind;      // codec within the family (job_kind)
  u32 mode;      // BM_FIXED or BM_VAR (reversible jobs are variable-rate)
  u32 dir;       // 0 compression, 1 decompression
  u32 type;      // zfp_type (1 int32, 2 int64, 3 float, 4 double)
  u32 dims;      // 1-4
  u32 wide;      // family: 0 for 1-3D lossy, 1 for 4D or reversible
  u32 rev;       // reversible mode
  u32 stage;     // ENC or DEC stages the blocks through LDS slots (fixed
                 // rate, word-aligned blocks of at most BATCH_STAGE_WORDS
                 // words, lanes in stream order, 1-3D lossy)
};

// the longest blocks, in words, that ENC and DEC stage through LDS: a
// workgroup of 256 lanes then needs at most 256 x 19 words (38 KiB)
constexpr u32 BATCH_STAGE_WORDS = 16;

// the kind of a job within its family
inline u32 job_kind(u32 type, u32 dims, bool rev)
{
  if (!rev && dims < 4)
    return 3u * (type - 1) + (dims - 1);
  return rev ? 4u + 4u * (type - 1) + (dims - 1) : type - 1;
}

// slot: the lane's LDS slot (staged jobs): ENC codes the block into it
// (zeroed, BATCH_STAGE_WORDS + 2 words at most), DEC reads the block from
// it; null otherwise
template <typename S, int D, int PASS, bool REV = false>
__device__ __forceinline__ void batch_one(const BJob& J, u32 t, u64* lens, const u64* offs, u64* slot)
{
  u32 bc[4] = {0, 0, 0, 0};
  const u32 b = lane_block<D>(t, J.g, bc);
  if constexpr (PASS == BK_LEN) {
    BitCounter c;
    c.pos = 0;
    if constexpr (REV)
      lens[J.vbase + b] = encode_block_rev<S, D>(c, (const S*)J.data, J.g, bc, J.z);
    else
      lens[J.vbase + b] = encode_block<S, D>(c, (const S*)J.data, J.g, bc, J.z);
  }
  else if constexpr (PASS == BK_ENC) {
    if (!REV && slot) {
      SlotWriter w;
      w.begin(slot, (J.z.maxbits >> 6) + 2);
      encode_block<S, D>(w, (const S*)J.data, J.g, bc, J.z);
      return;
    }
    u64 s0, s1;
    if (J.mode == BM_FIXED) {
      s0 = (u64)b * J.z.maxbits;
      s1 = s0 + J.z.maxbits;
    }
    else {
      const u64 base = offs[J.vbase];
      s0 = offs[J.vbase + b] - base;
      s1 = offs[J.vbase + b + 1] - base;
      if (J.index)
        J.index[b] = s0;
    }
    BitWriter<false> w;
    w.begin(J.words, J.offset + s0, J.offset + s1);
    if constexpr (REV)
      encode_block_rev<S, D>(w, (const S*)J.data, J.g, bc, J.z);
    else
      encode_block<S, D>(w, (const S*)J.data, J.g, bc, J.z);
    w.finish();
  }
  else if (!REV && slot) {
    SlotReader r;
    r.begin(slot, 0);
    decode_block<S, D, true>(r, (S*)J.data, J.g, bc, J.z);
  }
  else {
    const u64 s0 = J.mode == BM_FIXED ? (u64)b * J.z.maxbits : J.index[b];
    BitReader r;
    r.begin(J.words, J.nwords, J.offset + s0);
    if constexpr (REV)
      decode_block_rev<S, D, true>(r, (S*)J.data, J.g, bc, J.z);
    else
      decode_block<S, D, true>(r, (S*)J.data, J.g, bc, J.z);
  }
}

// 4D lossy (kinds 0-3, by type) and reversible (4 + 4 * (type - 1) +
// dims - 1) jobs
template <int PASS, typename S>
__device__ __forceinline__ void batch_wide_type(const BJob& J, u32 t, u64* lens, const u64* offs)
{
  if (!J.rev) {
    batch_one<S, 4, PASS>(J, t, lens, offs, nullptr);
    return;
  }
  switch (J.dims) {
    case 1: batch_one<S, 1, PASS, true>(J, t, lens, offs, nullptr); break;
    case 2: batch_one<S, 2, PASS, true>(J, t, lens, offs, nullptr); break;
    case 3: batch_one<S, 3, PASS, true>(J, t, lens, offs, nullptr); break;
    default: batch_one<S, 4, PASS, true>(J, t, lens, offs, nullptr); break;
  }
}

template <int PASS, bool WIDE>
__device__ __forceinline__ void batch_dispatch(const BJob& J, u32 t, u64* lens, const u64* offs, u64* slot)
{
  if constexpr (WIDE) {
    switch (J.type) {
      case 1: batch_wide_type<PASS, i32>(J, t, lens, offs); break;
      case 2: batch_wide_type<PASS, i64>(J, t, lens, offs); break;
      case 3: batch_wide_type<PASS, float>(J, t, lens, offs); break;
      default: batch_wide_type<PASS, double>(J, t, lens, offs); break;
    }
    return;
  }
  switch (J.kind) {
    case 0: batch_one<i32, 1, PASS>(J, t, lens, offs, slot); break;
    case 1: batch_one<i32, 2, PASS>(J, t, lens, offs, slot); break;
    case 2: batch_one<i32, 3, PASS>(J, t, lens, offs, slot); break;
    case 3: batch_one<i64, 1, PASS>(J, t, lens, offs, slot); break;
    case 4: batch_one<i64, 2, PASS>(J, t, lens, offs, slot); break;
    case 5: batch_one<i64, 3, PASS>(J, t, lens, offs, slot); break;
    case 6: batch_one<float, 1, PASS>(J, t, lens, offs, slot); break;
    case 7: batch_one<float, 2, PASS>(J, t, lens, offs, slot); break;
    case 8: batch_one<float, 3, PASS>(J, t, lens, offs, slot); break;
    case 9: batch_one<double, 1, PASS>(J, t, lens, offs, slot); break;
    case 10: batch_one<double, 2, PASS>(J, t, lens, offs, slot); break;
    default: batch_one<double, 3, PASS>(J, t, lens, offs, slot); break;
  }
}

// persistent kernel of one pass and family; tasks[i] = {job, first lane}.
// stride (ENC and DEC of the 1-3D family): LDS slot stride in words, 0 when
// no job is staged; the launch gives blockDim.x * stride words of LDS.
template <int PASS, bool WIDE>
__global__ void __launch_bounds__(256, 2)
k_batch(const BJob* __restrict__ jobs, const uint2* __restrict__ tasks, u32 ntasks, u32 per_task,
        u64* __restrict__ lens, const u64* __restrict__ offs, u32 stride)
{
  extern __shared__ u64 kslots[];
  const u32 lane = threadIdx.x & 63;
  const u32 wave = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const u32 nwaves = (gridDim.x * blockDim.x) >> 6;
  for (u32 task = wave; task < ntasks; task += nwaves) {
    const uint2 tk = tasks[task];
    const BJob& J = jobs[tk.x];
    const u32 end = umin(tk.y + per_task, J.nblocks);
    // staged job: the wave's blocks t0 .. t0 + cnt - 1 of each chunk are
    // one run of stream words from (offset / 64) + t0 * nw
    const bool staged = PASS != BK_LEN && !WIDE && stride && J.stage;
    const u32 nw = J.z.maxbits >> 6;
    u64* wslots = kslots + (u64)(threadIdx.x - lane) * stride;
    for (u32 t0 = tk.y; t0 < end; t0 += 64) {
      u64* slot = nullptr;
      const u32 cnt = umin(64u, end - t0);
      if (staged) {
        // the previous chunk's slots are read before they are overwritten
        wave_lds_sync();
        slot = wslots + (u64)lane * stride;
        if constexpr (PASS == BK_DEC) {
          wave_copy<false, true>(J.words, J.nwords, wslots, stride, nw, cnt, (J.offset >> 6) + (u64)t0 * nw,
                                 nullptr);
          wave_lds_sync();
        }
        else {
          for (u32 i = 0; i < nw + 2; i++)
            slot[i] = 0;
        }
      }
      const u32 t = t0 + lane;
      if (t < end)
        batch_dispatch<PASS, WIDE>(J, t, lens, offs, slot);
      if constexpr (PASS == BK_ENC) {
        if (staged) {
          wave_lds_sync();
          wave_copy<true, true>(J.words, 0, wslots, stride, nw, cnt, (J.offset >> 6) + (u64)t0 * nw, nullptr);
        }
      }
    }
  }
}

// host side: launch one pass of one family.  Each pass and family is
// instantiated in its own translation unit (batch_{len,enc,dec}.hip and
// batch_wide_{len,enc,dec}.hip) so that they compile in parallel.
template <int PASS, bool WIDE>
int launch_batch_pass(unsigned grid, unsigned wg, hipStream_t st, const BJob* jobs, const uint2* tasks,
                      u32 ntasks, u32 per_task, u64* lens, const u64* offs, u32 stride)
{
  if (PASS == BK_LEN || WIDE)
    stride = 0;
  hipLaunchKernelGGL((k_batch<PASS, WIDE>), dim3(grid), dim3(wg), (size_t)wg * stride * 8, st, jobs, tasks, ntasks,
                     per_task, lens, offs, stride);
  HIPCHK(hipGetLastError());
  return ZFP_HIP_OK;
}

// checks a job and fills its device descriptor; variable-rate compression
// jobs take the next nblocks slots of the length/offset arrays from vbase
int job_to_device(const zfp_hip_job& in, int dir, BJob& J, u64& vbase);

// zero the stream words of the listed jobs that blocks OR into
// (batch.hip); offs holds the scanned offsets of variable-rate jobs
int launch_batch_prepare(unsigned grid, hipStream_t st, const BJob* jobs, const u32* pjobs, u32 npjobs,
                         const u64* offs);
// bits[j]: bits written (compression jobs) or consumed (decompression
// jobs) by job j; completes the index of variable-rate compression jobs
int launch_batch_totals(hipStream_t st, const BJob* jobs, u32 njobs, const u64* offs, u64* bits);

#define ZH_BATCH_PASS(P, W)                                                                                \
  int launch_batch_pass<P, W>(unsigned, unsigned, hipStream_t, const BJob*, const uint2*, u32, u32, u64*,       \
                              const u64*, u32)
extern template ZH_BATCH_PASS(BK_LEN, false);
extern template ZH_BATCH_PASS(BK_ENC, false);
extern template ZH_BATCH_PASS(BK_DEC, false);
extern template ZH_BATCH_PASS(BK_LEN, true);
extern template ZH_BATCH_PASS(BK_ENC, true);
extern template ZH_BATCH_PASS(BK_DEC, true);

// both families' launches of one pass over task lists split by family
// (tasks[0, nfast) fast, then nwide wide tasks); stride: DEC's slot stride
// for staged jobs (batch_stage_stride)
template <int PASS>
int launch_batch_families(unsigned grid, unsigned wg, hipStream_t st, const BJob* jobs, const uint2* tasks,
                          u32 nfast, u32 nwide, u32 per_task, u64* lens, const u64* offs, u32 stride)
{
  int rc = ZFP_HIP_OK;
  if (nfast)
    rc = launch_batch_pass<PASS, false>(grid, wg, st, jobs, tasks, nfast, per_task, lens, offs, stride);
  if (!rc && nwide)
    rc = launch_batch_pass<PASS, true>(grid, wg, st, jobs, tasks + nfast, nwide, per_task, lens, offs, 0);
  return rc;
}

// The LDS slot stride for the staged jobs among jobs[0, n) of one
// direction: odd, and past the longest block by one word (DEC: the reader's
// lookahead) or two (ENC: the slot coder's overflow); 0 when none is staged
inline u32 batch_stage_stride(const BJob* jobs, size_t n, int dir)
{
  u32 nw = 0;
  for (size_t j = 0; j < n; j++)
    if (jobs[j].stage && (int)jobs[j].dir == dir)
      nw = std::max(nw, jobs[j].z.maxbits >> 6);
  return nw ? (nw + (dir ? 1u : 2u)) | 1u : 0u;
}

} // namespace zfphip
