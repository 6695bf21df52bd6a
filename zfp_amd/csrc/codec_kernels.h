// codec_kernels.h -- kernel templates of the MI355X zfp block codec and the
// host-side launchers that instantiate them.  Included by one translation unit
// per dimensionality (inst_d1.hip .. inst_d4.hip) so that the 1D-4D kernel
// families compile in parallel.
//
// Kernels (one lane = one zfp block; lanes walk the blocks along the axis of
// smallest stride, each block's bits go to its place in stream order):
//   k_encode<S,D,MODE>  gather -> BFP cast -> lifting -> reorder -> planes ->
//                       embedded coder, writing the block's bits at its offset
//                       (fixed rate: off + b*maxbits; otherwise from a scan)
//                       or, in the length pass, only counting them
//   k_decode<S,D>       the inverse, reading each block from its offset;
//                       for 3D doubles at fixed rates up to 16 bits/value a
//                       first pass over planes 63..32 (HI_ONLY) lists the
//                       blocks that k_decode_list then decodes in full
//   k_encode_waves<S,D> variable-length write pass whose waves assemble
//                       their words in LDS (lanes in stream order)
//   k_chain<S,D>        block starts of a variable-rate stream that arrives
//                       without an index, by chains of block lengths
//                       followed in parallel segments (rounds until exact)
//   k_parse<S,D>        one-lane sequential parse, the fallback of k_chain
//   k_prepare           zeroes the words that blocks will OR into
//   k_scan_*            exclusive prefix sum of block lengths
#pragma once

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "backend.h"
#include "codec_block.h"
#include "codec_device.h"

namespace zfphip {

// ENC_FIXED_PLAIN: fixed rate with every block range word-aligned, so each
// lane owns whole words and stores them without atomics; ENC_FIXED_AT: the
// same with block b placed at bit offset + boffs[b] (chosen slots)
enum { ENC_FIXED = 0, ENC_OFFSETS = 1, ENC_LENGTH = 2, ENC_FIXED_PLAIN = 3, ENC_FIXED_AT = 4 };

// the block codec of the stream's mode: lossy, or reversible (REV)
template <typename S, int D, bool REV, typename W>
__device__ __forceinline__ u32 encode_any(W& w, const S* data, const Geom& g, const u32* bc, const Params& z)
{
  if constexpr (REV)
    return encode_block_rev<S, D>(w, data, g, bc, z);
  else
    return encode_block<S, D>(w, data, g, bc, z);
}

template <typename S, int D, bool STORE, bool REV, typename R>
__device__ __forceinline__ u32 decode_any(R& r, S* data, const Geom& g, const u32* bc, const Params& z)
{
  if constexpr (REV)
    return decode_block_rev<S, D, STORE>(r, data, g, bc, z);
  else
    return decode_block<S, D, STORE>(r, data, g, bc, z);
}

// OCC: waves per SIMD the register budget allows (2: up to 256 VGPRs, 3: up to
// 168).  Three waves hide more latency when coding stays in the high half of
// the planes; with the low half (high rates) the tighter budget spills.
template <typename S, int D, int MODE, bool REV = false, int OCC = 2>
__global__ void __launch_bounds__(256, OCC)
k_encode(const S* __restrict__ data, Geom g, u32 nblocks, Params z,
         u64* __restrict__ out, u64 offset, const u64* __restrict__ boffs, u64* __restrict__ lens)
{
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblocks)
    return;
  u32 bc[4] = {0, 0, 0, 0};
  const u32 b = lane_block<D>(t, g, bc);
  if (MODE == ENC_LENGTH) {
    BitCounter cnt;
    cnt.pos = 0;
    lens[b] = encode_any<S, D, REV>(cnt, data, g, bc, z);
    if (t == 0)
      lens[nblocks] = 0; // so that the exclusive scan's last entry is the total
    return;
  }
  u64 start, end;
  if (MODE == ENC_FIXED || MODE == ENC_FIXED_PLAIN) {
    start = offset + (u64)b * z.maxbits;
    end = start + z.maxbits;
  }
  else if (MODE == ENC_FIXED_AT) {
    start = offset + boffs[b];
    end = start + z.maxbits;
  }
  else {
    start = offset + boffs[b];
    end = offset + boffs[b + 1];
  }
  BitWriter<MODE == ENC_FIXED_PLAIN || MODE == ENC_FIXED_AT> w;
  w.begin(out, start, end);
  encode_any<S, D, REV>(w, data, g, bc, z);
  w.finish();
}

// LDS operations of the lanes of one wave, ordered before what follows
ZH_INLINE void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Word runs between the lanes' LDS slots (nw words each, slot stride
// `stride`) and the stream, when the wave's blocks lie anywhere in the
// stream (lanes not in stream order): lane i's block starts at stream word
// start_i.  Each pass moves 64 words -- 64 / nw whole blocks when nw divides
// 64 -- so every load or store instruction covers whole runs of consecutive
// words instead of one word of each of 64 distant blocks.  The starts go
// through LDS (`starts`, 64 words per wave).
//
// Both this and the single-run copies below go 8 passes at a time: the
// eight loads (or LDS reads) are all in flight before the first result is
// used, so a wave waits out one memory latency per 8 passes, not per pass.
// `first` (one run) or `starts` (one run per block, in LDS) give the stream
// word of each (block, word) pair.
constexpr u32 XFER_BATCH = 8;

template <bool TO_STREAM, bool ONE_RUN>
ZH_INLINE void wave_copy(u64* __restrict__ stream, u64 nwords, u64* wslots, u32 stride, u32 nw, u32 cnt,
                         u64 first, const u64* starts)
{
  const u32 lane = threadIdx.x & 63u;
  const u32 total = cnt * nw;
  u32 blk = lane / nw, wd = lane - blk * nw;
  const u32 db = 64u / nw, dw = 64u - db * nw;
  for (u32 q0 = 0; q0 < total; q0 += 64 * XFER_BATCH) {
    u64 v[XFER_BATCH];
    u64 gi[XFER_BATCH];
    u32 li[XFER_BATCH];
#pragma unroll
    for (u32 k = 0; k < XFER_BATCH; k++) {
      // pass k: word wd of block blk
      li[k] = blk * stride + wd;
      gi[k] = ONE_RUN ? first + q0 + 64 * k + lane : starts[blk < cnt ? blk : 0] + wd;
      wd += dw;
      blk += db;
      if (wd >= nw) {
        wd -= nw;
        blk++;
      }
    }
#pragma unroll
    for (u32 k = 0; k < XFER_BATCH; k++) {
      const bool in_run = q0 + 64 * k + lane < total;
      if constexpr (TO_STREAM)
        v[k] = in_run ? wslots[li[k]] : 0ull;
      else
        v[k] = stream[gi[k] < nwords ? gi[k] : nwords - 1]; // clamped: always a valid address
    }
#pragma unroll
    for (u32 k = 0; k < XFER_BATCH; k++) {
      const bool in_run = q0 + 64 * k + lane < total;
      if (in_run) {
        if constexpr (TO_STREAM)
          stream[gi[k]] = v[k];
        else
          wslots[li[k]] = gi[k] < nwords ? v[k] : 0ull;
      }
    }
  }
}

template <bool TO_STREAM>
ZH_INLINE void wave_transfer(u64* __restrict__ stream, u64 nwords, u64* wslots, u32 stride, u32 nw, u32 cnt,
                             u64 my_start, u64* starts)
{
  const u32 lane = threadIdx.x & 63u;
  starts[lane] = my_start;
  wave_lds_sync();
  wave_copy<TO_STREAM, false>(stream, nwords, wslots, stride, nw, cnt, 0, starts);
  wave_lds_sync();
}

// Fixed rate with word-aligned blocks (maxbits = 64 * nw, offset a multiple
// of 64), 1-3D: each lane codes its block into its own slot of LDS (nw words
// plus two for the last plane's overflow, slot stride `stride`, odd), with
// the straight-line plane coder; then the block's nw words go to the stream.
// When the lanes hold consecutive stream blocks (lane order = stream order)
// the wave's blocks are one run of words, copied with coalesced stores;
// otherwise the wave stores whole blocks per pass (wave_transfer).
template <typename S, int D, int OCC>
__global__ void __launch_bounds__(256, OCC)
k_encode_slot(const S* __restrict__ data, Geom g, u32 nblocks, Params z, u64* __restrict__ out, u64 word0,
              u32 stride, u32 stream_order)
{
  extern __shared__ u64 slots[];
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63u;
  const u32 t0 = t - lane;
  if (t0 >= nblocks)
    return; // the whole wave lies past the field
  const u32 nw = z.maxbits >> 6;
  u64* slot = slots + (u64)threadIdx.x * stride;
  for (u32 i = 0; i < nw + 2; i++)
    slot[i] = 0;
  u32 b = 0;
  if (t < nblocks) {
    u32 bc[4] = {0, 0, 0, 0};
    b = lane_block<D>(t, g, bc);
    SlotWriter w;
    w.begin(slot, nw + 2);
    encode_block<S, D>(w, data, g, bc, z);
  }
  if (stream_order) {
    // blocks t0 .. t0 + cnt - 1 are stream blocks t0 ..: words
    // word0 + t0 * nw .. of the stream; word q of the run is word q % nw of
    // the slot of lane q / nw
    wave_lds_sync();
    wave_copy<true, true>(out, 0, slots + (u64)(threadIdx.x - lane) * stride, stride, nw, umin(64u, nblocks - t0),
                          word0 + (u64)t0 * nw, nullptr);
  }
  else {
    // the wave's blocks lie apart in the stream: whole runs per pass
    u64* starts = slots + (u64)blockDim.x * stride + (threadIdx.x - lane);
    wave_transfer<true>(out, 0, slots + (u64)(threadIdx.x - lane) * stride, stride, nw, umin(64u, nblocks - t0),
                        word0 + (u64)b * nw, starts);
  }
}

// Bit writer of the variable-length write pass whose wave assembles its
// words in LDS: a wave's lanes hold consecutive blocks of the stream, so
// their bits cover one contiguous range of words; words shared by two lanes
// are OR-ed in LDS (cheap) instead of in global memory.  `staged` is the
// same in every lane of the wave; when false (the wave's range is larger
// than its LDS area) the writer stores to the stream directly, OR-ing
// shared words with global atomics.
struct WaveWriter {
  u64* out;   // the stream
  u64* st;    // the wave's LDS words; st[0] is stream word wbase
  u64 wbase;
  u64 wi;     // index of the word being filled
  u64 acc;    // pending bits of word wi
  u32 fill;   // number of pending bits (< 64)
  u64 lo, hi; // the block's bit range
  bool staged;

  ZH_INLINE void emit(u64 word)
  {
    const u64 b0 = wi << 6;
    if (b0 < hi) {
      if (b0 + 64 > hi)
        word &= lowmask((u32)(hi - b0));
      const bool whole = b0 >= lo && b0 + 64 <= hi;
      // two branches, so that each keeps its address space (LDS or global)
      if (staged) {
        if (whole)
          st[wi - wbase] = word;
        else if (word)
          atomicOr((unsigned long long*)(st + (wi - wbase)), (unsigned long long)word);
      }
      else {
        if (whole)
          out[wi] = word;
        else if (word)
          atomicOr((unsigned long long*)(out + wi), (unsigned long long)word);
      }
    }
    wi++;
  }
  ZH_INLINE void put(u64 v, u32 n)
  {
    acc |= v << fill;
    u32 nf = fill + n;
    if (nf >= 64) {
      emit(acc);
      acc = (v >> 1) >> (63 - fill);
      nf -= 64;
    }
    fill = nf;
  }
  ZH_INLINE void zeros(u64 n)
  {
    while (n >= 64) { put(0, 64); n -= 64; }
    if (n)
      put(0, (u32)n);
  }
  ZH_INLINE void finish()
  {
    if (fill)
      emit(acc);
  }
};

// Write pass of variable-length coding when lane t holds stream block t (the
// lane order is the stream's): each wave zeroes its LDS area (wcap words),
// its lanes write their blocks there, then the wave copies the words to the
// stream with whole-word stores; only the range's first and last words,
// which neighbouring waves share, are OR-ed in global memory (into words
// k_prepare zeroed).  boffs: block offsets relative to `offset`.
template <typename S, int D, bool REV>
__global__ void __launch_bounds__(256)
k_encode_waves(const S* __restrict__ data, Geom g, u32 nblocks, Params z, u64* __restrict__ out, u64 offset,
               const u64* __restrict__ boffs, u32 wcap)
{
  extern __shared__ u64 wstage[];
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63u;
  const u32 t0 = t - lane;
  if (t0 >= nblocks)
    return; // the whole wave lies past the field
  const u32 tl = umin(t0 + 63u, nblocks - 1u);
  const u64 first = offset + boffs[t0], end = offset + boffs[tl + 1];
  const u64 wlo = first >> 6, whi = (end + 63) >> 6;
  const bool staged = whi - wlo <= wcap;
  const u32 nw = staged ? (u32)(whi - wlo) : 0u;
  u64* st = wstage + (u64)(threadIdx.x >> 6) * wcap;
  for (u32 i = lane; i < nw; i += 64)
    st[i] = 0;
  wave_lds_sync();
  if (t < nblocks) {
    u32 bc[4] = {0, 0, 0, 0};
    const u32 b = lane_block<D>(t, g, bc); // == t
    WaveWriter w;
    w.out = out; w.st = st; w.wbase = wlo; w.staged = staged;
    w.lo = offset + boffs[b]; w.hi = offset + boffs[b + 1];
    w.wi = w.lo >> 6; w.fill = (u32)(w.lo & 63); w.acc = 0;
    encode_any<S, D, REV>(w, data, g, bc, z);
    w.finish();
  }
  wave_lds_sync();
  for (u32 i = lane; i < nw; i += 64) {
    const u64 v = st[i];
    const bool edge = (i == 0 && (first & 63)) || (i + 1 == nw && (end & 63));
    if (!edge)
      out[wlo + i] = v;
    else if (v)
      atomicOr((unsigned long long*)(out + wlo + i), (unsigned long long)v);
  }
}

// Variable-length coding of long blocks (more than 32 words, beyond the LDS
// slots of k_encode_stage) with lanes in stream order, every block coded
// once:
//
//   k_encode_blocks  codes block b into its own bw words of `stage` (from
//                    bit 0, plain stores) and writes its length;
//   exclusive scan   of the lengths: the block offsets;
//   k_place_blocks   k_encode_waves with the staged bits in place of a
//                    second coding: each wave assembles its blocks' bits in
//                    LDS and stores whole words, OR-ing only its two edge
//                    words (zeroed by k_prepare).
template <typename S, int D, bool REV>
__global__ void __launch_bounds__(256)
k_encode_blocks(const S* __restrict__ data, Geom g, u32 nblocks, Params z, u64* __restrict__ stage,
                u64* __restrict__ lens, u32 bw)
{
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nblocks)
    return;
  u32 bc[4] = {0, 0, 0, 0};
  const u32 b = lane_block<D>(t, g, bc); // == t: stream order
  const u64 lo = (u64)b * bw * 64;
  BitWriter<true> w;
  w.begin(stage, lo, lo + (u64)bw * 64);
  lens[b] = encode_any<S, D, REV>(w, data, g, bc, z);
  w.finish();
  if (t == 0)
    lens[nblocks] = 0; // so that the exclusive scan's last entry is the total
}

template <int UNUSED = 0>
__global__ void __launch_bounds__(256)
k_place_blocks(const u64* __restrict__ stage, u32 bw, u32 nblocks, u64* __restrict__ out, u64 offset,
               const u64* __restrict__ boffs, u32 wcap)
{
  extern __shared__ u64 wstage[];
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  const u32 lane = threadIdx.x & 63u;
  const u32 t0 = t - lane;
  if (t0 >= nblocks)
    return; // the whole wave lies past the field
  const u32 tl = umin(t0 + 63u, nblocks - 1u);
  const u64 first = offset + boffs[t0], end = offset + boffs[tl + 1];
  const u64 wlo = first >> 6, whi = (end + 63) >> 6;
  const bool staged = whi - wlo <= wcap;
  const u32 nw = staged ? (u32)(whi - wlo) : 0u;
  u64* st = wstage + (u64)(threadIdx.x >> 6) * wcap;
  for (u32 i = lane; i < nw; i += 64)
    st[i] = 0;
  wave_lds_sync();
  if (t < nblocks) {
    WaveWriter w;
    w.out = out; w.st = st; w.wbase = wlo; w.staged = staged;
    w.lo = offset + boffs[t]; w.hi = offset + boffs[t + 1];
    w.wi = w.lo >> 6; w.fill = (u32)(w.lo & 63); w.acc = 0;
    const u64 len = boffs[t + 1] - boffs[t];
    const u64* src = stage + (u64)t * bw;
    // eight loads in flight at a time (words of the block's own stride,
    // always readable), then their bits in order
    for (u64 i0 = 0; 64 * i0 < len; i0 += 8) {
      u64 v[8];
#pragma unroll
      for (u32 j = 0; j < 8; j++)
        v[j] = i0 + j < bw ? src[i0 + j] : 0ull;
#pragma unroll
      for (u32 j = 0; j < 8; j++) {
        const u64 at = 64 * (i0 + j);
        if (at < len) {
          const u32 n = (u32)(len - at < 64 ? len - at : 64);
          w.put(v[j] & lowmask(n), n);
        }
      }
    }
    w.finish();
  }
  wave_lds_sync();
  for (u32 i = lane; i < nw; i += 64) {
    const u64 v = st[i];
    const bool edge = (i == 0 && (first & 63)) || (i + 1 == nw && (end & 63));
    if (!edge)
      out[wlo + i] = v;
    else if (v)
      atomicOr((unsigned long long*)(out + wlo + i), (unsigned long long)v);
  }
}

// Bits one block can take in the stream's mode: the per-block term of
// zfp_stream_maximum_size (src/zfp.c:711-742).
template <typename S, int D, bool REV>
__host__ __device__ inline u32 block_bits_bound(const Params& z)
{
  typedef Traits<S> T;
  constexpr u32 N = 1u << (2 * D);
  const u32 hdr = T::FLOAT ? (REV ? 2 + T::EBITS + T::PBITS : 1 + T::EBITS) : (REV ? T::PBITS : 0);
  const u32 prec = z.maxprec < (u32)T::P ? z.maxprec : (u32)T::P;
  u64 b = hdr + (N - 1) + (u64)N * prec;
  b = b < z.maxbits ? b : z.maxbits;
  b = b > z.minbits ? b : z.minbits;
  return (u32)b;
}

// Variable-length coding that codes every block once, for fields whose
// lanes hold consecutive stream blocks of at most bw words each:
//
//   k_encode_stage  each wave codes its 64 blocks into LDS slots,
//                   concatenates their bits in LDS and stores them, from bit
//                   0, in its own stride of 64 * bw words of `stage`; it
//                   writes its bit count to totals[wave] and each block's
//                   offset within the wave to boffs;
//   exclusive scan  of the totals (one per wave): each wave's first bit;
//   k_place         each staged wave's bits are shifted into place (the
//                   first word keeps the bits before them), and the block
//                   offsets get their wave's first bit.
//
// No wave waits for another and no word is shared between writers.
template <typename S, int D, bool REV>
__global__ void __launch_bounds__(256)
k_encode_stage(const S* __restrict__ data, Geom g, u32 nblocks, Params z, u64* __restrict__ stage,
               u64* __restrict__ totals, u64* __restrict__ boffs, u32 bw)
{
  extern __shared__ u64 lds[];
  const u32 lane = threadIdx.x & 63u;
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x, wn = t >> 6;
  if ((t & ~63u) >= nblocks)
    return; // the whole wave lies past the field
  const u32 sstride = (bw + 2) | 1u, awords = 64 * bw + 2;
  u64* slot_base = lds + (u64)(threadIdx.x >> 6) * (64 * sstride + awords);
  u64* area = slot_base + 64 * sstride;
  const bool act = t < nblocks;
  u64* slot = slot_base + lane * sstride;
  for (u32 i = 0; i < sstride; i++)
    slot[i] = 0;
  u64 len = 0;
  if (act) {
    u32 bc[4] = {0, 0, 0, 0};
    lane_block<D>(t, g, bc); // == t: stream order
    SlotWriter w;
    w.begin(slot, bw + 2);
    len = encode_any<S, D, REV>(w, data, g, bc, z);
  }
  // bit offsets of the lanes' blocks within the wave
  u64 incl = len;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u64 v = __shfl_up(incl, d, 64);
    if (lane >= d)
      incl += v;
  }
  const u64 excl = incl - len;
  const u64 T = __shfl(incl, 63, 64);
  const u32 nwa = (u32)((T + 63) >> 6);
  for (u32 i = lane; i <= nwa; i += 64)
    area[i] = 0;
  wave_lds_sync();
  if (act && len) {
    // area words fb..lb hold the block; those it fills whole are stored,
    // the two it shares with its neighbours are OR-ed
    const u32 sh = (u32)(excl & 63), fb = (u32)(excl >> 6), lb = (u32)((excl + len - 1) >> 6);
    const u32 lw = (u32)((len + 63) >> 6);
    u64 prev = 0;
    for (u32 k = fb, i = 0; k <= lb; k++, i++) {
      u64 cur = i < lw ? slot[i] : 0ull;
      if (i + 1 == lw && (len & 63))
        cur &= lowmask((u32)(len & 63));
      const u64 v = sh ? (cur << sh) | (prev >> (64 - sh)) : cur;
      prev = cur;
      const bool whole = (u64)k * 64 >= excl && (u64)k * 64 + 64 <= excl + len;
      if (whole)
        area[k] = v;
      else if (v)
        atomicOr((unsigned long long*)(area + k), (unsigned long long)v);
    }
  }
  wave_lds_sync();
  u64* dst = stage + (u64)wn * 64 * bw;
  for (u32 i = lane; i < nwa; i += 64)
    dst[i] = area[i];
  if (lane == 0)
    totals[wn] = T;
  if (act)
    boffs[t] = excl;
}

// Bits [pos, pos + n) (n <= 64) of wave w's staged bits, low bit first
ZH_INLINE u64 staged_bits(const u64* __restrict__ stage, u64 stride, u64 w, u64 pos, u32 n)
{
  const u64* q = stage + w * stride + (pos >> 6);
  const u32 s = (u32)(pos & 63);
  u64 v = q[0] >> s;
  if (s && s + n > 64)
    v |= q[1] << (64 - s);
  return v & lowmask(n);
}

// One wave per staged wave w: its bits shifted to their place in `out`.
// The first word takes the bits before them (the previous wave's last
// bits, or the caller's below `offset`); the last word, when the next wave
// shares it, is left to that wave.  Every staged wave but the last holds at
// least 64 bits (each block takes at least one), so the previous wave's
// last bits are in its own stride.  (A template, so that every translation
// unit that includes this header may instantiate it.)
template <int UNUSED = 0>
__global__ void __launch_bounds__(256)
k_place(const u64* __restrict__ stage, u64 stride, const u64* __restrict__ first, u32 nwaves, u32 nblocks,
        u64* __restrict__ out, u64 offset, u64* __restrict__ boffs)
{
  const u32 lane = threadIdx.x & 63u;
  const u32 w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  if (w >= nwaves)
    return;
  const u64 T = first[w + 1] - first[w];
  const u64 G = offset + first[w];
  const u32 s = (u32)(G & 63);
  const u64 a = G >> 6, nout = (s + T + 63) >> 6;
  const bool last = w + 1 == nwaves;
  const u64* src = stage + (u64)w * stride;
  for (u64 j = lane; j < nout; j += 64) {
    const u64 lo = j < (T + 63) >> 6 ? src[j] : 0ull;
    const u64 prev = j ? src[j - 1] : 0ull;
    u64 v = s ? (lo << s) | (prev >> (64 - s)) : lo;
    if (j == 0 && s) {
      if (w) {
        const u64 Tp = first[w] - first[w - 1];
        v |= staged_bits(stage, stride, w - 1, Tp - s, s);
      }
      else
        v |= out[a] & lowmask(s);
    }
    const bool shared = j + 1 == nout && ((s + T) & 63) && !last;
    if (!shared)
      out[a + j] = v;
  }
  // block offsets: the wave's first bit added
  const u32 t = w * 64 + lane;
  if (t < nblocks)
    boffs[t] += first[w];
  if (last && lane == 0)
    boffs[nblocks] = first[nwaves];
}

template <typename S, int D, bool REV, bool HI_ONLY, typename R>
__device__ __forceinline__ void decode_lane(R& r, S* __restrict__ data, const Geom& g, const u32* bc,
                                            const Params& z, u32 t, u32 b, const u64* __restrict__ boffs,
                                            u32* __restrict__ redo, u32* __restrict__ nredo, u32* __restrict__ bad);

// STAGE (fixed rate, blocks of at most `stride` words): each lane first
// copies its block's words into its own slot of LDS, with all loads in
// flight at once, and the bit reader then refills from LDS.  Reading the
// stream word by word from global memory inside the coder loop would wait
// out the full memory latency at every refill.
//
// HI_ONLY (fixed rate, 64-bit types): decode bit planes 63..32 only, which
// fits three waves per SIMD instead of two, and append the lane numbers of
// blocks that go on below plane 32 to redo[] (*nredo counts them, one atomic
// per wave); k_decode_list then decodes those in full, over the partial
// values written here.
//
// bad (variable length with a block index): set to 1 when a block consumes
// other than boffs[b+1] - boffs[b] bits.  Block 0 starts at the stream's
// start, so when no block disagrees every offset is, by induction, the true
// start of its block; one disagreement means the index does not belong to
// this stream and the caller decodes again from a parse.
//
// coal (STAGE, word-aligned blocks of nw = maxbits / 64 words, lanes in
// stream order): the wave's blocks are one run of stream words, which the
// wave copies into its lanes' slots with coalesced loads.
template <typename S, int D, bool STAGE, bool REV = false, bool HI_ONLY = false>
__global__ void __launch_bounds__(256, HI_ONLY ? 3 : 2)
k_decode(S* __restrict__ data, Geom g, u32 nblocks, Params z, const u64* __restrict__ in,
         u64 nwords, u64 offset, const u64* __restrict__ boffs, u32 stride,
         u32* __restrict__ redo = nullptr, u32* __restrict__ nredo = nullptr, u32* __restrict__ bad = nullptr,
         u32 coal = 0)
{
  extern __shared__ u64 stage[];
  const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
  if constexpr (STAGE) {
    if (coal) {
      const u32 lane = threadIdx.x & 63u;
      const u32 t0 = t - lane;
      if (t0 >= nblocks)
        return; // the whole wave lies past the field
      const u32 nw = z.maxbits >> 6;
      const u32 cnt = umin(64u, nblocks - t0);
      u64* wslots = stage + (u64)(threadIdx.x - lane) * stride;
      if (coal == 1) {
        wave_copy<false, true>((u64*)in, nwords, wslots, stride, nw, cnt, (offset >> 6) + (u64)t0 * nw, nullptr);
        wave_lds_sync();
      }
      else {
        // the wave's blocks lie apart in the stream: whole runs per pass
        u64 first = 0;
        if (t < nblocks) {
          u32 c[4] = {0, 0, 0, 0};
          const u32 bb = lane_block<D>(t, g, c);
          first = (offset + (boffs ? boffs[bb] : (u64)bb * z.maxbits)) >> 6;
        }
        wave_transfer<false>((u64*)in, nwords, wslots, stride, nw, cnt, first,
                             stage + (u64)blockDim.x * stride + (threadIdx.x - lane));
      }
    }
  }
  if (t >= nblocks)
    return;
  u32 bc[4] = {0, 0, 0, 0};
  const u32 b = lane_block<D>(t, g, bc);
  const u64 start = offset + (boffs ? boffs[b] : (u64)b * z.maxbits);
  if constexpr (STAGE) {
    u64* slot = stage + (u64)threadIdx.x * stride;
    u32 bit0 = 0;
    if (!coal) {
      const u64 first = start >> 6;
      const u32 cnt = (u32)(((start + z.maxbits + 63) >> 6) - first); // <= stride
      // batches of 16 loads with clamped (always valid) addresses, then the
      // stores of the words that exist; the host guarantees nwords >= 1
      for (u32 j0 = 0; j0 < cnt; j0 += 16) {
        u64 v[16];
#pragma unroll
        for (u32 q = 0; q < 16; q++) {
          const u64 i = first + j0 + q;
          v[q] = in[i < nwords ? i : nwords - 1];
        }
#pragma unroll
        for (u32 q = 0; q < 16; q++)
          if (j0 + q < cnt)
            slot[j0 + q] = first + j0 + q < nwords ? v[q] : 0ull;
      }
      bit0 = (u32)(start & 63);
    }
    // the slot's stride leaves a readable word after the block's last
    SlotReader r;
    r.begin(slot, bit0);
    decode_lane<S, D, REV, HI_ONLY>(r, data, g, bc, z, t, b, boffs, redo, nredo, bad);
  }
  else {
    BitReader r;
    r.begin(in, nwords, start);
    decode_lane<S, D, REV, HI_ONLY>(r, data, g, bc, z, t, b, boffs, redo, nredo, bad);
  }
}

// the decode of one lane's block, and the bookkeeping of k_decode's modes
template <typename S, int D, bool REV, bool HI_ONLY, typename R>
__device__ __forceinline__ void decode_lane(R& r, S* __restrict__ data, const Geom& g, const u32* bc,
                                            const Params& z, u32 t, u32 b, const u64* __restrict__ boffs,
                                            u32* __restrict__ redo, u32* __restrict__ nredo, u32* __restrict__ bad)
{
  if constexpr (HI_ONLY) {
    static_assert(!REV && Traits<S>::P == 64 && D < 4, "HI_ONLY: fixed-rate 64-bit types, 1-3D");
    bool more = false;
    decode_block<S, D, true, true>(r, data, g, bc, z, &more);
    const u64 m = __ballot(more);
    if (m) {
      const u32 below = __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
      const u32 lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
      const u32 leader = (u32)__builtin_ctzll(m);
      u32 base = 0;
      if (lane == leader)
        base = atomicAdd(nredo, (u32)__builtin_popcountll(m));
      base = __builtin_amdgcn_readlane(base, leader);
      if (more)
        redo[base + below] = t;
    }
  }
  else {
    const u32 c = decode_any<S, D, true, REV>(r, data, g, bc, z);
    if (bad && (u64)c != boffs[b + 1] - boffs[b])
      *bad = 1u;
  }
}

// Full fixed-rate decode of the lanes listed in redo[0 .. *nredo), a
// grid-stride loop over the list.
template <typename S, int D>
__global__ void __launch_bounds__(256)
k_decode_list(S* __restrict__ data, Geom g, Params z, const u64* __restrict__ in, u64 nwords, u64 offset,
              const u32* __restrict__ redo, const u32* __restrict__ nredo)
{
  const u32 n = *nredo;
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    u32 bc[4] = {0, 0, 0, 0};
    const u32 b = lane_block<D>(redo[i], g, bc);
    BitReader r;
    r.begin(in, nwords, offset + (u64)b * z.maxbits);
    decode_block<S, D, true>(r, data, g, bc, z);
  }
}

// One lane walks the stream and records every block's offset (relative to
// `offset`); boffs receives nblocks + 1 entries.
template <typename S, int D, bool REV>
__global__ void k_parse(Geom g, u32 nblocks, Params z, const u64* __restrict__ in, u64 nwords,
                        u64 offset, u64* __restrict__ boffs)
{
  if (blockIdx.x || threadIdx.x)
    return;
  u64 pos = 0;
  u32 bc[4] = {0, 0, 0, 0};
  for (u32 b = 0; b < nblocks; b++) {
    boffs[b] = pos;
    BitReader r;
    r.begin(in, nwords, offset + pos);
    pos += decode_any<S, D, false, REV>(r, (S*)nullptr, g, bc, z);
  }
  boffs[nblocks] = pos;
}

// one lane per segment: the phases of chain_lane (codec_block.h)
template <typename S, int D, bool REV>
__global__ void __launch_bounds__(64) k_chain(ChainArgs a, u32 phase)
{
  const u32 i = blockIdx.x * 64 + threadIdx.x;
  if (i < a.P)
    chain_lane<S, D, REV>(a, i, phase);
}

static inline unsigned grid_for(u64 n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// reversible mode (zfp_stream_set_reversible, src/zfp.c:750-757) is marked by
// minexp below ZFP_MIN_EXP; its blocks have variable length
static inline bool is_reversible(const Params& z) { return z.minexp < -1074; }

// Block offsets (nblocks + 1, relative to `offset`) of a variable-length
// stream without an index, by the chain rounds of k_chain over the stream's
// readable bits.  Returns CHAIN_FALLBACK (and writes nothing) when those
// bits hold fewer than nblocks blocks or the rounds do not settle within
// ROUNDS_MAX; the caller then parses sequentially.  Waits for `st`.
constexpr int CHAIN_FALLBACK = 101;

template <typename S, int D>
int chain_parse(const Desc& d, const Params& z, const u64* in, u64 nwords, u64 offset, u64* boffs, hipStream_t st)
{
  constexpr u32 ROUNDS_MAX = 256;
  const u64 span = nwords * 64 > offset ? nwords * 64 - offset : 0;
  if (!span || z.maxbits == 0 || d.nblocks == 0)
    return CHAIN_FALLBACK;
  // segments of at least 32 Kibit (chains from arbitrary bits mostly meet
  // the true one within that), at most 65536 of them
  u64 L = std::max<u64>(32768, (span + 65535) / 65536);
  L = (L + 63) & ~63ull;
  const u32 P = (u32)((span + L - 1) / L);
  void* tp;
  int rc = g_scan_tmp.get(((u64)5 * P + 1) * sizeof(u64), &tp);
  if (rc)
    return rc;
  u64* start = (u64*)tp;
  u64* ex[2] = {start + P, start + 2 * (u64)P};
  u64* cnt = start + 3 * (u64)P;
  u64* base = start + 4 * (u64)P;
  u32* changed = (u32*)(start + 5 * (u64)P);
  ChainArgs a;
  a.g = d.g;
  a.z = z;
  a.in = in;
  a.nwords = nwords;
  a.offset = offset;
  a.span = span;
  a.L = L;
  a.start = start;
  a.cnt = cnt;
  a.base = base;
  a.boffs = boffs;
  a.changed = changed;
  a.P = P;
  a.nblocks = d.nblocks;
  const bool rev = is_reversible(z);
  auto launch = [&](u32 phase) -> int {
    if (rev)
      hipLaunchKernelGGL((k_chain<S, D, true>), dim3(grid_for(P, 64)), dim3(64), 0, st, a, phase);
    else
      hipLaunchKernelGGL((k_chain<S, D, false>), dim3(grid_for(P, 64)), dim3(64), 0, st, a, phase);
    HIPCHK(hipGetLastError());
    return 0;
  };
  a.exit_prev = ex[1];
  a.exit_new = ex[0];
  if ((rc = launch(0)))
    return rc;
  int cur = 0; // ex[cur] holds the latest exits
  u32 rounds = 0;
  for (;; rounds++) {
    if (rounds == ROUNDS_MAX)
      return CHAIN_FALLBACK;
    HIPCHK(hipMemsetAsync(changed, 0, sizeof(u32), st));
    a.exit_prev = ex[cur];
    a.exit_new = ex[cur ^ 1];
    if ((rc = launch(1)))
      return rc;
    cur ^= 1;
    u32 h = 0;
    HIPCHK(hipMemcpyAsync(&h, changed, sizeof(u32), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!h)
      break;
  }
  // block numbers of the segments' first starts
  std::vector<u64> c(P);
  HIPCHK(hipMemcpyAsync(c.data(), cnt, (u64)P * sizeof(u64), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  u64 total = 0;
  for (u32 i = 0; i < P; i++) {
    const u64 v = c[i];
    c[i] = total;
    total += v;
  }
  if (total < d.nblocks)
    return CHAIN_FALLBACK;
  HIPCHK(hipMemcpyAsync(base, c.data(), (u64)P * sizeof(u64), hipMemcpyHostToDevice, st));
  a.exit_prev = ex[cur];
  a.exit_new = ex[cur ^ 1];
  if ((rc = launch(2)))
    return rc;
  // the host vector must outlive the copy
  HIPCHK(hipStreamSynchronize(st));
  return 0;
}

template <typename S, int D, bool REV>
int launch_encode_var(const Desc& d, const void* data, const Params& z, u64* out, u64 offset, u64* bits_out,
                      hipStream_t st, u64** boffs_out, const Workspace* ws);

// at (fixed rate, word-aligned blocks only): block b is written at bit
// offset + at[b] instead of offset + b * maxbits, and no other word is touched
template <typename S, int D>
int launch_encode(const Desc& d, const void* data, const Params& z, u64* out, u64 offset,
                         u64* bits_out, hipStream_t st, u64** boffs_out, const u64* at, const Workspace* ws)
{
  const unsigned BS = 256;
  const unsigned grid = grid_for(d.nblocks, BS);
  if (is_reversible(z))
    return launch_encode_var<S, D, true>(d, data, z, out, offset, bits_out, st, boffs_out, ws);
  const bool fixed = z.minbits == z.maxbits;
  if (!fixed)
    return launch_encode_var<S, D, false>(d, data, z, out, offset, bits_out, st, boffs_out, ws);
  if (at) {
    if ((offset & 63) || (z.maxbits & 63))
      return record_error(ZFP_HIP_EINVAL, "blocks at chosen places need word-aligned blocks");
    hipLaunchKernelGGL((k_encode<S, D, ENC_FIXED_AT>), dim3(grid), dim3(BS), 0, st, (const S*)data, d.g,
                       d.nblocks, z, out, offset, at, (u64*)nullptr);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    *bits_out = (u64)d.nblocks * z.maxbits;
    if (boffs_out)
      *boffs_out = nullptr;
    return ZFP_HIP_OK;
  }
  {
    u64 total = (u64)d.nblocks * z.maxbits;
    u64 w0 = offset >> 6, w1 = (offset + total + 63) >> 6;
    if ((offset & 63) || (z.maxbits & 63)) {
      hipLaunchKernelGGL(k_prepare, dim3(grid_for(w1 - w0, 256)), dim3(256), 0, st, out, w0, w1,
                         (u32)(offset & 63));
      hipLaunchKernelGGL((k_encode<S, D, ENC_FIXED>), dim3(grid), dim3(BS), 0, st, (const S*)data, d.g,
                         d.nblocks, z, out, offset, (const u64*)nullptr, (u64*)nullptr);
    }
    else if (D < 4 && z.maxbits <= 124u * 64u) {
      // 1-3D: blocks coded in LDS slots, then stored whole; a workgroup's
      // slots (and one block start per lane) take at most 64 KiB, which
      // holds one wave's slots up to 124 words a block.  Up to 16 bits per
      // value with lanes along x, the three-wave variant (higher rates need
      // the registers)
      if constexpr (D < 4) {
        const u32 nw = z.maxbits >> 6;
        const u32 stride = (nw + 2) | 1u;
        unsigned wg = 256;
        while (wg > 64 && (u64)wg * (stride + 1) * 8 > 65536)
          wg /= 2;
        bool order = true; // lane t holds stream block t
        for (int i = 0; i < d.dims; i++)
          order = order && d.g.lax[i] == (u32)i;
        const size_t lds = (size_t)wg * (stride + 1) * 8; // slots, then one block start per lane
        if (z.maxbits <= (16u << (2 * D)) && d.g.lax[0] == 0)
          hipLaunchKernelGGL((k_encode_slot<S, D, 3>), dim3(grid_for(d.nblocks, wg)), dim3(wg), lds, st,
                             (const S*)data, d.g, d.nblocks, z, out, offset >> 6, stride, (u32)order);
        else
          hipLaunchKernelGGL((k_encode_slot<S, D, 2>), dim3(grid_for(d.nblocks, wg)), dim3(wg), lds, st,
                             (const S*)data, d.g, d.nblocks, z, out, offset >> 6, stride, (u32)order);
      }
    }
    else {
      hipLaunchKernelGGL((k_encode<S, D, ENC_FIXED_PLAIN>), dim3(grid), dim3(BS), 0, st, (const S*)data,
                         d.g, d.nblocks, z, out, offset, (const u64*)nullptr, (u64*)nullptr);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    *bits_out = total;
    if (boffs_out)
      *boffs_out = nullptr;
    return ZFP_HIP_OK;
  }
}

// Coding every block once (k_encode_stage, a scan of the waves' bit counts,
// k_place) for a field whose lanes hold consecutive stream blocks, each of at
// most bw words; block offsets go to offs (nblocks + 1).  Waits for `st`.
template <typename S, int D, bool REV>
int encode_once(const Desc& d, const void* data, const Params& z, u64* out, u64 offset, u64* offs, u32 bw,
                u64* total, hipStream_t st)
{
  const u32 nwaves = (u32)((d.nblocks + 63) / 64);
  const u64 stride = (u64)64 * bw;
  void* tp;
  // the waves' bit counts and first bits, then the staged bits
  int rc = g_stage.get(((u64)2 * (nwaves + 1) + (u64)nwaves * stride) * sizeof(u64), &tp);
  if (rc)
    return rc;
  u64* totals = (u64*)tp;
  u64* first = totals + nwaves + 1;
  u64* stage = first + nwaves + 1;
  const u32 sstride = (bw + 2) | 1u;
  const size_t per_wave = ((size_t)64 * sstride + 64 * bw + 2) * 8;
  unsigned wg = 256;
  while (wg > 64 && (wg / 64) * per_wave > 65536)
    wg /= 2;
  HIPCHK(hipMemsetAsync(totals + nwaves, 0, sizeof(u64), st));
  hipLaunchKernelGGL((k_encode_stage<S, D, REV>), dim3(grid_for(d.nblocks, wg)), dim3(wg), (wg / 64) * per_wave,
                     st, (const S*)data, d.g, d.nblocks, z, stage, totals, offs, bw);
  HIPCHK(hipGetLastError());
  if ((rc = exclusive_scan(totals, first, (u64)nwaves + 1, st)))
    return rc;
  hipLaunchKernelGGL(k_place<0>, dim3(grid_for(nwaves, 4)), dim3(256), 0, st, (const u64*)stage, stride,
                     (const u64*)first, nwaves, d.nblocks, out, offset, offs);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(total, first + nwaves, sizeof(u64), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return ZFP_HIP_OK;
}

// variable rate (and reversible): length pass, exclusive scan, prepare,
// write pass -- or, for small blocks with lanes in stream order, every block
// coded once (encode_once)
template <typename S, int D, bool REV>
int launch_encode_var(const Desc& d, const void* data, const Params& z, u64* out, u64 offset, u64* bits_out,
                      hipStream_t st, u64** boffs_out, const Workspace* ws)
{
  const unsigned BS = 256;
  const unsigned grid = grid_for(d.nblocks, BS);
  u64 nb1 = (u64)d.nblocks + 1;
  u64* lens;
  u64* offs;
  int rc;
  if (ws) {
    // the caller's workspace: nothing is allocated, no index is kept
    lens = ws->lens;
    offs = ws->offs;
    boffs_out = nullptr;
  }
  else {
    void* lp;
    rc = g_lens.get(2 * nb1 * sizeof(u64), &lp);
    if (rc)
      return rc;
    lens = (u64*)lp;
    offs = lens + nb1;
  }
  // the caller's copy of the offsets: the scan writes it directly
  struct Held {
    u64* p = nullptr;
    u64 bytes = 0;
    ~Held()
    {
      if (p)
        index_release(p, bytes);
    }
  } held;
  if (boffs_out) {
    *boffs_out = nullptr;
    held.bytes = nb1 * sizeof(u64);
    if (index_buffer(held.bytes, &held.p) == ZFP_HIP_OK)
      offs = held.p;
  }
  bool stream_order = true; // lane t holds stream block t
  for (int i = 0; i < d.dims; i++)
    stream_order = stream_order && d.g.lax[i] == (u32)i;
  // every block coded once when a wave's slots fit LDS (blocks of at most
  // 32 words); not for the caller's workspace (no allocation, a fixed set of
  // passes)
  const u32 bw = (block_bits_bound<S, D, REV>(z) + 63) / 64;
  if (stream_order && !ws && bw >= 1 && bw <= 32 && d.nblocks > 0) {
    u64 total = 0;
    rc = encode_once<S, D, REV>(d, data, z, out, offset, offs, bw, &total, st);
    if (rc == ZFP_HIP_OK) {
      *bits_out = total;
      if (boffs_out) {
        *boffs_out = held.p;
        held.p = nullptr;
      }
      return ZFP_HIP_OK;
    }
    // no room for the staged bits (nblocks * bw words): the multi-pass
    // coder needs only the lengths and offsets
    if (rc != ZFP_HIP_ENOMEM)
      return rc;
  }
  // longer blocks in stream order: each coded once into a staging stride
  // of bw words when those fit in device memory (k_encode_blocks), else
  // a length pass now and a second coding in the write pass
  u64* bstage = nullptr;
  if (stream_order && !ws && bw > 32 && d.nblocks > 0) {
    void* tp;
    if (g_stage.get((u64)d.nblocks * bw * sizeof(u64), &tp) == ZFP_HIP_OK)
      bstage = (u64*)tp;
  }
  if (bstage)
    hipLaunchKernelGGL((k_encode_blocks<S, D, REV>), dim3(grid), dim3(BS), 0, st, (const S*)data, d.g, d.nblocks,
                       z, bstage, lens, bw);
  else
    hipLaunchKernelGGL((k_encode<S, D, ENC_LENGTH, REV>), dim3(grid), dim3(BS), 0, st, (const S*)data, d.g,
                       d.nblocks, z, out, offset, (const u64*)nullptr, lens);
  HIPCHK(hipGetLastError());
  rc = ws ? exclusive_scan_tmp(lens, offs, nb1, ws->tmp, st) : exclusive_scan(lens, offs, nb1, st);
  if (rc)
    return rc;
  // a wave's LDS area in the write pass: twice the words 64 average blocks
  // take (64 blocks of b bits take b words), at least 64 and at most 8192
  // (64 KiB); k_encode_waves, four waves per workgroup, takes at most 2048
  // of them, k_place_blocks, one wave per workgroup, all.  A wave whose
  // range is longer writes directly, so the size sets the speed, never the
  // stream.  It is taken from the last encode with the same codec
  // parameters; only the first such encode waits for the stream length on
  // the host.
  auto wcap_for = [&](u64 bits) {
    const u64 avg_bits = bits / (d.nblocks ? d.nblocks : 1) + 1;
    return (u32)std::min<u64>(8192, std::max<u64>(64, 2 * avg_bits + 4));
  };
  // (The workspace variant always waits: it keeps no host-side state.)
  static std::map<std::tuple<u32, u32, u32, int>, u32> last_wcap;
  const auto key = std::make_tuple(z.minbits, z.maxbits, z.maxprec, z.minexp);
  u32 wcap = 0;
  if (stream_order && !ws) {
    auto it = last_wcap.find(key);
    if (it != last_wcap.end())
      wcap = it->second;
  }
  const bool wait = stream_order && !wcap;
  u64 total = 0;
  if (wait) {
    HIPCHK(hipMemcpyAsync(&total, offs + d.nblocks, sizeof(u64), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    wcap = wcap_for(total);
    u64 w0 = offset >> 6, w1 = (offset + total + 63) >> 6;
    if (w1 > w0)
      hipLaunchKernelGGL(k_prepare, dim3(grid_for(w1 - w0, 256)), dim3(256), 0, st, out, w0, w1,
                         (u32)(offset & 63));
  }
  else {
    hipLaunchKernelGGL(k_prepare_dev, dim3(std::min(grid, 2048u)), dim3(256), 0, st, out, offset,
                       (const u64*)(offs + d.nblocks));
  }
  if (bstage) {
    hipLaunchKernelGGL(k_place_blocks<0>, dim3(grid_for(d.nblocks, 64)), dim3(64), (size_t)wcap * 8, st,
                       (const u64*)bstage, bw, d.nblocks, out, offset, (const u64*)offs, wcap);
  }
  else if (stream_order) {
    const u32 wc = std::min<u32>(wcap, 2048);
    hipLaunchKernelGGL((k_encode_waves<S, D, REV>), dim3(grid), dim3(BS), (size_t)4 * wc * 8, st, (const S*)data,
                       d.g, d.nblocks, z, out, offset, (const u64*)offs, wc);
  }
  else {
    hipLaunchKernelGGL((k_encode<S, D, ENC_OFFSETS, REV>), dim3(grid), dim3(BS), 0, st, (const S*)data, d.g,
                       d.nblocks, z, out, offset, (const u64*)offs, (u64*)nullptr);
  }
  HIPCHK(hipGetLastError());
  if (!wait)
    HIPCHK(hipMemcpyAsync(&total, offs + d.nblocks, sizeof(u64), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (stream_order && !ws) {
    if (last_wcap.size() >= 64)
      last_wcap.clear();
    last_wcap[key] = wcap_for(total);
  }
  *bits_out = total;
  if (boffs_out) {
    *boffs_out = held.p;
    held.p = nullptr;
  }
  return ZFP_HIP_OK;
}

template <typename S, int D>
int launch_decode(const Desc& d, void* data, const Params& z, const u64* in, u64 nwords,
                         u64 offset, const u64* boffs, u64* bits_out, hipStream_t st)
{
  const unsigned BS = 256;
  const unsigned grid = grid_for(d.nblocks, BS);
  const bool rev = is_reversible(z);
  const bool fixed = !rev && z.minbits == z.maxbits;
  u64 total = 0;
  // a variable-length stream decoded with the caller's index checks every
  // block against it (k_decode's `bad`)
  u32* bad = nullptr;
  if (!fixed && boffs) {
    void* fp;
    int rc = g_flag.get(sizeof(u64), &fp);
    if (rc)
      return rc;
    bad = (u32*)fp;
    HIPCHK(hipMemsetAsync(bad, 0, sizeof(u32), st));
  }
  if (!fixed && !boffs) {
    // no index: recover block offsets by a sequential parse
    void* lp;
    int rc = g_lens.get(((u64)d.nblocks + 1) * sizeof(u64), &lp);
    if (rc)
      return rc;
    rc = chain_parse<S, D>(d, z, in, nwords, offset, (u64*)lp, st);
    if (rc == CHAIN_FALLBACK) {
      if (rev)
        hipLaunchKernelGGL((k_parse<S, D, true>), dim3(1), dim3(64), 0, st, d.g, d.nblocks, z, in, nwords,
                           offset, (u64*)lp);
      else
        hipLaunchKernelGGL((k_parse<S, D, false>), dim3(1), dim3(64), 0, st, d.g, d.nblocks, z, in, nwords,
                           offset, (u64*)lp);
      HIPCHK(hipGetLastError());
    }
    else if (rc)
      return rc;
    boffs = (const u64*)lp;
  }
  // fixed rate: stage blocks through LDS when a workgroup's slots fit in
  // 64 KiB (an odd slot stride spreads the lanes over the LDS banks)
  const u32 cmax = (u32)((z.maxbits + 63) / 64) + ((offset & 63) || (z.maxbits & 63) ? 1u : 0u);
  const u32 stride = cmax | 1u;
  // slots, then one block start per lane (the staging of scattered blocks)
  unsigned sbs = 256;
  while (sbs > 64 && (u64)sbs * (stride + 1) * 8 > 65536)
    sbs /= 2;
  const bool staged = fixed && nwords > 0 && (u64)sbs * (stride + 1) * 8 <= 65536;
  // word-aligned blocks: waves stage their blocks' words with whole-run
  // loads, one run when the lanes are in stream order (1), else a run per
  // block (2)
  bool order = true;
  for (int i = 0; i < d.dims; i++)
    order = order && d.g.lax[i] == (u32)i;
  const bool aligned = !(offset & 63) && !(z.maxbits & 63);
  const u32 coal = !staged || !aligned ? 0u : order && !boffs ? 1u : 2u;
  if constexpr (std::is_same<S, double>::value && D == 3) {
    // 3D doubles at up to 16 bits per value (the 1D and 2D decoders already
    // run at five or more waves per SIMD): the cast puts each block's largest
    // value at plane 62, so a block seldom reaches plane 31 (near-constant
    // blocks do).  Decode the high planes at three waves per SIMD, then the
    // few blocks that go lower.  (int64 values keep their own scale: small
    // ones live in the low planes, so int64 keeps the one-pass decode.)
    if (staged && !boffs && z.maxbits <= (16u << (2 * D))) {
      void* lp;
      int rc = g_lens.get(((u64)d.nblocks + 1) * sizeof(u32), &lp);
      if (rc)
        return rc;
      u32* nredo = (u32*)lp;
      u32* redo = nredo + 1;
      HIPCHK(hipMemsetAsync(nredo, 0, sizeof(u32), st));
      hipLaunchKernelGGL((k_decode<S, D, true, false, true>), dim3(grid_for(d.nblocks, sbs)), dim3(sbs),
                         (size_t)sbs * (stride + 1) * 8, st, (S*)data, d.g, d.nblocks, z, in, nwords, offset,
                         (const u64*)nullptr, stride, redo, nredo, (u32*)nullptr, coal);
      // 512 workgroups fill the chip at two waves per SIMD
      hipLaunchKernelGGL((k_decode_list<S, D>), dim3(std::min(grid, 512u)), dim3(BS), 0, st, (S*)data, d.g, z,
                         in, nwords, offset, (const u32*)redo, (const u32*)nredo);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st));
      *bits_out = (u64)d.nblocks * z.maxbits;
      return ZFP_HIP_OK;
    }
  }
  if (staged) {
    // fixed rate: boffs, when given, places block b at offset + boffs[b]
    hipLaunchKernelGGL((k_decode<S, D, true>), dim3(grid_for(d.nblocks, sbs)), dim3(sbs),
                       (size_t)sbs * (stride + 1) * 8, st, (S*)data, d.g, d.nblocks, z, in, nwords, offset,
                       boffs, stride, (u32*)nullptr, (u32*)nullptr, (u32*)nullptr, coal);
  }
  else if (rev) {
    hipLaunchKernelGGL((k_decode<S, D, false, true>), dim3(grid), dim3(BS), 0, st, (S*)data, d.g, d.nblocks, z,
                       in, nwords, offset, boffs, 0u, (u32*)nullptr, (u32*)nullptr, bad);
  }
  else {
    hipLaunchKernelGGL((k_decode<S, D, false>), dim3(grid), dim3(BS), 0, st, (S*)data, d.g, d.nblocks, z, in,
                       nwords, offset, boffs, 0u, (u32*)nullptr, (u32*)nullptr, bad);
  }
  HIPCHK(hipGetLastError());
  u32 mismatch = 0;
  if (fixed)
    total = (u64)d.nblocks * z.maxbits;
  else
    HIPCHK(hipMemcpyAsync(&total, boffs + d.nblocks, sizeof(u64), hipMemcpyDeviceToHost, st));
  if (bad)
    HIPCHK(hipMemcpyAsync(&mismatch, bad, sizeof(u32), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  if (mismatch)
    return ZFP_HIP_ESTALE;
  *bits_out = total;
  return ZFP_HIP_OK;
}

template <int D>
int dispatch_encode_d(const Desc& d, const void* data, const Params& z, u64* out, u64 offset,
                             u64* bits, hipStream_t st, u64** boffs, const u64* at, const Workspace* ws)
{
  switch (d.type) {
    case zfp_type_int32: return launch_encode<i32, D>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case zfp_type_int64: return launch_encode<i64, D>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case zfp_type_float: return launch_encode<float, D>(d, data, z, out, offset, bits, st, boffs, at, ws);
    case zfp_type_double: return launch_encode<double, D>(d, data, z, out, offset, bits, st, boffs, at, ws);
  }
  return ZFP_HIP_EINVAL;
}


template <int D>
int dispatch_decode_d(const Desc& d, void* data, const Params& z, const u64* in, u64 nwords,
                             u64 offset, const u64* boffs, u64* bits, hipStream_t st)
{
  switch (d.type) {
    case zfp_type_int32: return launch_decode<i32, D>(d, data, z, in, nwords, offset, boffs, bits, st);
    case zfp_type_int64: return launch_decode<i64, D>(d, data, z, in, nwords, offset, boffs, bits, st);
    case zfp_type_float: return launch_decode<float, D>(d, data, z, in, nwords, offset, boffs, bits, st);
    case zfp_type_double: return launch_decode<double, D>(d, data, z, in, nwords, offset, boffs, bits, st);
  }
  return ZFP_HIP_EINVAL;
}


} // namespace zfphip
