// Host build of the codec's per-block device code (zfp_amd/csrc/codec_block.h)
// for the CPU tests: the same encode_block/decode_block the gfx950 kernels run,
// executed one block after another, so that their streams can be compared
// with the oracle without a GPU.  Fixed-rate blocks go to b * maxbits;
// variable-rate blocks are laid out by a length pass and a prefix sum, as on
// the device.
#include <cstdint>
#include <cstddef>
#include <cstdlib>
#include <vector>

#include "codec_block.h"

using namespace zfphip;

namespace {

bool make_geom(int dims, const size_t* n, const ptrdiff_t* s, Geom& g, u32& nblocks)
{
  int64_t def[4] = {1, (int64_t)n[0], (int64_t)(n[0] * (n[1] ? n[1] : 1)),
                    (int64_t)(n[0] * (n[1] ? n[1] : 1) * (n[2] ? n[2] : 1))};
  uint64_t nb = 1;
  for (int i = 0; i < 4; i++) {
    if (i < dims) {
      g.n[i] = (u32)n[i];
      g.s[i] = (s && s[i]) ? s[i] : def[i];
      g.nb[i] = (u32)((n[i] + 3) / 4);
      nb *= g.nb[i];
    }
    else {
      g.n[i] = 1; g.s[i] = 0; g.nb[i] = 1;
    }
  }
  g.dims = dims;
  // rows run along the axis of smallest |stride|, as the library sets it
  for (int i = 0; i < 4; i++)
    g.lax[i] = (u32)i;
  for (int i = 1; i < dims; i++)
    for (int j = i; j > 0; j--) {
      int64_t a = g.s[g.lax[j]], b = g.s[g.lax[j - 1]];
      if ((a < 0 ? -a : a) < (b < 0 ? -b : b)) {
        u32 t = g.lax[j]; g.lax[j] = g.lax[j - 1]; g.lax[j - 1] = t;
      }
    }
  for (int i = 0; i < 4; i++)
    g.lnb[i] = g.nb[g.lax[i]];
  nblocks = (u32)nb;
  return true;
}

template <typename S, int D>
uint64_t enc(const void* data, const size_t* n, const ptrdiff_t* s, const Params& z, uint64_t* words,
             uint64_t offset)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  std::vector<uint64_t> off(nb + 1, 0);
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    BitCounter c;
    c.pos = 0;
    off[b + 1] = off[b] + (z.minexp < -1074 ? encode_block_rev<S, D>(c, (const S*)data, g, bc, z)
                                            : encode_block<S, D>(c, (const S*)data, g, bc, z));
  }
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    BitWriter<false> w;
    w.begin(words, offset + off[b], offset + off[b + 1]);
    if (z.minexp < -1074)
      encode_block_rev<S, D>(w, (const S*)data, g, bc, z);
    else
      encode_block<S, D>(w, (const S*)data, g, bc, z);
    w.finish();
  }
  return off[nb];
}

template <typename S, int D>
uint64_t dec(void* data, const size_t* n, const ptrdiff_t* s, const Params& z, const uint64_t* words,
             uint64_t nwords, uint64_t offset)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  uint64_t pos = offset;
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    BitReader r;
    r.begin(words, nwords, pos);
    pos += z.minexp < -1074 ? decode_block_rev<S, D, true>(r, (S*)data, g, bc, z)
                            : decode_block<S, D, true>(r, (S*)data, g, bc, z);
  }
  return pos - offset;
}

// the two-pass fixed-rate decode of 64-bit types: planes 63..32 of every
// block, then every block that goes on below plane 32 in full
template <typename S, int D>
uint64_t dec_split(void* data, const size_t* n, const ptrdiff_t* s, const Params& z, const uint64_t* words,
                   uint64_t nwords, uint64_t offset, uint64_t* nredo)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  std::vector<u32> redo;
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    BitReader r;
    r.begin(words, nwords, offset + (uint64_t)b * z.maxbits);
    bool more = false;
    decode_block<S, D, true, true>(r, (S*)data, g, bc, z, &more);
    if (more)
      redo.push_back(b);
  }
  for (u32 b : redo) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    BitReader r;
    r.begin(words, nwords, offset + (uint64_t)b * z.maxbits);
    decode_block<S, D, true>(r, (S*)data, g, bc, z);
  }
  *nredo = redo.size();
  return (uint64_t)nb * z.maxbits;
}

// fixed rate with word-aligned blocks through the slot writer and its
// straight-line plane coder: each block coded into its own zeroed slot of
// maxbits/64 + 2 words, whose first maxbits/64 words are the block's
template <typename S, int D>
uint64_t enc_slot(const void* data, const size_t* n, const ptrdiff_t* s, const Params& z, uint64_t* words,
                  uint64_t offset)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  const u32 nw = z.maxbits / 64;
  std::vector<uint64_t> slot(nw + 2);
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    std::fill(slot.begin(), slot.end(), 0ull);
    SlotWriter w;
    w.begin(slot.data(), nw + 2);
    encode_block<S, D>(w, (const S*)data, g, bc, z);
    for (u32 i = 0; i < nw; i++)
      words[offset / 64 + (uint64_t)b * nw + i] = slot[i];
  }
  return (uint64_t)nb * z.maxbits;
}

// fixed rate through the slot reader: each block's words copied into a slot
// (plus one word after them), read by bit index; blocks at b * maxbits from
// bit `offset`
template <typename S, int D>
uint64_t dec_slot(void* data, const size_t* n, const ptrdiff_t* s, const Params& z, const uint64_t* words,
                  uint64_t nwords, uint64_t offset)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  std::vector<uint64_t> slot(z.maxbits / 64 + 3);
  for (u32 b = 0; b < nb; b++) {
    u32 bc[4] = {0, 0, 0, 0};
    block_coords<D>(b, g, bc);
    const uint64_t start = offset + (uint64_t)b * z.maxbits;
    for (size_t i = 0; i < slot.size(); i++) {
      const uint64_t w = start / 64 + i;
      slot[i] = w < nwords ? words[w] : 0;
    }
    SlotReader r;
    r.begin(slot.data(), (u32)(start & 63));
    decode_block<S, D, true>(r, (S*)data, g, bc, z);
  }
  return (uint64_t)nb * z.maxbits;
}

template <typename S>
uint64_t enc_t(int dims, const void* data, const size_t* n, const ptrdiff_t* s, const Params& z,
               uint64_t* words, uint64_t offset)
{
  switch (dims) {
    case 1: return enc<S, 1>(data, n, s, z, words, offset);
    case 2: return enc<S, 2>(data, n, s, z, words, offset);
    case 3: return enc<S, 3>(data, n, s, z, words, offset);
    default: return enc<S, 4>(data, n, s, z, words, offset);
  }
}

template <typename S>
uint64_t dec_t(int dims, void* data, const size_t* n, const ptrdiff_t* s, const Params& z,
               const uint64_t* words, uint64_t nwords, uint64_t offset)
{
  switch (dims) {
    case 1: return dec<S, 1>(data, n, s, z, words, nwords, offset);
    case 2: return dec<S, 2>(data, n, s, z, words, nwords, offset);
    case 3: return dec<S, 3>(data, n, s, z, words, nwords, offset);
    default: return dec<S, 4>(data, n, s, z, words, nwords, offset);
  }
}

} // namespace

extern "C" uint64_t hostcodec_compress(int type, const size_t* n, const ptrdiff_t* s, const void* data,
                                       unsigned minbits, unsigned maxbits, unsigned maxprec, int minexp,
                                       uint64_t* words, uint64_t offset)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{minbits, maxbits, maxprec, minexp};
  switch (type) {
    case 1: return enc_t<i32>(dims, data, n, s, z, words, offset);
    case 2: return enc_t<i64>(dims, data, n, s, z, words, offset);
    case 3: return enc_t<float>(dims, data, n, s, z, words, offset);
    default: return enc_t<double>(dims, data, n, s, z, words, offset);
  }
}

extern "C" uint64_t hostcodec_decompress(int type, const size_t* n, const ptrdiff_t* s, void* data,
                                         unsigned minbits, unsigned maxbits, unsigned maxprec, int minexp,
                                         const uint64_t* words, uint64_t nwords, uint64_t offset)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{minbits, maxbits, maxprec, minexp};
  switch (type) {
    case 1: return dec_t<i32>(dims, data, n, s, z, words, nwords, offset);
    case 2: return dec_t<i64>(dims, data, n, s, z, words, nwords, offset);
    case 3: return dec_t<float>(dims, data, n, s, z, words, nwords, offset);
    default: return dec_t<double>(dims, data, n, s, z, words, nwords, offset);
  }
}

// Fixed-rate 1-3D fields with maxbits and offset multiples of 64 through the
// slot writer (the LDS-slot encoder); returns 0 for other requests.
extern "C" uint64_t hostcodec_compress_slot(int type, const size_t* n, const ptrdiff_t* s, const void* data,
                                            unsigned maxbits, unsigned maxprec, int minexp, uint64_t* words,
                                            uint64_t offset)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{maxbits, maxbits, maxprec, minexp};
  if (dims < 1 || dims > 3 || (maxbits & 63) || (offset & 63))
    return 0;
#define ZH_SLOT(S) \
  return dims == 1 ? enc_slot<S, 1>(data, n, s, z, words, offset) \
       : dims == 2 ? enc_slot<S, 2>(data, n, s, z, words, offset) : enc_slot<S, 3>(data, n, s, z, words, offset)
  switch (type) {
    case 1: ZH_SLOT(i32);
    case 2: ZH_SLOT(i64);
    case 3: ZH_SLOT(float);
    default: ZH_SLOT(double);
  }
#undef ZH_SLOT
}

// Fixed-rate 1-4D fields through the slot reader (the decoder of blocks
// staged in LDS); returns the bits consumed.
extern "C" uint64_t hostcodec_decompress_slot(int type, const size_t* n, const ptrdiff_t* s, void* data,
                                              unsigned maxbits, unsigned maxprec, int minexp, const uint64_t* words,
                                              uint64_t nwords, uint64_t offset)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{maxbits, maxbits, maxprec, minexp};
#define ZH_DSLOT(S) \
  return dims == 1 ? dec_slot<S, 1>(data, n, s, z, words, nwords, offset) \
       : dims == 2 ? dec_slot<S, 2>(data, n, s, z, words, nwords, offset) \
       : dims == 3 ? dec_slot<S, 3>(data, n, s, z, words, nwords, offset) \
                   : dec_slot<S, 4>(data, n, s, z, words, nwords, offset)
  switch (type) {
    case 1: ZH_DSLOT(i32);
    case 2: ZH_DSLOT(i64);
    case 3: ZH_DSLOT(float);
    default: ZH_DSLOT(double);
  }
#undef ZH_DSLOT
}

// Fixed-rate int64/double fields of 1-3 dimensions through the two-pass
// decode; *nredo receives the number of blocks decoded twice.
extern "C" uint64_t hostcodec_decompress_split(int type, const size_t* n, const ptrdiff_t* s, void* data,
                                               unsigned maxbits, unsigned maxprec, int minexp,
                                               const uint64_t* words, uint64_t nwords, uint64_t offset,
                                               uint64_t* nredo)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{maxbits, maxbits, maxprec, minexp};
  if (dims < 1 || dims > 3 || (type != 2 && type != 4))
    return 0;
  if (type == 2)
    return dims == 1 ? dec_split<i64, 1>(data, n, s, z, words, nwords, offset, nredo)
         : dims == 2 ? dec_split<i64, 2>(data, n, s, z, words, nwords, offset, nredo)
                     : dec_split<i64, 3>(data, n, s, z, words, nwords, offset, nredo);
  return dims == 1 ? dec_split<double, 1>(data, n, s, z, words, nwords, offset, nredo)
       : dims == 2 ? dec_split<double, 2>(data, n, s, z, words, nwords, offset, nredo)
                   : dec_split<double, 3>(data, n, s, z, words, nwords, offset, nredo);
}

// The chain rounds of the parallel index recovery (chain_lane, the body of
// k_chain), run lane after lane on the host with the launcher's schedule:
// phase 0, rounds of phase 1 until no exit changes, the scan of the counts,
// phase 2.  Segments are L bits (a multiple of 64).  Returns the number of
// phase-1 rounds, or -1 when the stream's bits hold fewer than nblocks
// blocks; boffs receives nblocks + 1 offsets relative to `offset`.
namespace {
template <typename S, int D, bool REV>
int chain_t(const size_t* n, const ptrdiff_t* s, const Params& z, const uint64_t* words, uint64_t nwords,
            uint64_t offset, uint64_t L, uint64_t* boffs)
{
  Geom g;
  u32 nb;
  make_geom(D, n, s, g, nb);
  const uint64_t span = nwords * 64 - offset;
  const u32 P = (u32)((span + L - 1) / L);
  std::vector<uint64_t> start(P), ex0(P), ex1(P), cnt(P), base(P);
  u32 changed = 0;
  ChainArgs a;
  a.g = g; a.z = z; a.in = words; a.nwords = nwords; a.offset = offset; a.span = span; a.L = L;
  a.start = start.data(); a.cnt = cnt.data(); a.base = base.data(); a.boffs = boffs; a.changed = &changed;
  a.P = P; a.nblocks = nb;
  uint64_t* ex[2] = {ex0.data(), ex1.data()};
  a.exit_prev = ex[1]; a.exit_new = ex[0];
  for (u32 i = 0; i < P; i++) chain_lane<S, D, REV>(a, i, 0);
  int cur = 0, rounds = 0;
  for (;; rounds++) {
    changed = 0;
    a.exit_prev = ex[cur]; a.exit_new = ex[cur ^ 1];
    for (u32 i = 0; i < P; i++) chain_lane<S, D, REV>(a, i, 1);
    cur ^= 1;
    if (!changed) break;
  }
  uint64_t total = 0;
  for (u32 i = 0; i < P; i++) { base[i] = total; total += cnt[i]; }
  if (total < nb) return -1;
  for (u32 i = 0; i < P; i++) chain_lane<S, D, REV>(a, i, 2);
  return rounds + 1;
}
} // namespace

extern "C" int hostcodec_chain_offsets(int type, const size_t* n, const ptrdiff_t* s, unsigned minbits,
                                       unsigned maxbits, unsigned maxprec, int minexp, const uint64_t* words,
                                       uint64_t nwords, uint64_t offset, uint64_t L, uint64_t* boffs)
{
  int dims = n[0] ? n[1] ? n[2] ? n[3] ? 4 : 3 : 2 : 1 : 0;
  Params z{minbits, maxbits, maxprec, minexp};
  const bool rev = minexp < -1074;
#define ZH_CHAIN(S) \
  return dims == 1 ? (rev ? chain_t<S, 1, true> : chain_t<S, 1, false>)(n, s, z, words, nwords, offset, L, boffs) \
       : dims == 2 ? (rev ? chain_t<S, 2, true> : chain_t<S, 2, false>)(n, s, z, words, nwords, offset, L, boffs) \
       : dims == 3 ? (rev ? chain_t<S, 3, true> : chain_t<S, 3, false>)(n, s, z, words, nwords, offset, L, boffs) \
                   : (rev ? chain_t<S, 4, true> : chain_t<S, 4, false>)(n, s, z, words, nwords, offset, L, boffs)
  switch (type) {
    case 1: ZH_CHAIN(i32);
    case 2: ZH_CHAIN(i64);
    case 3: ZH_CHAIN(float);
    default: ZH_CHAIN(double);
  }
#undef ZH_CHAIN
}
