// hipcodec.hpp -- block codecs for the reference's compressed arrays whose
// block coding runs on the GPU (libzfp_hip.so).
//
// The compressed arrays of LLNL zfp 1.0.1 take their codec as a template
// parameter (zfp::array3<Scalar, Codec, Index>, include/zfp/array3.hpp:21-24)
// and code one block per call through BlockStore<d>::encode/decode
// (include/zfp/internal/array/store3.hpp:94-120) into
// zfp::codec::zfp<d>::encode_block/decode_block and their strided forms
// (include/zfp/codec/zfpcodec.hpp:146-190, 370-420).  The codecs here are
// those classes with the block coding replaced:
//
//   zfp::array3<double, zfp::codec::hip3<double>> a(nx, ny, nz, rate);
//
// Outside a batch every call codes its block on the GPU at once.  Inside a
// zfp::codec::hip_batch scope, the block encodes and the strided block
// decodes (an array's set() and get(), and the write-back of its cache) are
// queued, and the scope codes them in one launch per stream when it ends:
//
//   {
//     zfp::codec::hip_batch batch;
//     a.set(p);             // every block queued
//   }                       // one launch encodes them all
//   {
//     zfp::codec::hip_batch batch;
//     a.get(q);             // destinations are written when the scope ends
//   }
//
// Inside a scope, memory that a queued strided decode writes, and the
// compressed data of a stream with queued encodes (compressed_data()), must
// not be read before the scope ends (or batch.flush()).  Contiguous block
// decodes (the fill of a cache line, on element access) always run at once,
// after the stream's queued work, so element access stays correct anywhere.
//
// Streams, rates and values are those of the default codecs
// (zfp::codec::zfp1..zfp4): the arrays' compressed data is the same, byte
// for byte.  The codecs take fixed-rate word-aligned blocks only (what the
// arrays' set_rate(rate, align = true) gives); other modes throw
// zfp::exception.  Each scope belongs to the thread that made it.
#ifndef ZFP_HIPCODEC_HPP
#define ZFP_HIPCODEC_HPP

#include <cstddef>
#include <cstring>
#include <exception>
#include <vector>

#include "zfp/codec/zfpcodec.hpp"
#include "zfp/internal/array/exception.hpp"
#include "zfp_hip.h"

namespace zfp {
namespace codec {

namespace hip_detail {

// Source positions of the four block positions along an axis of which only
// m (1..4) lie in the array: zfp's padding of partial blocks (the pattern
// of src/template/encode.c:8-27 applied per axis), as position indices
inline void pad_positions(unsigned m, unsigned* o)
{
  o[0] = 0;
  o[1] = m >= 2 ? 1 : 0;
  o[2] = m >= 3 ? 2 : (m == 2 ? 1 : 0);
  o[3] = m >= 4 ? 3 : 0;
}

// extents of a block from its shape code (BlockStore::block_shape: 2 bits
// per axis, 4 - extent)
inline void shape_extents(unsigned dims, unsigned shape, unsigned* n)
{
  for (unsigned d = 0; d < 4; d++) {
    n[d] = d < dims ? 4 - (shape & 3u) : 1;
    shape >>= 2;
  }
}

// a whole block (x fastest) from strided storage, padded like zfp pads a
// partial block
template <typename Scalar>
void gather(unsigned dims, unsigned shape, const Scalar* p, const std::ptrdiff_t* s, Scalar* block)
{
  unsigned n[4], o[4][4];
  shape_extents(dims, shape, n);
  const unsigned nw = dims > 3 ? 4 : 1, nz = dims > 2 ? 4 : 1, ny = dims > 1 ? 4 : 1;
  if (!shape && s[0] == 1) {
    // a whole block with contiguous rows
    for (unsigned w = 0; w < nw; w++)
      for (unsigned z = 0; z < nz; z++)
        for (unsigned y = 0; y < ny; y++, block += 4)
          std::memcpy(block, p + (std::ptrdiff_t)y * s[1] + (std::ptrdiff_t)z * s[2] + (std::ptrdiff_t)w * s[3],
                      4 * sizeof(Scalar));
    return;
  }
  for (unsigned d = 0; d < 4; d++)
    pad_positions(d < dims ? n[d] : 1, o[d]);
  for (unsigned w = 0; w < nw; w++)
    for (unsigned z = 0; z < nz; z++)
      for (unsigned y = 0; y < ny; y++)
        for (unsigned x = 0; x < 4; x++)
          *block++ = p[(std::ptrdiff_t)o[0][x] * s[0] + (std::ptrdiff_t)o[1][y] * s[1] +
                       (std::ptrdiff_t)o[2][z] * s[2] + (std::ptrdiff_t)o[3][w] * s[3]];
}

// the part of a whole block that lies in the array, to strided storage
template <typename Scalar>
void scatter(unsigned dims, unsigned shape, const Scalar* block, Scalar* p, const std::ptrdiff_t* s)
{
  unsigned n[4];
  shape_extents(dims, shape, n);
  const unsigned bw = dims > 3 ? 4 : 1, bz = dims > 2 ? 4 : 1, by = dims > 1 ? 4 : 1;
  if (!shape && s[0] == 1) {
    for (unsigned w = 0; w < bw; w++)
      for (unsigned z = 0; z < bz; z++)
        for (unsigned y = 0; y < by; y++, block += 4)
          std::memcpy(p + (std::ptrdiff_t)y * s[1] + (std::ptrdiff_t)z * s[2] + (std::ptrdiff_t)w * s[3], block,
                      4 * sizeof(Scalar));
    return;
  }
  for (unsigned w = 0; w < bw; w++)
    for (unsigned z = 0; z < bz; z++)
      for (unsigned y = 0; y < by; y++)
        for (unsigned x = 0; x < 4; x++, block++)
          if (x < n[0] && y < n[1] && z < n[2] && w < n[3])
            p[(std::ptrdiff_t)x * s[0] + (std::ptrdiff_t)y * s[1] + (std::ptrdiff_t)z * s[2] +
              (std::ptrdiff_t)w * s[3]] = *block;
}

inline void check(int rc)
{
  if (rc)
    throw zfp::exception("zfp::codec::hip: block coding on the GPU failed");
}

// the queued work of one stream: all encodes or all decodes
struct Group {
  uint64_t* words;
  uint64_t nwords;
  int type;
  unsigned dims;
  uint minbits, maxbits, maxprec;
  int minexp;
  bool decode;
  size_t esize;
  std::vector<uint64_t> slots;
  std::vector<unsigned char> blocks;   // whole blocks, one after another
  struct Dest {
    void* p;
    std::ptrdiff_t s[4];
    unsigned shape;
  };
  std::vector<Dest> dests;             // decodes: where each block goes

  void run()
  {
    const size_t count = slots.size();
    if (!count)
      return;
    if (decode) {
      check(zfp_hip_decode_blocks_host(type, dims, count, blocks.data(), slots.data(), minbits, maxbits, maxprec,
                                       minexp, words, nwords, 0));
      const size_t bsize = esize << (2 * dims);
      for (size_t k = 0; k < count; k++) {
        if (esize == 4)
          scatter<float>(dims, dests[k].shape, (const float*)(blocks.data() + k * bsize), (float*)dests[k].p,
                         dests[k].s);
        else
          scatter<double>(dims, dests[k].shape, (const double*)(blocks.data() + k * bsize), (double*)dests[k].p,
                          dests[k].s);
      }
    }
    else
      check(zfp_hip_encode_blocks_host(type, dims, count, blocks.data(), slots.data(), minbits, maxbits, maxprec,
                                       minexp, words, nwords, 0));
    slots.clear();
    blocks.clear();
    dests.clear();
  }
};

// the current thread's batch
struct Batch {
  std::vector<Group> groups;
  unsigned depth = 0;

  Group* find(const uint64_t* words)
  {
    for (Group& g : groups)
      if (g.words == words)
        return &g;
    return nullptr;
  }
  void flush()
  {
    // groups of different streams are independent; a group's own order is
    // kept by running its queue whenever its kind changes
    for (Group& g : groups)
      g.run();
    groups.clear();
  }
};

inline Batch& batch()
{
  static thread_local Batch b;
  return b;
}

} // namespace hip_detail

// Scope in which block encodes and strided block decodes are queued and
// coded together (see the top of this file).  Scopes nest: the outermost
// one codes the queue.
class hip_batch {
public:
  hip_batch() { hip_detail::batch().depth++; }
  // codes the queue when it is the outermost scope; a failure throws
  // zfp::exception unless an exception is already propagating
  ~hip_batch() noexcept(false)
  {
    hip_detail::Batch& b = hip_detail::batch();
    if (--b.depth == 0) {
      try {
        b.flush();
      }
      catch (...) {
        b.groups.clear();
        if (!std::uncaught_exceptions())
          throw;
      }
    }
  }
  // code everything queued so far
  void flush() { hip_detail::batch().flush(); }

  hip_batch(const hip_batch&) = delete;
  hip_batch& operator=(const hip_batch&) = delete;
};

// The codec zfp<dims> (Base, one of zfp1..zfp4 of Scalar) with its block
// coding on the GPU.
template <class Base, unsigned dims, typename Scalar>
class hip_codec : public Base {
public:
  size_t encode_block(bitstream_offset offset, unsigned shape, const Scalar* block) const
  {
    const std::ptrdiff_t s[4] = {1, 4, 16, 64};
    return put(offset, shape, block, s);
  }

  size_t decode_block(bitstream_offset offset, unsigned shape, Scalar* block) const
  {
    // a cache line is used as soon as this returns: decoded at once
    const std::ptrdiff_t s[4] = {1, 4, 16, 64};
    return get(offset, shape, block, s, true);
  }

protected:
  using Base::stream;

  // the stream's parameters; fixed rate with word-aligned blocks only
  void target(hip_detail::Group& g) const
  {
    bitstream* bs = zfp_stream_bit_stream(stream);
    if (!bs)
      throw zfp::exception("zfp::codec::hip: no stream is open");
    g.words = (uint64_t*)stream_data(bs);
    g.nwords = stream_capacity(bs) / sizeof(uint64_t);
    zfp_stream_params(stream, &g.minbits, &g.maxbits, &g.maxprec, &g.minexp);
    if (zfp_stream_compression_mode(stream) != zfp_mode_fixed_rate || g.minbits != g.maxbits || !g.maxbits ||
        g.maxbits % 64)
      throw zfp::exception("zfp::codec::hip: fixed-rate word-aligned blocks only");
    g.type = (int)Base::type;
    g.dims = dims;
    g.esize = sizeof(Scalar);
  }

  // the group that takes this call's block: the batch's group of this
  // stream (flushed first when it holds the other kind of work), or a
  // one-block group run at once
  size_t code(bitstream_offset offset, unsigned shape, const Scalar* src, Scalar* dst, const std::ptrdiff_t* s,
              bool decode, bool now) const
  {
    hip_detail::Group one;
    target(one);
    one.decode = decode;
    hip_detail::Batch& b = hip_detail::batch();
    hip_detail::Group* g = &one;
    if (b.depth) {
      hip_detail::Group* q = b.find(one.words);
      if (q && (q->decode != decode || q->maxbits != one.maxbits || q->maxprec != one.maxprec ||
                q->minexp != one.minexp || q->type != one.type || q->nwords != one.nwords || now))
        q->run();
      if (!now) {
        if (!q) {
          b.groups.push_back(one);
          q = &b.groups.back();
        }
        q->decode = decode;
        q->maxbits = one.maxbits;
        q->minbits = one.minbits;
        q->maxprec = one.maxprec;
        q->minexp = one.minexp;
        q->type = one.type;
        q->nwords = one.nwords;
        g = q;
      }
    }
    // block position: a whole slot when the offset is a multiple of the
    // block size, else a one-block group at that (word-aligned) offset
    if (offset % 64)
      throw zfp::exception("zfp::codec::hip: block offsets must be word-aligned");
    if (offset % g->maxbits) {
      if (g != &one)
        g->run();
      hip_detail::Group at = one;
      at.words = one.words + offset / 64;
      at.nwords = one.nwords - offset / 64;
      add(at, 0, shape, src, dst, s);
      at.run();
      return one.maxbits;
    }
    add(*g, offset / g->maxbits, shape, src, dst, s);
    if (g == &one)
      one.run();
    return one.maxbits;
  }

  void add(hip_detail::Group& g, uint64_t slot, unsigned shape, const Scalar* src, Scalar* dst,
           const std::ptrdiff_t* s) const
  {
    const size_t bsize = sizeof(Scalar) << (2 * dims);
    const size_t at = g.blocks.size();
    g.blocks.resize(at + bsize);
    g.slots.push_back(slot);
    if (g.decode) {
      typename hip_detail::Group::Dest d;
      d.p = dst;
      for (int i = 0; i < 4; i++)
        d.s[i] = s[i];
      d.shape = shape;
      g.dests.push_back(d);
    }
    else
      hip_detail::gather<Scalar>(dims, shape, src, s, (Scalar*)(g.blocks.data() + at));
  }

  size_t put(bitstream_offset offset, unsigned shape, const Scalar* p, const std::ptrdiff_t* s) const
  {
    return code(offset, shape, p, nullptr, s, false, false);
  }

  size_t get(bitstream_offset offset, unsigned shape, Scalar* p, const std::ptrdiff_t* s, bool now) const
  {
    return code(offset, shape, nullptr, p, s, true, now);
  }
};

template <typename Scalar>
class hip1 : public hip_codec<zfp1<Scalar>, 1, Scalar> {
public:
  size_t encode_block_strided(bitstream_offset offset, unsigned shape, const Scalar* p, std::ptrdiff_t sx) const
  {
    const std::ptrdiff_t s[4] = {sx, 0, 0, 0};
    return this->put(offset, shape, p, s);
  }
  size_t decode_block_strided(bitstream_offset offset, unsigned shape, Scalar* p, std::ptrdiff_t sx) const
  {
    const std::ptrdiff_t s[4] = {sx, 0, 0, 0};
    return this->get(offset, shape, p, s, false);
  }
};

template <typename Scalar>
class hip2 : public hip_codec<zfp2<Scalar>, 2, Scalar> {
public:
  size_t encode_block_strided(bitstream_offset offset, unsigned shape, const Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, 0, 0};
    return this->put(offset, shape, p, s);
  }
  size_t decode_block_strided(bitstream_offset offset, unsigned shape, Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, 0, 0};
    return this->get(offset, shape, p, s, false);
  }
};

template <typename Scalar>
class hip3 : public hip_codec<zfp3<Scalar>, 3, Scalar> {
public:
  size_t encode_block_strided(bitstream_offset offset, unsigned shape, const Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy, std::ptrdiff_t sz) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, sz, 0};
    return this->put(offset, shape, p, s);
  }
  size_t decode_block_strided(bitstream_offset offset, unsigned shape, Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy, std::ptrdiff_t sz) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, sz, 0};
    return this->get(offset, shape, p, s, false);
  }
};

template <typename Scalar>
class hip4 : public hip_codec<zfp4<Scalar>, 4, Scalar> {
public:
  size_t encode_block_strided(bitstream_offset offset, unsigned shape, const Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy, std::ptrdiff_t sz, std::ptrdiff_t sw) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, sz, sw};
    return this->put(offset, shape, p, s);
  }
  size_t decode_block_strided(bitstream_offset offset, unsigned shape, Scalar* p, std::ptrdiff_t sx,
                              std::ptrdiff_t sy, std::ptrdiff_t sz, std::ptrdiff_t sw) const
  {
    const std::ptrdiff_t s[4] = {sx, sy, sz, sw};
    return this->get(offset, shape, p, s, false);
  }
};

} // namespace codec
} // namespace zfp

#endif
