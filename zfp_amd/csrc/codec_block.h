// codec_block.h -- one zfp block, encoded or decoded by one lane: the
// composition of the primitives in codec_device.h into the reference's block
// codec (LLNL zfp 1.0.1, cited per function).  Shared by the gfx950 kernels
// (codec_kernels.h) and by a host build of the same source used in CPU tests.
#pragma once

#include "codec_device.h"

namespace zfphip {

// codec parameters of a zfp_stream (include/zfp.h:86-93)
struct Params {
  u32 minbits, maxbits, maxprec;
  i32 minexp;
};

// number of bit planes coded for a float block (src/template/codecf.c:5-13)
__device__ __forceinline__ u32 plane_count(int emax, u32 maxprec, int minexp, int dims)
{
  int p = emax - minexp + 2 * dims + 2;
  if (p < 0)
    p = 0;
  return (u32)p < maxprec ? (u32)p : maxprec;
}

// block coordinates of linear block id b (x fastest, src/template/compress.c)
template <int D>
__device__ __forceinline__ void block_coords(u32 b, const Geom& g, u32* c)
{
  c[0] = b % g.nb[0];
  u32 t = b / g.nb[0];
  if (D > 1) { c[1] = t % g.nb[1]; t /= g.nb[1]; }
  if (D > 2) { c[2] = t % g.nb[2]; t /= g.nb[2]; }
  if (D > 3) { c[3] = t; }
}

// ---------------------------------------------------------------------------
// encode one block (src/template/encodef.c:62-90, encodei.c:6-15,
// encode.c:259-280); returns the block's length in bits
// ---------------------------------------------------------------------------
template <typename S, int D, typename W>
__device__ __forceinline__ u32 encode_block(W& w, const S* data, const Geom& g, const u32* bc, const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  constexpr int N = 1 << (2 * D);
  S v[N];
  Int q[N];
  gather<S, D>(v, data, g, bc);
  u32 hdr = 0;
  u32 minbits = z.minbits, maxbits = z.maxbits, maxprec = z.maxprec;
  if constexpr (T::FLOAT) {
    int emax = block_emax<N>(v);
    u32 mp = plane_count(emax, z.maxprec, z.minexp, D);
    u32 e = mp ? (u32)(emax + T::EBIAS) : 0u;
    if (!e) {
      // all-zero (or below tolerance) block: a single zero bit
      w.put(0, 1);
      if (z.minbits > 1) {
        w.zeros(z.minbits - 1);
        return z.minbits;
      }
      return 1;
    }
    hdr = 1 + T::EBITS;
    w.put(2 * (u64)e + 1, hdr);
#pragma unroll
    for (int i = 0; i < N; i++)
      q[i] = fwd_cast_one(v[i], emax);
    maxprec = mp;
    minbits = z.minbits - (hdr < z.minbits ? hdr : z.minbits);
    maxbits = z.maxbits - hdr;
  }
  else {
#pragma unroll
    for (int i = 0; i < N; i++)
      q[i] = (Int)v[i];
  }
  fwd_xform<Int, D>(q);
  UInt u[N];
#pragma unroll
  for (int i = 0; i < N; i++)
    u[i] = ((UInt)q[order_at<D>(i)] + NB<UInt>::mask) ^ NB<UInt>::mask;
  u32 c;
  if constexpr (D < 4) {
    PlaneEncoder<N> pe;
    pe.init(T::P, maxbits, maxprec);
    u64 X[T::P];
    if constexpr (N == 64 && T::P == 64) {
      // high planes first; the low-plane transposes only when still coding
      planes64_half<1>(u, X);
#pragma unroll
      for (int k = 63; k >= 32; k--)
        pe.plane(w, X[k], (u32)k);
      if (pe.active() && pe.kmin < 32) {
        planes64_half<0>(u, X);
#pragma unroll
        for (int k = 31; k >= 0; k--)
          pe.plane(w, X[k], (u32)k);
      }
    }
    else {
      to_planes<N>(u, X);
#pragma unroll
      for (int k = T::P - 1; k >= 0; k--)
        pe.plane(w, X[k], (u32)k);
    }
    c = pe.pos;
  }
  else {
    c = encode_planes256<UInt, T::P>(w, u, maxbits, maxprec);
  }
  if (c < minbits) {
    w.zeros(minbits - c);
    c = minbits;
  }
  return hdr + c;
}

// ---------------------------------------------------------------------------
// decode one block (src/template/decodef.c:6-44, decodei.c:6-10,
// decode.c:281-298); returns the block's length in bits
// ---------------------------------------------------------------------------
template <typename S, int D, bool STORE>
__device__ __forceinline__ u32 decode_block(BitReader& r, S* data, const Geom& g, const u32* bc, const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  constexpr int N = 1 << (2 * D);
  u32 hdr = 0;
  u32 minbits = z.minbits, maxbits = z.maxbits, maxprec = z.maxprec;
  int emax = 0;
  if constexpr (T::FLOAT) {
    if (!r.read(1)) {
      if (STORE) {
        S v[N];
#pragma unroll
        for (int i = 0; i < N; i++)
          v[i] = (S)0;
        scatter<S, D>(v, data, g, bc);
      }
      return z.minbits > 1 ? z.minbits : 1;
    }
    emax = (int)r.read(T::EBITS) - T::EBIAS;
    hdr = 1 + T::EBITS;
    maxprec = plane_count(emax, z.maxprec, z.minexp, D);
    minbits = z.minbits - (hdr < z.minbits ? hdr : z.minbits);
    maxbits = z.maxbits - hdr;
  }
  UInt u[N];
  u32 c;
  if constexpr (D < 4) {
    PlaneDecoder<N> pd;
    pd.init(T::P, maxbits, maxprec);
    u64 X[T::P];
    if constexpr (N == 64 && T::P == 64) {
#pragma unroll
      for (int k = 63; k >= 32; k--)
        X[k] = pd.plane(r, (u32)k);
      const bool low = pd.active() && pd.kmin < 32;
      if (low) {
#pragma unroll
        for (int k = 31; k >= 0; k--)
          X[k] = pd.plane(r, (u32)k);
      }
      if (STORE) {
#pragma unroll
        for (int i = 0; i < N; i++)
          u[i] = 0;
        coeffs64_half<1>(X, u);
        if (low)
          coeffs64_half<0>(X, u);
      }
    }
    else {
#pragma unroll
      for (int k = T::P - 1; k >= 0; k--)
        X[k] = pd.plane(r, (u32)k);
      if (STORE)
        from_planes<N>(X, u);
    }
    c = pd.used();
  }
  else {
    c = decode_planes256<UInt, T::P>(r, u, maxbits, maxprec);
  }
  if (c < minbits)
    c = minbits;
  if (STORE) {
    Int q[N];
#pragma unroll
    for (int i = 0; i < N; i++)
      q[order_at<D>(i)] = (Int)((u[i] ^ NB<UInt>::mask) - NB<UInt>::mask);
    inv_xform<Int, D>(q);
    S v[N];
    if constexpr (T::FLOAT) {
      S s;
      if constexpr (T::P == 64)
        s = exp2_exact(emax - 62);
      else
        s = exp2_exact_f(emax - 30);
#pragma unroll
      for (int i = 0; i < N; i++)
        v[i] = inv_cast_one(q[i], s);
    }
    else {
#pragma unroll
      for (int i = 0; i < N; i++)
        v[i] = (S)q[i];
    }
    scatter<S, D>(v, data, g, bc);
  }
  return hdr + c;
}

} // namespace zfphip
