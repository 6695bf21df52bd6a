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
// embedded coding of the bit planes of N <= 64 coefficients of P bits
// (src/template/encode.c:90-256): planes from P-1 down, the high half of
// 64-bit coefficients before the low half, each half a loop over a plane
// vector with the plane number uniform across the wave.  Returns the bits of
// the budgeted code.
// ---------------------------------------------------------------------------
// The plane loops have no early exit: an exit taken on a wave-wide vote
// around the lanes' own (divergent) group-test loops makes the compiler
// treat the plane number as varying across lanes, and index the plane
// vectors lane by lane.  A plane no lane wants costs a vote and a branch.
template <int N, int P, typename UInt, typename PE, typename W>
__device__ __forceinline__ void encode_half(PE& pe, W& w, const PlaneVec& lo, const PlaneVec& hi, u32 base)
{
  for (int k = 31; k >= 0; k--) {
    const bool act = pe.wants(base + (u32)k);
    if (wave_any(act)) {
      const int ku = __builtin_amdgcn_readfirstlane(k); // uniform plane number
      const u64 x = (u64)lo[ku] | (N == 64 ? (u64)hi[ku] << 32 : 0ull);
      if (wave_all(!act || pe.n >= (u32)N)) {
        // every coefficient significant in every lane still coding: the
        // plane verbatim, nothing else
        if (act) {
          w.put(x, (u32)N);
          pe.pos += (u32)N;
        }
      }
      else if (act)
        pe.plane(w, x);
    }
  }
}

template <int N, int P, typename UInt, typename W>
__device__ __forceinline__ u32 encode_planes(W& w, const UInt* u, u32 maxbits, u32 maxprec)
{
  // a slot writer takes the straight-line coder
  typename std::conditional<std::is_same<W, SlotWriter>::value, SlotPlaneEncoder<N>, PlaneEncoder<N>>::type pe;
  pe.init(P, maxbits, maxprec);
  PlaneVec lo, hi;
  if constexpr (P == 64) {
    if (wave_any(pe.wants(63))) {
      planes_half<N, 1>(u, lo, hi);
      encode_half<N, P, UInt>(pe, w, lo, hi, 32);
    }
  }
  if (wave_any(pe.wants(31))) {
    planes_half<N, 0>(u, lo, hi);
    encode_half<N, P, UInt>(pe, w, lo, hi, 0);
  }
  return pe.length();
}

// decoding mirror (src/template/decode.c:78-278); u receives the
// coefficients' bits, zero where no plane was coded
template <int N, typename R>
__device__ __forceinline__ void decode_half(PlaneDecoder<N>& pd, R& r, PlaneVec& lo, PlaneVec& hi,
                                            u32 base)
{
  for (int k = 31; k >= 0; k--) {
    const bool act = pd.wants(base + (u32)k);
    u64 x = 0;
    if (wave_any(act)) {
      if (act)
        x = pd.plane(r);
    }
    // every plane is written (zero when not coded); the plane number is
    // uniform, which readfirstlane states to the compiler
    const int ku = __builtin_amdgcn_readfirstlane(k);
    lo[ku] = (u32)x;
    hi[ku] = (u32)(x >> 32);
  }
}

// HI_ONLY (64-bit coefficients): decode planes 63..32 only and set *more
// when the block goes on below plane 32; such a block's coefficients lack
// their low half and it has to be decoded again in full.  Leaving the low
// half out frees the registers that hold it.
template <int N, int P, typename UInt, bool STORE, bool HI_ONLY = false, typename R = BitReader>
__device__ __forceinline__ u32 decode_planes(R& r, UInt* u, u32 maxbits, u32 maxprec,
                                             bool* more = nullptr)
{
  static_assert(!HI_ONLY || P == 64, "HI_ONLY splits 64-bit coefficients");
  PlaneDecoder<N> pd;
  pd.init(P, maxbits, maxprec);
  u32 w[N];
  PlaneVec lo, hi;
  if constexpr (P == 64) {
    decode_half<N>(pd, r, lo, hi, 32);
    if (STORE)
      coeffs_half<N>(lo, hi, w);
#pragma unroll
    for (int i = 0; i < N; i++)
      u[i] = STORE ? (UInt)w[i] << (P / 2) : 0;
    if constexpr (HI_ONLY) {
      *more = pd.wants(31);
      return pd.used();
    }
  }
  else {
#pragma unroll
    for (int i = 0; i < N; i++)
      u[i] = 0;
  }
  if (wave_any(pd.wants(31))) {
    decode_half<N>(pd, r, lo, hi, 0);
    if (STORE) {
      coeffs_half<N>(lo, hi, w);
#pragma unroll
      for (int i = 0; i < N; i++)
        u[i] |= (UInt)w[i];
    }
  }
  return pd.used();
}

// ---------------------------------------------------------------------------
// 4D blocks (256 values) in four w-slices of 64: the block exponent from a
// first pass over the values; then per slice the cast and the lifts along
// x, y and z in registers (the 3D code); then the lifts along w across the
// slices, and the coefficients in zfp's order straight into the plane
// table, one 64-coefficient chunk at a time.  Lifts along different axes
// of different slices are independent, so this keeps the reference's order
// x, y, z, w (encode4.c:41-65); decoding runs w, then z, y, x per slice
// (decode4.c:27-52).  Only the w lifts, the reordering and the plane table
// go through lane-private memory.
// ---------------------------------------------------------------------------
template <typename S, typename W>
__device__ __forceinline__ u32 encode_block4(W& w, const S* data, const Geom& g, const u32* bc, const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  u32 hdr = 0;
  u32 minbits = z.minbits, maxbits = z.maxbits, maxprec = z.maxprec;
  S s = 0;
  if constexpr (T::FLOAT) {
    // the largest |x| first, then its exponent: exponent(inf) is 0, so the
    // slices' exponents cannot be combined instead
    S mx = 0;
#pragma unroll
    for (int ws = 0; ws < 4; ws++) {
      S v[64];
      gather_slice4<S>(v, data, g, bc, ws);
      const S m = max_abs<64>(v);
      mx = mx > m ? mx : m; // both are NaN-free
    }
    const int emax = emax_of(mx);
    const u32 mp = plane_count(emax, z.maxprec, z.minexp, 4);
    const u32 e = mp ? (u32)(emax + T::EBIAS) : 0u;
    hdr = e ? 1 + T::EBITS : 1;
    w.put(e ? 2 * (u64)e + 1 : 0ull, hdr);
    if constexpr (T::P == 64)
      s = exp2_exact(62 - emax);
    else
      s = exp2_exact_f(30 - emax);
    maxprec = e ? mp : 0u;
    minbits = z.minbits - umin(hdr, z.minbits);
    maxbits = z.maxbits - hdr;
  }
  Int q[256];
#pragma unroll
  for (int ws = 0; ws < 4; ws++) {
    S v[64];
    gather_slice4<S>(v, data, g, bc, ws);
    Int t[64];
#pragma unroll
    for (int i = 0; i < 64; i++) {
      if constexpr (T::FLOAT)
        t[i] = fwd_cast_one(v[i], s);
      else
        t[i] = (Int)v[i];
    }
    fwd_xform<Int, 3>(t);
#pragma unroll
    for (int i = 0; i < 64; i++)
      q[64 * ws + i] = t[i];
  }
#pragma unroll
  for (int i = 0; i < 64; i++)
    fwd_lift(q[i], q[64 + i], q[128 + i], q[192 + i]);
  u64 X[4 * T::P];
  // planes 0-31 of 64-bit coefficients only when the block codes any of them
  const bool low = T::P == 32 || maxprec > 32;
#pragma unroll
  for (int c = 0; c < 4; c++) {
    UInt u[64];
#pragma unroll
    for (int i = 0; i < 64; i++)
      u[i] = ((UInt)q[order_at<4>(64 * c + i)] + NB<UInt>::mask) ^ NB<UInt>::mask;
    planes_chunk<UInt, T::P>(u, X, c, low);
  }
  u32 c;
  if constexpr (std::is_same<W, BitCounter>::value) {
    // a length pass: the code's length in closed form per plane
    c = count_planes256<T::P>(X, maxbits, maxprec);
    w.zeros(c);
  }
  else
    c = code_planes256<T::P>(w, X, maxbits, maxprec);
  if (c < minbits) {
    w.zeros(minbits - c);
    c = minbits;
  }
  return hdr + c;
}

template <typename S, bool STORE, typename R = BitReader>
__device__ __forceinline__ u32 decode_block4(R& r, S* data, const Geom& g, const u32* bc, const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  u32 hdr = 0;
  u32 minbits = z.minbits, maxbits = z.maxbits, maxprec = z.maxprec;
  int emax = 0;
  if constexpr (T::FLOAT) {
    const bool coded = r.read(1) != 0;
    if (coded) {
      emax = (int)r.read(T::EBITS) - T::EBIAS;
      hdr = 1 + T::EBITS;
      maxprec = plane_count(emax, z.maxprec, z.minexp, 4);
    }
    else {
      hdr = 1;
      maxprec = 0;
    }
    minbits = z.minbits - umin(hdr, z.minbits);
    maxbits = z.maxbits - hdr;
  }
  u64 X[4 * T::P];
  // planes 0-31 of 64-bit coefficients only when the block codes any of them
  const bool low = T::P == 32 || maxprec > 32;
  u32 c = decode_table256<T::P>(r, X, maxbits, maxprec, low);
  if (c < minbits)
    c = minbits;
  if (STORE) {
    Int q[256];
#pragma unroll
    for (int ch = 0; ch < 4; ch++) {
      UInt u[64];
      coeffs_chunk<UInt, T::P>(X, ch, u, low);
#pragma unroll
      for (int i = 0; i < 64; i++)
        q[order_at<4>(64 * ch + i)] = (Int)((u[i] ^ NB<UInt>::mask) - NB<UInt>::mask);
    }
#pragma unroll
    for (int i = 0; i < 64; i++)
      inv_lift(q[i], q[64 + i], q[128 + i], q[192 + i]);
    S sc = 0;
    if constexpr (T::FLOAT) {
      if constexpr (T::P == 64)
        sc = exp2_exact(emax - 62);
      else
        sc = exp2_exact_f(emax - 30);
    }
#pragma unroll
    for (int ws = 0; ws < 4; ws++) {
      Int t[64];
#pragma unroll
      for (int i = 0; i < 64; i++)
        t[i] = q[64 * ws + i];
      inv_xform<Int, 3>(t);
      S v[64];
#pragma unroll
      for (int i = 0; i < 64; i++) {
        if constexpr (T::FLOAT)
          v[i] = inv_cast_one(t[i], sc);
        else
          v[i] = (S)t[i];
      }
      scatter_slice4<S>(v, data, g, bc, ws);
    }
  }
  return hdr + c;
}

// ---------------------------------------------------------------------------
// encode one block (src/template/encodef.c:62-90, encodei.c:6-15,
// encode.c:259-280); returns the block's length in bits.  A float block
// with nothing to code (header bit 0) runs the same path with no planes, so
// that the lanes of a wave do not split on it.
// ---------------------------------------------------------------------------
template <typename S, int D, typename W>
__device__ __forceinline__ u32 encode_block(W& w, const S* data, const Geom& g, const u32* bc, const Params& z)
{
  if constexpr (D == 4)
    return encode_block4<S>(w, data, g, bc, z);
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
    const int emax = block_emax<N>(v);
    const u32 mp = plane_count(emax, z.maxprec, z.minexp, D);
    const u32 e = mp ? (u32)(emax + T::EBIAS) : 0u;
    hdr = e ? 1 + T::EBITS : 1;
    w.put(e ? 2 * (u64)e + 1 : 0ull, hdr);
    S s;
    if constexpr (T::P == 64)
      s = exp2_exact(62 - emax);
    else
      s = exp2_exact_f(30 - emax);
#pragma unroll
    for (int i = 0; i < N; i++)
      q[i] = fwd_cast_one(v[i], s);
    maxprec = e ? mp : 0u;
    minbits = z.minbits - umin(hdr, z.minbits);
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
  if constexpr (D < 4)
    c = encode_planes<N, T::P>(w, u, maxbits, maxprec);
  else
    c = encode_planes256<UInt, T::P>(w, u, maxbits, maxprec);
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
template <typename S, int D, bool STORE, bool HI_ONLY = false, typename R = BitReader>
__device__ __forceinline__ u32 decode_block(R& r, S* data, const Geom& g, const u32* bc, const Params& z,
                                            bool* more = nullptr)
{
  if constexpr (D == 4)
    return decode_block4<S, STORE>(r, data, g, bc, z);
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  constexpr int N = 1 << (2 * D);
  u32 hdr = 0;
  u32 minbits = z.minbits, maxbits = z.maxbits, maxprec = z.maxprec;
  int emax = 0;
  if constexpr (T::FLOAT) {
    // header bit 0: a block of zeros, decoded as a block with no planes
    const bool coded = r.read(1) != 0;
    if (coded) {
      emax = (int)r.read(T::EBITS) - T::EBIAS;
      hdr = 1 + T::EBITS;
      maxprec = plane_count(emax, z.maxprec, z.minexp, D);
    }
    else {
      hdr = 1;
      maxprec = 0;
    }
    minbits = z.minbits - umin(hdr, z.minbits);
    maxbits = z.maxbits - hdr;
  }
  UInt u[N];
  u32 c;
  if constexpr (D < 4)
    c = decode_planes<N, T::P, UInt, STORE, HI_ONLY>(r, u, maxbits, maxprec, more);
  else
    c = decode_planes256<UInt, T::P>(r, u, maxbits, maxprec);
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

// ---------------------------------------------------------------------------
// reversible (lossless) mode, selected by minexp < ZFP_MIN_EXP
// (src/template/revencode.c, revencodef.c, revdecode.c, revdecodef.c)
// ---------------------------------------------------------------------------

// IEEE bits as two's-complement integers: negative values keep their sign
// bit and have the magnitude bits flipped (revencodef.c:29-43); the map is
// its own inverse (revdecodef.c:9-22)
template <typename S>
__device__ __forceinline__ typename Traits<S>::Int pun_float(S x)
{
  typedef typename Traits<S>::Int Int;
  if constexpr (Traits<S>::P == 64) {
    const Int b = (Int)__double_as_longlong(x);
    return b < 0 ? (Int)((u64)b ^ 0x7fffffffffffffffull) : b;
  }
  else {
    const Int b = (Int)__float_as_uint(x);
    return b < 0 ? (Int)((u32)b ^ 0x7fffffffu) : b;
  }
}

template <typename S>
__device__ __forceinline__ S unpun_float(typename Traits<S>::Int b)
{
  if constexpr (Traits<S>::P == 64) {
    const i64 x = b < 0 ? (i64)((u64)b ^ 0x7fffffffffffffffull) : b;
    return __longlong_as_double((long long)x);
  }
  else {
    const i32 x = b < 0 ? (i32)((u32)b ^ 0x7fffffffu) : b;
    return __uint_as_float((u32)x);
  }
}

// the integer part of a reversible block (revencode.c:36-79): Lorenzo
// transform, reorder, then a precision of P minus the trailing zero planes
// of all coefficients (1 at least, maxprec at most) in PBITS bits, then the
// planes
template <typename S, int D, typename W>
__device__ __forceinline__ u32 rev_encode_ints(W& w, typename Traits<S>::Int* q, u32 minbits, u32 maxbits,
                                               u32 maxprec)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  constexpr int N = 1 << (2 * D);
  rev_fwd_xform<Int, D>(q);
  UInt u[N];
  UInt all = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    u[i] = ((UInt)q[order_at<D>(i)] + NB<UInt>::mask) ^ NB<UInt>::mask;
    all |= u[i];
  }
  u32 prec = 0;
  if (all)
    prec = (u32)T::P - (u32)(T::P == 64 ? __builtin_ctzll((u64)all) : __builtin_ctz((u32)all));
  prec = umin(prec, maxprec);
  if (prec < 1)
    prec = 1;
  w.put(prec - 1, T::PBITS);
  u32 c = T::PBITS;
  if constexpr (D < 4)
    c += encode_planes<N, T::P>(w, u, maxbits - T::PBITS, prec);
  else
    c += encode_planes256<UInt, T::P>(w, u, maxbits - T::PBITS, prec);
  if (c < minbits) {
    w.zeros(minbits - c);
    c = minbits;
  }
  return c;
}

template <typename S, int D, typename W>
__device__ __forceinline__ u32 encode_block_rev(W& w, const S* data, const Geom& g, const u32* bc,
                                                const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  constexpr int N = 1 << (2 * D);
  S v[N];
  Int q[N];
  gather<S, D>(v, data, g, bc);
  if constexpr (T::FLOAT) {
    // block-floating-point when the cast reproduces every value bit for bit
    // (revencodef.c:5-15, 45-80); all-zero blocks reconstruct as +0
    const int emax = block_emax<N>(v);
    const bool zero = emax == -T::EBIAS;
    S s, si;
    if constexpr (T::P == 64) {
      s = exp2_exact(62 - emax);
      si = exp2_exact(emax - 62);
    }
    else {
      s = exp2_exact_f(30 - emax);
      si = exp2_exact_f(emax - 30);
    }
    bool exact = true;
#pragma unroll
    for (int i = 0; i < N; i++) {
      q[i] = zero ? (Int)0 : fwd_cast_one(v[i], s);
      const S rec = zero ? (S)0 : inv_cast_one(q[i], si);
      if constexpr (T::P == 64)
        exact = exact && __double_as_longlong(rec) == __double_as_longlong(v[i]);
      else
        exact = exact && __float_as_uint(rec) == __float_as_uint(v[i]);
    }
    u32 hdr;
    if (exact) {
      const u32 e = (u32)(emax + T::EBIAS);
      if (!e) {
        // all-zero block: one zero bit, and no padding on this side
        w.put(0, 1);
        return 1;
      }
      w.put(1, 2);
      w.put(e, T::EBITS);
      hdr = 2 + T::EBITS;
    }
    else {
#pragma unroll
      for (int i = 0; i < N; i++)
        q[i] = pun_float<S>(v[i]);
      w.put(3, 2);
      hdr = 2;
    }
    return hdr + rev_encode_ints<S, D>(w, q, z.minbits - umin(hdr, z.minbits), z.maxbits - hdr, z.maxprec);
  }
  else {
#pragma unroll
    for (int i = 0; i < N; i++)
      q[i] = (Int)v[i];
    return rev_encode_ints<S, D>(w, q, z.minbits, z.maxbits, z.maxprec);
  }
}

template <typename S, int D, bool STORE, typename R = BitReader>
__device__ __forceinline__ u32 rev_decode_ints(R& r, typename Traits<S>::Int* q, u32 minbits,
                                               u32 maxbits)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  typedef typename T::UInt UInt;
  constexpr int N = 1 << (2 * D);
  const u32 prec = (u32)r.read(T::PBITS) + 1;
  UInt u[N];
  u32 c = T::PBITS;
  if constexpr (D < 4)
    c += decode_planes<N, T::P, UInt, STORE>(r, u, maxbits - T::PBITS, prec);
  else
    c += decode_planes256<UInt, T::P>(r, u, maxbits - T::PBITS, prec);
  if (c < minbits)
    c = minbits;
  if (STORE) {
#pragma unroll
    for (int i = 0; i < N; i++)
      q[order_at<D>(i)] = (Int)((u[i] ^ NB<UInt>::mask) - NB<UInt>::mask);
    rev_inv_xform<Int, D>(q);
  }
  return c;
}

template <typename S, int D, bool STORE, typename R = BitReader>
__device__ __forceinline__ u32 decode_block_rev(R& r, S* data, const Geom& g, const u32* bc,
                                                const Params& z)
{
  typedef Traits<S> T;
  typedef typename T::Int Int;
  constexpr int N = 1 << (2 * D);
  Int q[N];
  S v[N];
  u32 c;
  if constexpr (T::FLOAT) {
    if (!r.read(1)) {
      // all-zero block; the decoder skips to minbits (revdecodef.c:24-59)
      if (STORE) {
#pragma unroll
        for (int i = 0; i < N; i++)
          v[i] = (S)0;
        scatter<S, D>(v, data, g, bc);
      }
      return z.minbits > 1 ? z.minbits : 1;
    }
    if (r.read(1)) {
      c = 2 + rev_decode_ints<S, D, STORE>(r, q, z.minbits - umin(2, z.minbits), z.maxbits - 2);
      if (STORE) {
#pragma unroll
        for (int i = 0; i < N; i++)
          v[i] = unpun_float<S>(q[i]);
      }
    }
    else {
      const int emax = (int)r.read(T::EBITS) - T::EBIAS;
      const u32 hdr = 2 + T::EBITS;
      c = hdr + rev_decode_ints<S, D, STORE>(r, q, z.minbits - umin(hdr, z.minbits), z.maxbits - hdr);
      if (STORE) {
        S si;
        if constexpr (T::P == 64)
          si = exp2_exact(emax - 62);
        else
          si = exp2_exact_f(emax - 30);
        const bool zero = emax == -T::EBIAS;
#pragma unroll
        for (int i = 0; i < N; i++)
          v[i] = zero ? (S)0 : inv_cast_one(q[i], si);
      }
    }
  }
  else {
    c = rev_decode_ints<S, D, STORE>(r, q, z.minbits, z.maxbits);
    if (STORE) {
#pragma unroll
      for (int i = 0; i < N; i++)
        v[i] = (S)q[i];
    }
  }
  if (STORE)
    scatter<S, D>(v, data, g, bc);
  return c;
}

// ---------------------------------------------------------------------------
// Parallel recovery of the block starts of a variable-length stream that
// arrives without an index.  Following block lengths from a position p
// gives a chain p -> p + len(p) -> ...; the true chain starts at 0.  The
// stream's bits [0, span) are cut into P segments of L bits, one lane each:
//
//   phase 0  lane i follows the chain from bit i*L (lane 0 from 0) to its
//            exit, the first chain position at or past the segment's end,
//            counting the positions before it;
//   phase 1  (rounds) lane i moves its start to its predecessor's exit of
//            the round before, t.  If its chain passes through t the two
//            chains coincide from there and its exit stands (the positions
//            before t leave its count); otherwise it follows the chain again
//            from t.  A round in which no exit changes ends the rounds.
//   phase 2  each lane writes its chain positions at their block numbers
//            (the exclusive scan of the counts).
//
// At the end of the rounds start[0] = 0 and start[i] = exit[i-1], so by
// induction every start, count and offset is the true chain's: the result
// is exact, and only the number of rounds depends on how soon chains from
// arbitrary bits meet the true one (in zfp streams typically within a few
// hundred blocks).  A float block whose header bit is 0 is one bit long
// when minbits <= 1, so a run of zero bits is a run of one-bit blocks and
// is crossed in one step.
// ---------------------------------------------------------------------------
struct ChainArgs {
  Geom g;
  Params z;
  const u64* in;
  u64 nwords, offset, span, L;
  u64* start;
  const u64* exit_prev;
  u64* exit_new;
  u64* cnt;
  const u64* base;
  u64* boffs;
  u32* changed;
  u32 P, nblocks;
};

template <typename S, int D, bool REV>
__device__ __forceinline__ void chain_lane(const ChainArgs& a, u32 i, u32 phase)
{
  const u64 bound = i + 1 == a.P ? a.span : (u64)(i + 1) * a.L;
  const bool runs = Traits<S>::FLOAT && !REV && a.z.minbits <= 1;
  u64 pos, limit, t = 0;
  bool check = false; // phase 1: following the old chain up to t
  if (phase == 0) {
    pos = (u64)i * a.L;
    limit = bound;
  }
  else if (phase == 1) {
    if (i == 0) {
      a.exit_new[0] = a.exit_prev[0];
      return;
    }
    t = a.exit_prev[i - 1];
    const u64 s = a.start[i];
    if (t >= bound) {
      // the predecessor's chain steps over the whole segment
      a.start[i] = t;
      a.cnt[i] = 0;
      a.exit_new[i] = t;
      if (t != a.exit_prev[i])
        *a.changed = 1u;
      return;
    }
    if (s == t) {
      a.exit_new[i] = a.exit_prev[i];
      return;
    }
    // s < t: does the chain from s pass through t?  s > t: start again at t
    check = s < t;
    pos = check ? s : t;
    limit = check ? t : bound;
  }
  else {
    pos = a.start[i];
    limit = bound;
  }
  u64 k = phase == 2 ? a.base[i] : 0;
  u32 bc[4] = {0, 0, 0, 0};
#pragma unroll 1
  for (int pass = 0; pass < 2; pass++) {
    u64 steps = 0;
#pragma unroll 1
    while (pos < limit) {
      BitReader r;
      r.begin(a.in, a.nwords, a.offset + pos);
      u64 len = 0, n = 1;
      if (runs) {
        const u64 v = r.peek64();
        if (!(v & 1)) {
          n = v ? ctz64(v) : 64;
          n = n < limit - pos ? n : limit - pos;
          len = n;
        }
      }
      if (!len) {
        if constexpr (REV)
          len = decode_block_rev<S, D, false>(r, (S*)nullptr, a.g, bc, a.z);
        else
          len = decode_block<S, D, false>(r, (S*)nullptr, a.g, bc, a.z);
        len = len ? len : 1; // blocks are at least one bit when maxbits >= 1
      }
      if (phase == 2)
        for (u64 j = 0; j < n && k + j <= a.nblocks; j++)
          a.boffs[k + j] = n > 1 ? pos + j : pos;
      pos += len;
      steps += n;
      k += n;
    }
    if (phase == 0) {
      a.start[i] = (u64)i * a.L;
      a.exit_new[i] = pos;
      a.cnt[i] = steps;
      return;
    }
    if (phase == 2) {
      // the first position past the segment is the next segment's first
      // start, or the end of the last block
      if (k <= a.nblocks)
        a.boffs[k] = pos;
      return;
    }
    if (check && pos == t) {
      // the chain from the old start passes through t
      a.start[i] = t;
      a.cnt[i] -= steps;
      a.exit_new[i] = a.exit_prev[i];
      return;
    }
    if (!check) {
      // followed again from t
      a.start[i] = t;
      a.cnt[i] = steps;
      a.exit_new[i] = pos;
      if (pos != a.exit_prev[i])
        *a.changed = 1u;
      return;
    }
    check = false;
    pos = t;
    limit = bound;
  }
}


} // namespace zfphip
