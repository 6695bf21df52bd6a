// codec_device.h -- device-side building blocks of the zfp block codec for
// gfx950: gather with partial-block padding, block-floating-point cast,
// lifting transform, negabinary reorder, bit-plane transpose and the embedded
// plane coder, plus bit I/O on 64-bit stream words.
//
// Layout: one lane owns one zfp block.  A block's 4^d values live in VGPRs
// (fully unrolled loops, compile-time indices), so the transform, the reorder
// and the 32x32 bit-matrix transposes that turn coefficients into bit planes
// are register renames plus ALU work, never memory traffic.  The plane coder
// runs as a rolled loop over planes: the plane number is the same in every
// lane of a wave, so the 32-word plane vectors are indexed through the
// register-index mode (s_set_gpr_idx) instead of being unrolled, which keeps
// the kernels small enough for the instruction cache.
//
// Every routine cites the reference code whose result it reproduces
// (LLNL zfp 1.0.1, paths relative to its tree).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace zfphip {

typedef uint64_t u64;
typedef int64_t i64;
typedef uint32_t u32;
typedef int32_t i32;

#define ZH_INLINE __device__ __forceinline__

// --------------------------------------------------------------------------
// scalar traits (src/traitsf.h, traitsd.h, traitsi.h, traitsl.h)
// --------------------------------------------------------------------------
template <typename S> struct Traits;
template <> struct Traits<float> {
  typedef i32 Int; typedef u32 UInt;
  static constexpr int P = 32, EBITS = 8, EBIAS = 127, PBITS = 5;
  static constexpr bool FLOAT = true;
};
template <> struct Traits<double> {
  typedef i64 Int; typedef u64 UInt;
  static constexpr int P = 64, EBITS = 11, EBIAS = 1023, PBITS = 6;
  static constexpr bool FLOAT = true;
};
template <> struct Traits<i32> {
  typedef i32 Int; typedef u32 UInt;
  static constexpr int P = 32, EBITS = 0, EBIAS = 0, PBITS = 5;
  static constexpr bool FLOAT = false;
};
template <> struct Traits<i64> {
  typedef i64 Int; typedef u64 UInt;
  static constexpr int P = 64, EBITS = 0, EBIAS = 0, PBITS = 6;
  static constexpr bool FLOAT = false;
};

template <typename UInt> struct NB;
template <> struct NB<u32> { static constexpr u32 mask = 0xaaaaaaaau; };
template <> struct NB<u64> { static constexpr u64 mask = 0xaaaaaaaaaaaaaaaaull; };

// --------------------------------------------------------------------------
// coefficient orderings: format constants of the stream
// (src/template/codec1.c:2-4, codec2.c:4-30, codec3.c:4-88, codec4.c:4-295)
// --------------------------------------------------------------------------
struct Order {
  static constexpr unsigned char o2[16] = {0, 1, 4, 5, 2, 8, 6, 9, 3, 12, 10, 7, 13, 11, 14, 15};
  static constexpr unsigned char o3[64] = {
    0, 1, 4, 16, 20, 17, 5, 2, 8, 32, 21, 6, 18, 24, 9, 33,
    36, 3, 12, 48, 22, 25, 37, 40, 34, 10, 7, 19, 28, 13, 49, 52,
    41, 38, 26, 23, 29, 53, 11, 35, 44, 14, 50, 56, 42, 27, 39, 45,
    30, 54, 57, 60, 51, 15, 43, 46, 58, 61, 55, 31, 62, 59, 47, 63};
  static constexpr unsigned char o4[256] = {
    0, 1, 4, 16, 64, 5, 80, 17, 68, 65, 20, 2, 8, 32, 128, 84,
    81, 69, 21, 6, 18, 66, 24, 72, 9, 96, 33, 36, 129, 132, 144, 3,
    12, 48, 192, 85, 82, 70, 22, 73, 25, 88, 37, 100, 97, 148, 145, 133,
    10, 160, 34, 136, 130, 40, 7, 19, 67, 28, 76, 13, 112, 49, 52, 193,
    196, 208, 86, 89, 101, 149, 161, 137, 41, 134, 38, 164, 26, 152, 146, 104,
    98, 74, 83, 71, 23, 77, 29, 92, 53, 116, 113, 212, 209, 197, 11, 35,
    131, 44, 140, 14, 176, 50, 56, 194, 200, 224, 90, 165, 102, 153, 150, 105,
    168, 162, 138, 42, 87, 93, 117, 213, 27, 75, 99, 39, 135, 147, 108, 45,
    141, 156, 30, 78, 177, 180, 54, 114, 120, 57, 198, 210, 216, 201, 225, 228,
    15, 240, 51, 204, 195, 60, 169, 166, 154, 106, 91, 103, 151, 109, 157, 94,
    181, 118, 121, 214, 217, 229, 163, 139, 43, 142, 46, 172, 58, 184, 178, 232,
    226, 202, 241, 205, 61, 199, 55, 244, 31, 220, 211, 124, 115, 79, 170, 167,
    155, 107, 158, 110, 173, 122, 185, 182, 233, 230, 218, 95, 245, 119, 221, 215,
    125, 242, 206, 62, 203, 59, 248, 47, 236, 227, 188, 179, 143, 171, 174, 186,
    234, 246, 222, 126, 219, 123, 249, 111, 237, 231, 189, 183, 159, 252, 243, 207,
    63, 175, 250, 187, 238, 235, 190, 253, 247, 223, 127, 254, 251, 239, 191, 255};
};

template <int D> ZH_INLINE constexpr int order_at(int i)
{
  return D == 1 ? i : D == 2 ? Order::o2[i] : D == 3 ? Order::o3[i] : Order::o4[i];
}

// --------------------------------------------------------------------------
// small helpers
// --------------------------------------------------------------------------
ZH_INLINE u64 lowmask(u32 n) { return n >= 64 ? ~0ull : ((1ull << n) - 1); }
ZH_INLINE u64 shr64(u64 x, u32 n) { return n >= 64 ? 0ull : x >> n; }
ZH_INLINE u64 shl64(u64 x, u32 n) { return n >= 64 ? 0ull : x << n; }
ZH_INLINE u32 ctz64(u64 x) { return (u32)__builtin_ctzll(x); }
// trailing zeros of x, 64 when x is 0
ZH_INLINE u32 ctz64z(u64 x) { return x ? (u32)__builtin_ctzll(x) : 64u; }
ZH_INLINE u32 umin(u32 a, u32 b) { return a < b ? a : b; }
// true when the predicate holds in any active lane of the wave
ZH_INLINE bool wave_any(bool p) { return __any(p) != 0; }
// true when the predicate holds in every active lane of the wave
ZH_INLINE bool wave_all(bool p) { return __any(!p) == 0; }

// 32 words held in VGPRs; an index that is uniform across the wave selects
// a register without unrolling
typedef u32 PlaneVec __attribute__((ext_vector_type(32)));

// --------------------------------------------------------------------------
// block geometry
// --------------------------------------------------------------------------
struct Geom {
  i64 s[4];   // strides in elements
  u32 n[4];   // extents (1 for unused dims)
  u32 nb[4];  // blocks per dim
  u32 dims;
  // lane order: neighbouring lanes take neighbouring blocks along axis
  // lax[0] (the one of smallest |stride|), then lax[1], ...; lnb[i] = nb[lax[i]]
  u32 lax[4];
  u32 lnb[4];
};

// The block of lane t in lane order: its coordinates bc (x first) and its
// position in stream order (x fastest, src/template/compress.c), which fixes
// where its bits go.  Walking the blocks along the axis of smallest stride
// makes a wave's loads and stores cover contiguous memory for any axis order
// of the view, such as a transposed one.
template <int D>
ZH_INLINE u32 lane_block(u32 t, const Geom& g, u32* bc)
{
  u32 c[4];
#pragma unroll
  for (int i = 0; i < D; i++) {
    if (i < D - 1) {
      c[i] = t % g.lnb[i];
      t /= g.lnb[i];
    }
    else
      c[i] = t;
  }
#pragma unroll
  for (int d = 0; d < D; d++) {
    u32 x = 0;
#pragma unroll
    for (int i = 0; i < D; i++)
      x = g.lax[i] == (u32)d ? c[i] : x;
    bc[d] = x;
  }
  u32 b = bc[D - 1];
#pragma unroll
  for (int d = D - 2; d >= 0; d--)
    b = b * g.nb[d] + bc[d];
  return b;
}

// Element offsets of the four positions of a block along one axis of stride
// s when only m (1..4) of them lie inside the array.  The reference pads x,
// then y, z, w with the patterns p1 <- p0 (m = 1), p2 <- p1 and p3 <- p0
// (m = 2), p3 <- p0 (m = 3) (src/template/encode.c:8-27); per axis these
// compose to source positions {0,0,0,0}, {0,1,1,0}, {0,1,2,0}, {0,1,2,3}.
ZH_INLINE void axis_offsets(i64 s, u32 m, i64* o)
{
  o[0] = 0;
  o[1] = m >= 2 ? s : 0;
  o[2] = m >= 3 ? 2 * s : (m == 2 ? s : 0);
  o[3] = m >= 4 ? 3 * s : 0;
}

// A row of a block is its four values along one axis.  Rows along an axis of
// stride 1 that start on a 16-byte boundary move as 16-byte vectors: one
// access per row for 4-byte scalars, two for 8-byte ones.
template <typename S> struct Row16;
template <> struct Row16<float> { typedef float T __attribute__((ext_vector_type(4))); };
template <> struct Row16<i32> { typedef i32 T __attribute__((ext_vector_type(4))); };
template <> struct Row16<double> { typedef double T __attribute__((ext_vector_type(2))); };
template <> struct Row16<i64> { typedef i64 T __attribute__((ext_vector_type(2))); };

// Whether every row of the blocks along axis a can move as 16-byte vectors:
// a has stride 1, the field starts on a 16-byte boundary, and the other
// strides keep row starts on one (a row starts at a multiple of 4 along a).
template <typename S, int D>
ZH_INLINE bool rows16(const void* data, const Geom& g, int a)
{
  bool ok = g.s[a] == 1 && ((uintptr_t)data & 15) == 0;
#pragma unroll
  for (int d = 0; d < D; d++)
    ok = ok && (d == a || ((u64)g.s[d] * sizeof(S)) % 16 == 0);
  return ok;
}

// The index in v[] (x fastest) of value i of row r of a block whose rows run
// along axis A; r numbers the other axes' positions, lowest axis fastest
template <int D, int A>
ZH_INLINE constexpr int row_value(int r, int i)
{
  int k = 0, sh = 0;
  for (int d = 0; d < D; d++) {
    if (d == A)
      k += i << (2 * d);
    else {
      k += ((r >> sh) & 3) << (2 * d);
      sh += 2;
    }
  }
  return k;
}

// The block's rows along axis A as 16-byte vectors (whole block along A,
// rows16 holds): the values of gather<S,D>, loaded in memory order
template <typename S, int D, int A>
ZH_INLINE void gather_rows16(S* v, const S* p, const i64 (*o)[4])
{
  constexpr int N = 1 << (2 * D);
  typedef typename Row16<S>::T V;
  constexpr int L = 16 / sizeof(S);
#pragma unroll
  for (int r = 0; r < N / 4; r++) {
    const S* q = p;
#pragma unroll
    for (int d = 0, sh = 0; d < D; d++)
      if (d != A) { q += o[d][(r >> sh) & 3]; sh += 2; }
#pragma unroll
    for (int h = 0; h < 4 / L; h++) {
      const V x = *(const V*)(q + h * L);
#pragma unroll
      for (int i = 0; i < L; i++)
        v[row_value<D, A>(r, h * L + i)] = x[i];
    }
  }
}

// The axis whose rows may move as 16-byte vectors for this lane's block, or
// -1: the lane axis lax[0] (smallest |stride|) when rows16 holds for it and
// the block is whole along it; x only, for 4D blocks
template <typename S, int D>
ZH_INLINE int row_axis(const void* data, const Geom& g, const u32* b)
{
  const int a = D > 1 && g.lax[0] < (u32)D ? (int)g.lax[0] : 0;
  return (D < 4 || a == 0) && rows16<S, D>(data, g, a) && g.n[a] - 4 * b[a] >= 4 ? a : -1;
}

// Gather the 4^D values of block (bx,by,bz,bw) into v[] with partial-block
// padding (encode1.c:4-20, encode2.c:4-25, encode3.c:4-31, encode4.c:4-38).
// Whole and partial blocks take the same path: padding only changes the
// offset tables.  Where row_axis allows, a lane loads its rows along the
// axis of unit stride as vectors, in memory order; the order of the loads
// does not change the values.
template <typename S, int D>
ZH_INLINE void gather(S* v, const S* data, const Geom& g, const u32* b)
{
  constexpr int N = 1 << (2 * D);
  i64 o[4][4];
  const S* p = data;
#pragma unroll
  for (int d = 0; d < D; d++) {
    const u32 c = 4 * b[d];
    p += (i64)c * g.s[d];
    axis_offsets(g.s[d], umin(4u, g.n[d] - c), o[d]);
  }
  const int a = row_axis<S, D>(data, g, b);
  if (D > 1 && D < 4 && a == 1)
    gather_rows16<S, D, (D > 1 ? 1 : 0)>(v, p, o);
  else if (D > 2 && D < 4 && a == 2)
    gather_rows16<S, D, (D > 2 ? 2 : 0)>(v, p, o);
  else if (a == 0)
    gather_rows16<S, D, 0>(v, p, o);
  else {
#pragma unroll
    for (int r = 0; r < N / 4; r++) {
      const S* q = p;
      if (D > 1) q += o[1][r & 3];
      if (D > 2) q += o[2][(r >> 2) & 3];
      if (D > 3) q += o[3][r >> 4];
#pragma unroll
      for (int x = 0; x < 4; x++)
        v[4 * r + x] = q[o[0][x]];
    }
  }
}

// The block's rows along axis A as 16-byte vectors (the counterpart of
// gather_rows16; rows outside the array along the other axes are skipped)
template <typename S, int D, int A>
ZH_INLINE void scatter_rows16(const S* v, S* p, const Geom& g, const u32* m)
{
  constexpr int N = 1 << (2 * D);
  typedef typename Row16<S>::T V;
  constexpr int L = 16 / sizeof(S);
#pragma unroll
  for (int r = 0; r < N / 4; r++) {
    bool ok = true;
    S* q = p;
#pragma unroll
    for (int d = 0, sh = 0; d < D; d++)
      if (d != A) {
        const u32 c = (r >> sh) & 3;
        ok = ok && c < m[d];
        q += (i64)c * g.s[d];
        sh += 2;
      }
    if (ok) {
#pragma unroll
      for (int h = 0; h < 4 / L; h++) {
        V x;
#pragma unroll
        for (int i = 0; i < L; i++)
          x[i] = v[row_value<D, A>(r, h * L + i)];
        *(V*)(q + h * L) = x;
      }
    }
  }
}

// Write back only the valid part of a block (decode1.c:4-19 ... decode4.c:4-25).
// Where row_axis allows, a lane stores whole rows along the axis of unit
// stride as vectors, filling cache-line segments in consecutive stores
// whatever the axis order of the view.
template <typename S, int D>
ZH_INLINE void scatter(const S* v, S* data, const Geom& g, const u32* b)
{
  constexpr int N = 1 << (2 * D);
  u32 m[4] = {4, 4, 4, 4};
  S* p = data;
#pragma unroll
  for (int d = 0; d < D; d++) {
    u32 o = 4 * b[d];
    m[d] = umin(4u, g.n[d] - o);
    p += (i64)o * g.s[d];
  }
  const int a = row_axis<S, D>(data, g, b);
  if (D > 1 && D < 4 && a == 1)
    scatter_rows16<S, D, (D > 1 ? 1 : 0)>(v, p, g, m);
  else if (D > 2 && D < 4 && a == 2)
    scatter_rows16<S, D, (D > 2 ? 2 : 0)>(v, p, g, m);
  else if (a == 0)
    scatter_rows16<S, D, 0>(v, p, g, m);
  else {
#pragma unroll
    for (int r = 0; r < N / 4; r++) {
      const u32 y = r & 3, z = (r >> 2) & 3, w = r >> 4;
      bool ok = true;
      S* q = p;
      if (D > 1) { ok = ok && y < m[1]; q += (i64)y * g.s[1]; }
      if (D > 2) { ok = ok && z < m[2]; q += (i64)z * g.s[2]; }
      if (D > 3) { ok = ok && w < m[3]; q += (i64)w * g.s[3]; }
#pragma unroll
      for (u32 x = 0; x < 4; x++)
        if (ok && x < m[0])
          q[(i64)x * g.s[0]] = v[4 * r + x];
    }
  }
}

// The 64 values of w-slice ws (0..3) of a 4D block, with the padding of
// gather<S,4> (the slice's w position comes from the w axis's offset table)
template <typename S>
ZH_INLINE void gather_slice4(S* v, const S* data, const Geom& g, const u32* b, int ws)
{
  i64 o[4][4];
  const S* p = data;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const u32 c = 4 * b[d];
    p += (i64)c * g.s[d];
    axis_offsets(g.s[d], umin(4u, g.n[d] - c), o[d]);
  }
  p += o[3][ws];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const S* q = p + o[1][r & 3] + o[2][r >> 2];
#pragma unroll
    for (int x = 0; x < 4; x++)
      v[4 * r + x] = q[o[0][x]];
  }
}

// Write back the valid part of w-slice ws of a 4D block (scatter<S,4>)
template <typename S>
ZH_INLINE void scatter_slice4(const S* v, S* data, const Geom& g, const u32* b, int ws)
{
  u32 m[4];
  S* p = data;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    const u32 o = 4 * b[d];
    m[d] = umin(4u, g.n[d] - o);
    p += (i64)o * g.s[d];
  }
  p += (i64)ws * g.s[3];
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const u32 y = r & 3, z = r >> 2;
    const bool ok = (u32)ws < m[3] && y < m[1] && z < m[2];
    S* q = p + (i64)y * g.s[1] + (i64)z * g.s[2];
#pragma unroll
    for (u32 x = 0; x < 4; x++)
      if (ok && x < m[0])
        q[(i64)x * g.s[0]] = v[4 * r + x];
  }
}

// --------------------------------------------------------------------------
// block-floating-point cast, computed on the IEEE bit patterns
// --------------------------------------------------------------------------

// frexp-style exponent of the largest |x| over the block, NaNs ignored,
// 0 -> -EBIAS, subnormal -> 1-EBIAS, inf -> 0 (glibc frexp) -- the result of
// src/template/encodef.c:10-40.  IEEE magnitudes order like their bit
// patterns, so the max is an integer max of the sign-stripped bits.
// The largest |x| is an IEEE maxNum reduction, which skips NaNs as the
// reference's `max < fabs(x)` does; its biased exponent field ef then gives
// the result (ef - 1022 for normal values).
template <int N>
ZH_INLINE double max_abs(const double* v)
{
  double mx = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++)
    mx = __builtin_fmax(mx, __builtin_fabs(v[i]));
  return mx;
}

template <int N>
ZH_INLINE float max_abs(const float* v)
{
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < N; i++)
    mx = __builtin_fmaxf(mx, __builtin_fabsf(v[i]));
  return mx;
}

// the block exponent of a block whose largest |x| (NaNs skipped) is mx
ZH_INLINE int emax_of(double mx)
{
  const u64 bits = (u64)__double_as_longlong(mx);
  if (!bits)
    return -1023;
  const int ef = (int)(bits >> 52);
  if (ef == 2047)
    return 0;
  return ef == 0 ? -1022 : ef - 1022;
}

ZH_INLINE int emax_of(float mx)
{
  const u32 bits = __float_as_uint(mx);
  if (!bits)
    return -127;
  const int ef = (int)(bits >> 23);
  if (ef == 255)
    return 0;
  return ef == 0 ? -126 : ef - 126;
}

template <int N> ZH_INLINE int block_emax(const double* v) { return emax_of(max_abs<N>(v)); }
template <int N> ZH_INLINE int block_emax(const float* v) { return emax_of(max_abs<N>(v)); }

// 2^e as a double, exactly as ldexp(1.0, e) rounds it (0 below 2^-1074,
// +inf from 2^1024 up)
ZH_INLINE double exp2_exact(int e)
{
  if (e >= 1024)
    return __longlong_as_double((long long)0x7ff0000000000000ull);
  if (e >= -1022)
    return __longlong_as_double((long long)((u64)(e + 1023) << 52));
  if (e >= -1074)
    return __longlong_as_double((long long)(1ull << (e + 1074)));
  return 0.0;
}

ZH_INLINE float exp2_exact_f(int e)
{
  if (e >= 128)
    return __uint_as_float(0x7f800000u);
  if (e >= -126)
    return __uint_as_float((u32)(e + 127) << 23);
  if (e >= -149)
    return __uint_as_float(1u << (e + 149));
  return 0.0f;
}

// (Int)(s * x) with s = 2^(P-2-emax) (src/template/encodef.c:42-59).  The
// product is exact: a power-of-two scale of a value below 2^emax in
// magnitude, rounded only where it falls below 1 and truncates to 0 anyway.
// What the reference's conversion yields on x86-64 for a product outside
// the integer range, NaN (inf * 0) or infinity (inf inputs, or an infinite
// scale when emax < -961 for double, < -97 for float) is INT_MIN; the range
// test is false for NaN, so all of these take that value.
ZH_INLINE i64 fwd_cast_one(double x, double s)
{
  double y = s * x;
  return __builtin_fabs(y) < 0x1p63 ? (i64)y : (i64)0x8000000000000000ull;
}

ZH_INLINE i32 fwd_cast_one(float x, float s)
{
  float y = s * x;
  return __builtin_fabsf(y) < 0x1p31f ? (i32)y : (i32)0x80000000u;
}

// (Scalar)(2^(emax-(P-2)) * (Scalar)i): round-to-nearest-even conversion,
// then one IEEE multiply with subnormals preserved (src/template/codecf.c:15-32)
ZH_INLINE double inv_cast_one(i64 i, double s) { return s * (double)i; }
ZH_INLINE float inv_cast_one(i32 i, float s) { return s * (float)i; }

// --------------------------------------------------------------------------
// decorrelating transform (src/template/encode.c:30-56, decode.c:7-45)
// wrap-around arithmetic in the coefficient width, arithmetic right shifts
// --------------------------------------------------------------------------
template <typename Int>
ZH_INLINE void fwd_lift(Int& x, Int& y, Int& z, Int& w)
{
  typedef typename Traits<Int>::UInt U;
  x = (Int)((U)x + (U)w); x >>= 1; w = (Int)((U)w - (U)x);
  z = (Int)((U)z + (U)y); z >>= 1; y = (Int)((U)y - (U)z);
  x = (Int)((U)x + (U)z); x >>= 1; z = (Int)((U)z - (U)x);
  w = (Int)((U)w + (U)y); w >>= 1; y = (Int)((U)y - (U)w);
  w = (Int)((U)w + (U)(y >> 1)); y = (Int)((U)y - (U)(w >> 1));
}

template <typename Int>
ZH_INLINE void inv_lift(Int& x, Int& y, Int& z, Int& w)
{
  typedef typename Traits<Int>::UInt U;
  y = (Int)((U)y + (U)(w >> 1)); w = (Int)((U)w - (U)(y >> 1));
  y = (Int)((U)y + (U)w); w = (Int)((U)w - ((U)y - (U)w));
  z = (Int)((U)z + (U)x); x = (Int)((U)x - ((U)z - (U)x));
  y = (Int)((U)y + (U)z); z = (Int)((U)z - ((U)y - (U)z));
  w = (Int)((U)w + (U)x); x = (Int)((U)x - ((U)w - (U)x));
}

// forward transform: axis x, then y, z, w (encode3.c:147-164 and kin)
template <typename Int, int D>
ZH_INLINE void fwd_xform(Int* q)
{
  constexpr int N = 1 << (2 * D);
#pragma unroll
  for (int a = 0; a < D; a++) {
    const int st = 1 << (2 * a);
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i >> (2 * a)) & 3) == 0)
        fwd_lift(q[i], q[i + st], q[i + 2 * st], q[i + 3 * st]);
  }
}

// inverse transform: axes in reverse order (decode3.c:25-42 and kin)
template <typename Int, int D>
ZH_INLINE void inv_xform(Int* q)
{
  constexpr int N = 1 << (2 * D);
#pragma unroll
  for (int a = D - 1; a >= 0; a--) {
    const int st = 1 << (2 * a);
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i >> (2 * a)) & 3) == 0)
        inv_lift(q[i], q[i + st], q[i + 2 * st], q[i + 3 * st]);
  }
}

// reversible (Lorenzo) lifting in wrap-around arithmetic
// (src/template/revencode.c:6-33, revdecode.c:6-30)
template <typename Int>
ZH_INLINE void rev_fwd_lift(Int& x, Int& y, Int& z, Int& w)
{
  typedef typename Traits<Int>::UInt U;
  U a = (U)x, b = (U)y, c = (U)z, d = (U)w;
  d -= c; c -= b; b -= a;
  d -= c; c -= b;
  d -= c;
  x = (Int)a; y = (Int)b; z = (Int)c; w = (Int)d;
}

template <typename Int>
ZH_INLINE void rev_inv_lift(Int& x, Int& y, Int& z, Int& w)
{
  typedef typename Traits<Int>::UInt U;
  U a = (U)x, b = (U)y, c = (U)z, d = (U)w;
  d += c;
  c += b; d += c;
  b += a; c += b; d += c;
  x = (Int)a; y = (Int)b; z = (Int)c; w = (Int)d;
}

template <typename Int, int D>
ZH_INLINE void rev_fwd_xform(Int* q)
{
  constexpr int N = 1 << (2 * D);
#pragma unroll
  for (int a = 0; a < D; a++) {
    const int st = 1 << (2 * a);
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i >> (2 * a)) & 3) == 0)
        rev_fwd_lift(q[i], q[i + st], q[i + 2 * st], q[i + 3 * st]);
  }
}

template <typename Int, int D>
ZH_INLINE void rev_inv_xform(Int* q)
{
  constexpr int N = 1 << (2 * D);
#pragma unroll
  for (int a = D - 1; a >= 0; a--) {
    const int st = 1 << (2 * a);
#pragma unroll
    for (int i = 0; i < N; i++)
      if (((i >> (2 * a)) & 3) == 0)
        rev_inv_lift(q[i], q[i + st], q[i + 2 * st], q[i + 3 * st]);
  }
}

// --------------------------------------------------------------------------
// bit-plane transpose: 32x32 bit matrices, least significant bit = column 0.
// After transpose32(a), bit i of a[j] is bit j of the original a[i].
// --------------------------------------------------------------------------
// Stage J swaps the J x J off-diagonal sub-blocks of every 2J x 2J block:
// the bits of a[k] selected by ~M, moved down by J, trade places with the
// bits of a[k + J] selected by M.  Stages 16 and 8 move whole bytes (one
// byte permute per word on gfx950); the others are a shift and a bitfield
// insert per word.
template <int J, u32 M>
ZH_INLINE void tp_stage(u32* a)
{
#pragma unroll
  for (int k = 0; k < 32; k++) {
    if (k & J)
      continue;
    const u32 lo = a[k], hi = a[k + J];
    a[k] = (lo & M) | ((hi << J) & ~M);
    a[k + J] = ((lo >> J) & M) | (hi & ~M);
  }
}

ZH_INLINE void transpose32(u32* a)
{
  tp_stage<16, 0x0000ffffu>(a);
  tp_stage<8, 0x00ff00ffu>(a);
  tp_stage<4, 0x0f0f0f0fu>(a);
  tp_stage<2, 0x33333333u>(a);
  tp_stage<1, 0x55555555u>(a);
}

// Planes 32H .. 32H+31 of N <= 64 coefficients: bit i of lo[k] is bit
// 32H + k of coefficient i, bit i of hi[k] that of coefficient 32 + i (hi is
// zero unless N = 64).  64-bit coefficients have two halves of planes; the
// coder takes the high half first, so blocks whose budget runs out there
// never transpose the low half.
template <int N, int H, typename UInt>
ZH_INLINE void planes_half(const UInt* u, PlaneVec& lo, PlaneVec& hi)
{
  static_assert(H == 0 || sizeof(UInt) == 8, "32-bit coefficients have one half");
  if constexpr (N == 4) {
#pragma unroll
    for (int k = 0; k < 32; k++) {
      u32 x = 0;
#pragma unroll
      for (int i = 0; i < 4; i++)
        x |= (u32)((u[i] >> (32 * H + k)) & 1u) << i;
      lo[k] = x;
      hi[k] = 0;
    }
  }
  else {
    u32 a[32];
#pragma unroll
    for (int i = 0; i < 32; i++)
      a[i] = i < N ? (u32)(u[i] >> (32 * H)) : 0u;
    transpose32(a);
#pragma unroll
    for (int k = 0; k < 32; k++)
      lo[k] = a[k];
    if constexpr (N == 64) {
#pragma unroll
      for (int i = 0; i < 32; i++)
        a[i] = (u32)(u[32 + i] >> (32 * H));
      transpose32(a);
#pragma unroll
      for (int k = 0; k < 32; k++)
        hi[k] = a[k];
    }
    else {
#pragma unroll
      for (int k = 0; k < 32; k++)
        hi[k] = 0;
    }
  }
}

// inverse of planes_half: w[i] receives the 32 bits of coefficient i that
// one half of the planes holds (kept apart from the other half so that a
// 64-bit coefficient is not a register pair before both halves exist)
template <int N>
ZH_INLINE void coeffs_half(const PlaneVec& lo, const PlaneVec& hi, u32* w)
{
  if constexpr (N == 4) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      u32 x = 0;
#pragma unroll
      for (int k = 0; k < 32; k++)
        x |= ((lo[k] >> i) & 1u) << k;
      w[i] = x;
    }
  }
  else {
    u32 a[32];
#pragma unroll
    for (int k = 0; k < 32; k++)
      a[k] = lo[k];
    transpose32(a);
#pragma unroll
    for (int i = 0; i < (N < 32 ? N : 32); i++)
      w[i] = a[i];
    if constexpr (N == 64) {
#pragma unroll
      for (int k = 0; k < 32; k++)
        a[k] = hi[k];
      transpose32(a);
#pragma unroll
      for (int i = 0; i < 32; i++)
        w[32 + i] = a[i];
    }
  }
}

// --------------------------------------------------------------------------
// bit output: one lane appends to its own bit range [lo, hi) of the stream.
// Bits appended past hi are dropped, so a coder may write a budgeted block's
// unbudgeted code and let the range cut it (the budgeted code is a prefix of
// the unbudgeted one).  Words wholly inside the range are stored; boundary
// words, shared with a neighbouring block or with bits already in the
// stream, are OR-ed into memory the host has zeroed beforehand.  PLAIN
// writers (every block range whole words) store every word they own.
// --------------------------------------------------------------------------
template <bool PLAIN>
struct BitWriter {
  u64* out;
  u64 wi;    // index of the word being filled
  u64 acc;   // pending bits of word wi
  u32 fill;  // number of pending bits (< 64)
  u64 lo, hi;

  ZH_INLINE void begin(u64* o, u64 start, u64 end)
  {
    out = o; wi = start >> 6; fill = (u32)(start & 63); acc = 0; lo = start; hi = end;
  }
  ZH_INLINE void emit(u64 word)
  {
    const u64 b0 = wi << 6;
    if (PLAIN) {
      if (b0 < hi)
        out[wi] = word;
    }
    else if (b0 < hi) {
      if (b0 + 64 > hi)
        word &= lowmask((u32)(hi - b0));
      if (b0 >= lo && b0 + 64 <= hi)
        out[wi] = word;
      else if (word)
        atomicOr((unsigned long long*)(out + wi), (unsigned long long)word);
    }
    wi++;
  }
  // append the low n (<= 64) bits of v; v must have no bits at or above n
  ZH_INLINE void put(u64 v, u32 n)
  {
    acc |= v << fill;
    u32 nf = fill + n;
    if (nf >= 64) {
      emit(acc);
      acc = (v >> 1) >> (63 - fill); // bits of v that did not fit (none when fill = 0)
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

// Bit output into a lane's own run of words (a slot of LDS that is zero
// beforehand).  Every put stores both words it can touch, so the slot always
// holds every bit put so far: nothing is pending and no put branches on
// whether a word filled up.  The slot has two words past the block's own;
// puts beyond them (a budgeted coder's last plane) stay in those two.
struct SlotWriter {
  u64* slot;
  u64 acc;    // slot[wi]
  u32 wi;     // word being filled
  u32 fill;   // bits in it (< 64)
  u32 last;   // highest value of wi (slot words - 2)

  ZH_INLINE void begin(u64* s, u32 words)
  {
    slot = s; acc = 0; wi = 0; fill = 0; last = words - 2;
  }
  // append the low n (<= 64) bits of v; v must have no bits at or above n
  ZH_INLINE void put(u64 v, u32 n)
  {
    const u64 lo = acc | (v << fill);
    const u64 hi = (v >> 1) >> (63 - fill); // bits of v past the word (none when fill = 0)
    slot[wi] = lo;
    slot[wi + 1] = hi;
    const u32 nf = fill + n;
    const bool full = nf >= 64;
    acc = full ? hi : lo;
    wi = umin(wi + (full ? 1u : 0u), last);
    fill = nf & 63u;
  }
  // append n zero bits (the slot's words past wi are zero already)
  ZH_INLINE void zeros(u64 n)
  {
    const u64 nf = fill + n;
    const u64 adv = nf >> 6;
    acc = adv ? 0ull : acc;
    wi = (u32)(adv >= (u64)(last - wi) ? last : wi + adv);
    fill = (u32)(nf & 63);
  }
};

// Counts bits without writing (length pass of variable-rate modes).
struct BitCounter {
  u64 pos;
  ZH_INLINE void put(u64, u32 n) { pos += n; }
  ZH_INLINE void zeros(u64 n) { pos += n; }
};

// Sequential reader of one lane's bits with a three-word window (one word of
// lookahead so that refills do not wait on memory) and a bound on the words
// it may load.
struct BitReader {
  const u64* in;
  u64 nwords;      // readable words
  u64 next;        // next word index to load
  u64 w0, w1, w2;  // unread bits start at bit `pos` of w0
  u32 pos;

  ZH_INLINE u64 load(u64 i) const { return i < nwords ? in[i] : 0ull; }
  ZH_INLINE void begin(const u64* p, u64 n, u64 bit)
  {
    in = p; nwords = n;
    u64 wi = bit >> 6;
    w0 = load(wi); w1 = load(wi + 1); w2 = load(wi + 2); next = wi + 3; pos = (u32)(bit & 63);
  }
  // next 64 bits without consuming them
  ZH_INLINE u64 peek64() const { return (w0 >> pos) | ((w1 << 1) << (63 - pos)); }
  ZH_INLINE u64 peek(u32 n) const { return peek64() & lowmask(n); }
  ZH_INLINE void skip(u32 n) // n <= 64
  {
    pos += n;
    if (pos >= 64) {
      w0 = w1; w1 = w2; w2 = load(next++); pos -= 64;
    }
  }
  ZH_INLINE u64 read(u32 n) { u64 v = peek(n); skip(n); return v; }
  ZH_INLINE void skip_long(u64 n)
  {
    while (n >= 64) { skip(64); n -= 64; }
    if (n) skip((u32)n);
  }
};

// Reader of one block staged in LDS: the position is a bit index into the
// slot, and every peek reads the two words it spans, so reading keeps no
// window to shift and refill.  The word after the block's last must be
// readable (its bits are past the block and are masked off by the coder).
struct SlotReader {
  const u64* slot;
  u32 pos;

  ZH_INLINE void begin(const u64* s, u32 bit) { slot = s; pos = bit; }
  ZH_INLINE u64 peek64() const
  {
    const u32 i = pos >> 6, s = pos & 63u;
    return (slot[i] >> s) | ((slot[i + 1] << 1) << (63 - s));
  }
  ZH_INLINE u64 peek(u32 n) const { return peek64() & lowmask(n); }
  ZH_INLINE void skip(u32 n) { pos += n; }
  ZH_INLINE u64 read(u32 n) { u64 v = peek(n); skip(n); return v; }
  ZH_INLINE void skip_long(u64 n) { pos += (u32)n; }
};

// --------------------------------------------------------------------------
// embedded coder for N <= 64 coefficients, one bit plane at a time
// (src/template/encode.c:90-256).  Plane k of the coefficients is the word
// X[k] (bit i = bit k of coefficient i).  A plane is coded as the bits of
// the n coefficients already significant, then group tests over the rest:
// a 1 when a one-bit remains, followed by the run up to and including that
// one-bit; the run's last one-bit is implied when it is the last
// coefficient; a final 0 when none remains.
//
// Budget: when the worst case can exceed maxbits (codec.c:1-6) every emitted
// bit costs one unit and coding stops at zero, which makes the output exactly
// the first maxbits bits of the unbudgeted code.  The encoder therefore
// writes each plane's whole code, stops after the plane that reaches the
// budget, and leaves the cut to the writer's range.
// --------------------------------------------------------------------------
ZH_INLINE bool budget_binds(u32 maxbits, u32 maxprec, u32 size)
{
  return (maxprec + 1) * size - 1 > maxbits;
}

// {c0, c1} ^= v << s for a 128-bit value, v < 4, s < 128 (a bit pushed
// past position 127 is dropped)
ZH_INLINE void xor128(u64& c0, u64& c1, u64 v, u32 s)
{
  const u32 t = s & 63;
  const u64 a = v << t;
  const u64 spill = (v >> 1) >> (63 - t);
  const bool up = s >= 64;
  c0 ^= up ? 0ull : a;
  c1 ^= up ? a : spill;
}

template <int N>
struct PlaneEncoder {
  u32 n;      // coefficients already significant
  u32 pos;    // bits of unbudgeted code written
  u32 limit;  // budget (0xffffffff when it cannot bind)
  u32 kmin;   // lowest plane coded

  ZH_INLINE void init(u32 P, u32 maxbits, u32 maxprec)
  {
    n = 0;
    pos = 0;
    kmin = P > maxprec ? P - maxprec : 0;
    limit = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
  }
  // whether plane k is coded: within the precision and the budget not spent
  ZH_INLINE bool wants(u32 k) const { return k >= kmin && pos < limit; }
  // bits of the budgeted code
  ZH_INLINE u32 length() const { return umin(pos, limit); }

  // one plane (bit i of x = bit of coefficient i): the bits of the n
  // coefficients already significant, then the group-test code of the rest
  template <typename W>
  ZH_INLINE void plane(W& w, u64 x)
  {
    if (n >= (u32)N) {
      // every coefficient significant: the plane verbatim
      w.put(x, (u32)N);
      pos += (u32)N;
      return;
    }
    const u32 n0 = n;
    const u64 r = x >> n0; // n0 < N <= 64
    if (!r) {
      // no new one-bit: the verbatim bits, then a 0 group test
      w.put(x & lowmask(n0), n0 + 1);
      pos += n0 + 1;
      return;
    }
    // group-test code of the remainder: bit 0 is the first group test; the
    // j-th one-bit of r (at p) lands at p + j, followed by a 1 at p + j + 1
    // (the next group test).  With m one-bits, the highest at pm, the code
    // has pm + m + 2 bits, or pm + m when pm is the last coefficient (its
    // one-bit is implied and no group test follows it).
    const u32 m = (u32)__builtin_popcountll(r);
    const u32 pm = 63u - (u32)__builtin_clzll(r);
    const bool implied = pm + 1 >= (u32)N - n0;
    const u32 clen = implied ? pm + m : pm + m + 2;
    n = n0 + pm + 1;
    u64 c0 = 1, c1 = 0;
    if (pm + m < 63) {
      // the whole code lies below bit 64
      u32 j = 0;
      for (u64 rr = r; rr; rr &= rr - 1) {
        j++;
        c0 |= 3ull << (ctz64(rr) + j);
      }
      // drop the bits past the code: the group test after the last
      // one-bit (a 0), or that one-bit too when it is implied
      c0 &= lowmask(implied ? pm + m : pm + m + 1);
    }
    else {
      u32 j = 0, p = 0;
      for (u64 rr = r; rr; rr &= rr - 1) {
        p = ctz64(rr);
        j++;
        xor128(c0, c1, 3ull, p + j);
      }
      if (implied)
        xor128(c0, c1, 3ull, p + j);
      else
        xor128(c0, c1, 1ull, p + j + 1);
    }
    // the plane's code: the n0 verbatim bits, then the clen group-test
    // bits, appended as up to three words
    const u32 L = n0 + clen;
    w.put((x & lowmask(n0)) | (c0 << n0), umin(L, 64u));
    if (L > 64) {
      w.put(((c0 >> 1) >> (63 - n0)) | (c1 << n0), umin(L - 64, 64u));
      if (L > 128)
        w.put((c1 >> 1) >> (63 - n0), L - 128);
    }
    pos += L;
  }
};

// The same code written through a SlotWriter, with the same instructions for
// every case of a plane -- all coefficients significant, no new one-bit, new
// one-bits -- so that the lanes of a wave never split on them.  The n
// verbatim bits go out as one put, the group-test code as one (two when the
// code of some lane of the wave is longer than 64 bits).  The code's one-bits
// are placed by a loop over the plane's new one-bits, whose trip count is the
// largest in the wave.
template <int N>
struct SlotPlaneEncoder {
  u32 n;      // coefficients already significant
  u32 pos;    // bits of unbudgeted code written
  u32 limit;  // budget (0xffffffff when it cannot bind)
  u32 kmin;   // lowest plane coded

  ZH_INLINE void init(u32 P, u32 maxbits, u32 maxprec)
  {
    n = 0;
    pos = 0;
    kmin = P > maxprec ? P - maxprec : 0;
    limit = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
  }
  ZH_INLINE bool wants(u32 k) const { return k >= kmin && pos < limit; }
  ZH_INLINE u32 length() const { return umin(pos, limit); }

  ZH_INLINE void plane(SlotWriter& w, u64 x)
  {
    // the n significant coefficients' bits verbatim; r: the others' bits
    const u64 sig = lowmask(n);
    const u64 r = (x & ~sig) >> (n & 63u);
    const bool all = n >= (u32)N;
    const bool nz = r != 0;
    const u32 m = (u32)__builtin_popcountll(r);
    const u32 pm = nz ? 63u - (u32)__builtin_clzll(r) : 0u;
    const bool implied = nz && n + pm + 1 >= (u32)N;
    // group-test code: bit 0 the first test; the j-th one-bit of r (at p)
    // at p + j, then the next test (a 1) at p + j + 1; the last one-bit is
    // followed by a 0 test, or is itself implied (not written) when it is
    // the last coefficient
    const u32 keep = nz ? (implied ? pm + m : pm + m + 1) : 0u;
    const u32 glen = all ? 0u : !nz ? 1u : (implied ? pm + m : pm + m + 2);
    u64 c0 = nz ? 1ull : 0ull, c1 = 0;
    if (wave_any(nz && pm + m >= 63)) {
      // some lane's code reaches past bit 63
      u64 rr = r;
      for (u32 j = 1; wave_any(rr != 0); j++) {
        if (rr) {
          const u32 s = ctz64(rr) + j;   // < 128
          xor128(c0, c1, 3ull, s);
          rr &= rr - 1;
        }
      }
      c1 &= lowmask(keep > 64 ? keep - 64 : 0u);
      c0 &= lowmask(keep);
    }
    else {
      u64 rr = r;
      for (u32 j = 1; wave_any(rr != 0); j++) {
        if (rr) {
          c0 |= 3ull << (ctz64(rr) + j);   // < 64
          rr &= rr - 1;
        }
      }
      c0 &= lowmask(keep);
    }
    w.put(x & sig, all ? (u32)N : n);
    w.put(c0, umin(glen, 64u));
    if (wave_any(glen > 64))
      w.put(c1, glen > 64 ? glen - 64 : 0u);
    pos += (all ? (u32)N : n) + glen;
    n += nz ? pm + 1 : 0u;
  }
};

// decoder mirror (src/template/decode.c:78-278), with the reference's budget
// semantics: a scan that runs out of budget, or reaches the last
// coefficient, takes the current coefficient to be the one-bit.
template <int N>
struct PlaneDecoder {
  u32 n, budget, budget0, kmin;

  ZH_INLINE void init(u32 P, u32 maxbits, u32 maxprec)
  {
    n = 0;
    kmin = P > maxprec ? P - maxprec : 0;
    budget = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
    budget0 = budget;
  }
  ZH_INLINE bool wants(u32 k) const { return k >= kmin && budget != 0; }
  ZH_INLINE u32 used() const { return budget0 - budget; }

  // one plane; the caller checks wants(k)
  template <typename R>
  ZH_INLINE u64 plane(R& r)
  {
    // verbatim bits of the significant coefficients; when the group test
    // that follows them is a 0 (the usual end of a plane) it is taken in
    // the same step
    const u32 m = umin(n, budget);
    if (m >= 64) {
      // all 64 coefficients significant: the plane verbatim, no group test
      const u64 x = r.peek64();
      r.skip(64);
      budget -= 64;
      return x;
    }
    const u64 v = r.peek64();
    u64 x = v & lowmask(m);
    const bool stop = n < (u32)N && budget > m && !((v >> m) & 1);
    const u32 take = m + (stop ? 1u : 0u);
    budget -= take;
    if (stop) {
      r.skip(take);
      return x;
    }
    // group tests: a 1, then the run of zeros up to the next one-bit; a
    // scan that reaches the last coefficient or the end of the budget
    // stops there and that coefficient takes the one-bit.  One group test
    // and its run take at most 64 bits; they are decoded from a local
    // 64-bit window (first the verbatim step's, past its `take` bits),
    // which is refilled from the reader only when fewer bits than a test
    // may need remain in it.
    // The trip is written without branches but the refill: a 0 test
    // consumes one bit and ends the plane, a 1 test consumes its run too.
    u64 w = v;
    u32 used = take;
    bool more = budget && n < (u32)N;
    while (more) {
      const u32 av = umin((u32)N - 1 - n, budget - 1); // <= 63: scan length
      if (used + umin(av + 2, 64u) > 64) {
        r.skip(used);
        w = r.peek64();
        used = 0;
      }
      const u64 wl = w >> used;
      const u32 g = (u32)wl & 1u;
      // p1 - 1: the zeros after the test bit, at most av (a scan that
      // reaches its end takes the coefficient there)
      const u32 p1 = umin(ctz64z(wl & ~1ull), av + 1);
      // the test bit, the zeros, and the one-bit that ends them if read
      const u32 step = g ? p1 + (p1 <= av ? 1u : 0u) : 1u;
      used += step;
      budget -= step;
      x |= (u64)g << (n + p1 - 1); // n + p1 - 1 <= N - 1
      n += g ? p1 : 0u;
      more = g && budget && n < (u32)N;
    }
    r.skip(used);
    return x;
  }
};

// --------------------------------------------------------------------------
// 4D blocks (256 coefficients): planes are four 64-bit words; the coder
// follows encode_many_ints / decode_many_ints (encode.c:134-176,
// decode.c:133-184), which produce the same code as the N <= 64 variants.
// The block stays in lane-private memory.
// --------------------------------------------------------------------------
// The bit planes of 256 coefficients as a table: X[4k + c] is plane k of
// coefficients 64c .. 64c+63 (bit i = bit k of coefficient 64c + i).  Built
// with the 32x32 transposes of planes_half, one 64-coefficient chunk and
// one half at a time, so that the coder reads four words per plane instead
// of extracting a bit from each of the 256 coefficients.  low (64-bit
// coefficients): build planes 0-31 too; a block whose lowest coded plane is
// 32 or above leaves them out (and never reads them).
template <typename UInt, int P>
ZH_INLINE void planes_chunk(const UInt* u, u64* X, int c, bool low = true)
{
  PlaneVec lo, hi;
  if (P == 32 || low) {
    planes_half<64, 0>(u, lo, hi);
#pragma unroll
    for (int k = 0; k < 32; k++)
      X[4 * k + c] = (u64)lo[k] | ((u64)hi[k] << 32);
  }
  if constexpr (P == 64) {
    planes_half<64, 1>(u, lo, hi);
#pragma unroll
    for (int k = 0; k < 32; k++)
      X[4 * (32 + k) + c] = (u64)lo[k] | ((u64)hi[k] << 32);
  }
}

template <typename UInt, int P>
ZH_INLINE void planes256(const UInt* u, u64* X)
{
#pragma unroll
  for (int c = 0; c < 4; c++)
    planes_chunk<UInt, P>(u + 64 * c, X, c);
}

// inverse of planes_chunk: coefficients 64c .. 64c+63 from the plane table;
// without `low` (64-bit coefficients) planes 0-31 are taken as zero and not
// read
template <typename UInt, int P>
ZH_INLINE void coeffs_chunk(const u64* X, int c, UInt* u, bool low = true)
{
#pragma unroll
  for (int i = 0; i < 64; i++)
    u[i] = 0;
#pragma unroll
  for (int H = 0; H < P / 32; H++) {
    if (P == 64 && H == 0 && !low)
      continue;
    PlaneVec lo, hi;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const u64 x = X[4 * (32 * H + k) + c];
      lo[k] = (u32)x;
      hi[k] = (u32)(x >> 32);
    }
    u32 w[64];
    coeffs_half<64>(lo, hi, w);
#pragma unroll
    for (int i = 0; i < 64; i++)
      u[i] |= (UInt)w[i] << (32 * H);
  }
}

template <typename UInt, int P>
ZH_INLINE void coeffs256(const u64* X, UInt* u)
{
#pragma unroll
  for (int c = 0; c < 4; c++)
    coeffs_chunk<UInt, P>(X, c, u + 64 * c);
}

// bits of the 256-bit plane x starting at position n (n < 256), 64 at most
ZH_INLINE u64 bits256(const u64* x, u32 n)
{
  u32 c = n >> 6, s = n & 63;
  u64 v = x[c] >> s;
  if (s && c < 3)
    v |= x[c + 1] << (64 - s);
  return v;
}

// the embedded coder over a plane table (planes_chunk layout)
template <int P, typename W>
ZH_INLINE u32 code_planes256(W& w, const u64* X, u32 maxbits, u32 maxprec)
{
  const u32 N = 256;
  const u32 kmin = P > (int)maxprec ? P - maxprec : 0;
  u32 budget = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
  const u32 budget0 = budget;
  u32 n = 0;
  for (int k = P - 1; k >= (int)kmin && budget; k--) {
    u64 x[4];
#pragma unroll
    for (int c = 0; c < 4; c++)
      x[c] = X[4 * k + c];
    // verbatim part, 64 bits at a time
    u32 m = umin(n, budget);
    budget -= m;
    for (u32 i = 0; i < m; i += 64) {
      u32 c = umin(64u, m - i);
      w.put(x[i >> 6] & lowmask(c), c);
    }
    while (budget && n < N) {
      budget--;
      // any one-bit at or above n?
      bool any = false;
      for (u32 c = n >> 6; c < 4; c++)
        any = any || (c == (n >> 6) ? (x[c] >> (n & 63)) != 0 : x[c] != 0);
      if (!any) { w.put(0, 1); break; }
      w.put(1, 1);
      // distance p to the next one-bit
      u32 p = 0;
      for (;;) {
        u64 v = bits256(x, n + p);
        if (v) { p += ctz64(v); break; }
        p += 64 - ((n + p) & 63);
      }
      u32 cnt = (n + p < N - 1) ? p + 1 : p;
      if (cnt > budget) { w.zeros(budget); budget = 0; break; }
      if (cnt == p + 1) { w.zeros(p); w.put(1, 1); }
      else w.zeros(p);
      budget -= cnt;
      n += p + 1;
    }
  }
  return budget0 - budget;
}

// The length of code_planes256's code without writing it: per plane, the
// verbatim bits of the n significant coefficients, then (n < 256) the group
// tests of the rest: one 0 when none has a one-bit, else, with m one-bits
// and the highest at coefficient h, h - n + m + 2 bits (h - n + m when h is
// the last coefficient, whose one-bit is implied); cut to the budget, of
// which code_planes256's code is a prefix.
template <int P>
ZH_INLINE u32 count_planes256(const u64* X, u32 maxbits, u32 maxprec)
{
  const u32 N = 256;
  const u32 kmin = P > (int)maxprec ? P - maxprec : 0;
  const u32 limit = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
  u32 pos = 0, n = 0;
  for (int k = P - 1; k >= (int)kmin && pos < limit; k--) {
    if (n >= N) {
      pos += N;
      continue;
    }
    u32 m = 0, h = 0;
    bool any = false;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const u32 lo = 64u * (u32)c;
      const u64 mask = n <= lo ? ~0ull : n >= lo + 64 ? 0ull : ~0ull << (n - lo);
      const u64 v = X[4 * k + c] & mask;
      m += (u32)__builtin_popcountll(v);
      h = v ? lo + 63u - (u32)__builtin_clzll(v) : h;
      any = any || v != 0;
    }
    if (!any)
      pos += n + 1;
    else {
      pos += n + (h - n) + m + (h == N - 1 ? 0u : 2u);
      n = h + 1;
    }
  }
  return umin(pos, limit);
}

template <typename UInt, int P, typename W>
ZH_INLINE u32 encode_planes256(W& w, const UInt* u, u32 maxbits, u32 maxprec)
{
  u64 X[4 * P];
  planes256<UInt, P>(u, X);
  return code_planes256<P>(w, X, maxbits, maxprec);
}

// the decoder into a plane table (planes_chunk layout); planes that are not
// coded stay zero (without `low`, planes 32 and up only: the caller does not
// read the others).  Returns the bits consumed.
template <int P, typename R>
ZH_INLINE u32 decode_table256(R& r, u64* X, u32 maxbits, u32 maxprec, bool low = true)
{
  const u32 N = 256;
  const u32 kmin = P > (int)maxprec ? P - maxprec : 0;
  u32 budget = budget_binds(maxbits, maxprec, N) ? maxbits : 0xffffffffu;
  const u32 budget0 = budget;
  u32 n = 0;
  for (int i = P == 64 && !low ? 128 : 0; i < 4 * P; i++)
    X[i] = 0;
  for (int k = P - 1; k >= (int)kmin && budget; k--) {
    u64 x[4] = {0, 0, 0, 0};
    u32 m = umin(n, budget);
    budget -= m;
    for (u32 i = 0; i < m; i += 64) {
      u32 c = umin(64u, m - i);
      const u64 v = r.read(c);
#pragma unroll
      for (int q = 0; q < 4; q++)
        x[q] = (i >> 6) == (u32)q ? v : x[q];
    }
    while (budget && n < N) {
      budget--;
      if (!r.read(1))
        break;
      // scan, at most 64 bits per look
      for (;;) {
        u32 av = umin(umin(N - 1 - n, budget), 64u);
        if (!av)
          break;
        u64 look = r.peek(av);
        if (look) {
          u32 p = ctz64(look);
          r.skip(p + 1);
          budget -= p + 1;
          n += p;
          break;
        }
        r.skip(av);
        budget -= av;
        n += av;
      }
      const u64 bit = 1ull << (n & 63);
#pragma unroll
      for (int q = 0; q < 4; q++)
        x[q] |= (n >> 6) == (u32)q ? bit : 0ull;
      n++;
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
      X[4 * k + q] = x[q];
  }
  return budget0 - budget;
}

template <typename UInt, int P, typename R>
ZH_INLINE u32 decode_planes256(R& r, UInt* u, u32 maxbits, u32 maxprec)
{
  u64 X[4 * P];
  const u32 c = decode_table256<P>(r, X, maxbits, maxprec);
  coeffs256<UInt, P>(X, u);
  return c;
}

} // namespace zfphip
