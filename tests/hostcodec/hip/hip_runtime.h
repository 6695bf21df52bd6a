// Test-only stand-in for <hip/hip_runtime.h>: lets the codec's per-block
// device source (zfp_amd/csrc/codec_device.h, codec_block.h) compile as plain
// C++ on the host so the CPU tests can run the very same encoder and decoder
// block by block against the oracle.  Not used by the product build.
#pragma once
#include <cstdint>
#include <cstring>

#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))

static inline long long __double_as_longlong(double x) { long long v; std::memcpy(&v, &x, 8); return v; }
static inline double __longlong_as_double(long long v) { double x; std::memcpy(&x, &v, 8); return x; }
static inline unsigned int __float_as_uint(float x) { unsigned int v; std::memcpy(&v, &x, 4); return v; }
static inline float __uint_as_float(unsigned int v) { float x; std::memcpy(&x, &v, 4); return x; }
// one block at a time: the "wave" is a single lane
static inline int __any(int predicate) { return predicate; }
#define __builtin_amdgcn_readfirstlane(x) (x)
static inline unsigned long long atomicOr(unsigned long long* p, unsigned long long v)
{
  unsigned long long old = *p;
  *p = old | v;
  return old;
}
