// zfp_api.cpp -- the high-level zfp C API (include/zfp.h) with the MI355X
// execution policy.  Field, stream, mode and header functions reproduce the
// reference's semantics (src/zfp.c of LLNL zfp 1.0.1, cited per function);
// zfp_compress/zfp_decompress dispatch zfp_exec_hip to the device codec
// (include/zfp_hip.h) and report every other policy as unsupported (0).
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "zfp.h"
#include "zfp_hip.h"

extern "C" {

const uint zfp_codec_version = ZFP_CODEC;
const uint zfp_library_version = ZFP_VERSION;
const char* const zfp_version_string = "zfp version 1.0.1 (MI355X execution policy, zfp-amd)";

static size_t umax(size_t a, size_t b) { return a > b ? a : b; }

// src/zfp.c:89-103
size_t zfp_type_size(zfp_type type)
{
  switch (type) {
    case zfp_type_int32: return 4;
    case zfp_type_int64: return 8;
    case zfp_type_float: return 4;
    case zfp_type_double: return 8;
    default: return 0;
  }
}

// ---------------------------------------------------------------------------
// fields (src/zfp.c:105-463)
// ---------------------------------------------------------------------------

zfp_field* zfp_field_alloc(void)
{
  zfp_field* f = (zfp_field*)std::malloc(sizeof(zfp_field));
  if (f) {
    f->type = zfp_type_none;
    f->nx = f->ny = f->nz = f->nw = 0;
    f->sx = f->sy = f->sz = f->sw = 0;
    f->data = nullptr;
  }
  return f;
}

static zfp_field* field_make(void* data, zfp_type type, size_t nx, size_t ny, size_t nz, size_t nw)
{
  zfp_field* f = zfp_field_alloc();
  if (f) {
    f->type = type;
    f->nx = nx; f->ny = ny; f->nz = nz; f->nw = nw;
    f->data = data;
  }
  return f;
}

zfp_field* zfp_field_1d(void* p, zfp_type t, size_t nx) { return field_make(p, t, nx, 0, 0, 0); }
zfp_field* zfp_field_2d(void* p, zfp_type t, size_t nx, size_t ny) { return field_make(p, t, nx, ny, 0, 0); }
zfp_field* zfp_field_3d(void* p, zfp_type t, size_t nx, size_t ny, size_t nz) { return field_make(p, t, nx, ny, nz, 0); }
zfp_field* zfp_field_4d(void* p, zfp_type t, size_t nx, size_t ny, size_t nz, size_t nw) { return field_make(p, t, nx, ny, nz, nw); }
void zfp_field_free(zfp_field* f) { std::free(f); }
void* zfp_field_pointer(const zfp_field* f) { return f->data; }
zfp_type zfp_field_type(const zfp_field* f) { return f->type; }
uint zfp_field_precision(const zfp_field* f) { return (uint)(8 * zfp_type_size(f->type)); }

uint zfp_field_dimensionality(const zfp_field* f)
{
  return f->nx ? f->ny ? f->nz ? f->nw ? 4 : 3 : 2 : 1 : 0;
}

// element offsets spanned by the field (src/zfp.c:18-39)
static size_t index_span(const zfp_field* f, ptrdiff_t* lo, ptrdiff_t* hi)
{
  ptrdiff_t sx = f->sx ? f->sx : 1;
  ptrdiff_t sy = f->sy ? f->sy : (ptrdiff_t)f->nx;
  ptrdiff_t sz = f->sz ? f->sz : (ptrdiff_t)(f->nx * f->ny);
  ptrdiff_t sw = f->sw ? f->sw : (ptrdiff_t)(f->nx * f->ny * f->nz);
  ptrdiff_t e[4];
  e[0] = f->nx ? sx * (ptrdiff_t)(f->nx - 1) : 0;
  e[1] = f->ny ? sy * (ptrdiff_t)(f->ny - 1) : 0;
  e[2] = f->nz ? sz * (ptrdiff_t)(f->nz - 1) : 0;
  e[3] = f->nw ? sw * (ptrdiff_t)(f->nw - 1) : 0;
  ptrdiff_t a = 0, b = 0;
  for (int i = 0; i < 4; i++) {
    if (e[i] < 0)
      a += e[i];
    else
      b += e[i];
  }
  if (lo) *lo = a;
  if (hi) *hi = b;
  return (size_t)(b - a + 1);
}

void* zfp_field_begin(const zfp_field* f)
{
  if (!f->data)
    return nullptr;
  ptrdiff_t lo;
  index_span(f, &lo, nullptr);
  return (unsigned char*)f->data + lo * (ptrdiff_t)zfp_type_size(f->type);
}

size_t zfp_field_size(const zfp_field* f, size_t* size)
{
  if (size) {
    uint d = zfp_field_dimensionality(f);
    if (d >= 4) size[3] = f->nw;
    if (d >= 3) size[2] = f->nz;
    if (d >= 2) size[1] = f->ny;
    if (d >= 1) size[0] = f->nx;
  }
  return umax(f->nx, 1) * umax(f->ny, 1) * umax(f->nz, 1) * umax(f->nw, 1);
}

size_t zfp_field_size_bytes(const zfp_field* f) { return index_span(f, nullptr, nullptr) * zfp_type_size(f->type); }

size_t zfp_field_blocks(const zfp_field* f)
{
  size_t bx = (f->nx + 3) / 4, by = (f->ny + 3) / 4, bz = (f->nz + 3) / 4, bw = (f->nw + 3) / 4;
  switch (zfp_field_dimensionality(f)) {
    case 1: return bx;
    case 2: return bx * by;
    case 3: return bx * by * bz;
    case 4: return bx * by * bz * bw;
    default: return 0;
  }
}

zfp_bool zfp_field_stride(const zfp_field* f, ptrdiff_t* stride)
{
  if (stride) {
    uint d = zfp_field_dimensionality(f);
    if (d >= 4) stride[3] = f->sw ? f->sw : (ptrdiff_t)(f->nx * f->ny * f->nz);
    if (d >= 3) stride[2] = f->sz ? f->sz : (ptrdiff_t)(f->nx * f->ny);
    if (d >= 2) stride[1] = f->sy ? f->sy : (ptrdiff_t)f->nx;
    if (d >= 1) stride[0] = f->sx ? f->sx : 1;
  }
  return f->sx || f->sy || f->sz || f->sw;
}

zfp_bool zfp_field_is_contiguous(const zfp_field* f)
{
  return index_span(f, nullptr, nullptr) == zfp_field_size(f, nullptr);
}

// 52-bit metadata: extents, dimensionality, type (src/zfp.c:287-330)
uint64 zfp_field_metadata(const zfp_field* f)
{
  uint64 meta = 0;
  uint d = zfp_field_dimensionality(f);
  const size_t n[4] = {f->nx, f->ny, f->nz, f->nw};
  static const int width[5] = {0, 48, 24, 16, 12};
  if (d < 1 || d > 4)
    return ZFP_META_NULL; // a dimensionless field has no metadata
  for (int i = (int)d - 1; i >= 0; i--) {
    uint64 v = (uint64)(n[i] - 1);
    if (v >> width[d])
      return ZFP_META_NULL;
    meta = (meta << width[d]) + v;
  }
  meta = (meta << 2) + (d - 1);
  meta = (meta << 2) + (uint64)(f->type - 1);
  return meta;
}

void zfp_field_set_pointer(zfp_field* f, void* p) { f->data = p; }

zfp_type zfp_field_set_type(zfp_field* f, zfp_type type)
{
  switch (type) {
    case zfp_type_int32: case zfp_type_int64: case zfp_type_float: case zfp_type_double:
      f->type = type;
      return type;
    default:
      return zfp_type_none;
  }
}

void zfp_field_set_size_1d(zfp_field* f, size_t nx) { f->nx = nx; f->ny = f->nz = f->nw = 0; }
void zfp_field_set_size_2d(zfp_field* f, size_t nx, size_t ny) { f->nx = nx; f->ny = ny; f->nz = f->nw = 0; }
void zfp_field_set_size_3d(zfp_field* f, size_t nx, size_t ny, size_t nz) { f->nx = nx; f->ny = ny; f->nz = nz; f->nw = 0; }
void zfp_field_set_size_4d(zfp_field* f, size_t nx, size_t ny, size_t nz, size_t nw) { f->nx = nx; f->ny = ny; f->nz = nz; f->nw = nw; }
void zfp_field_set_stride_1d(zfp_field* f, ptrdiff_t sx) { f->sx = sx; f->sy = f->sz = f->sw = 0; }
void zfp_field_set_stride_2d(zfp_field* f, ptrdiff_t sx, ptrdiff_t sy) { f->sx = sx; f->sy = sy; f->sz = f->sw = 0; }
void zfp_field_set_stride_3d(zfp_field* f, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz) { f->sx = sx; f->sy = sy; f->sz = sz; f->sw = 0; }
void zfp_field_set_stride_4d(zfp_field* f, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz, ptrdiff_t sw) { f->sx = sx; f->sy = sy; f->sz = sz; f->sw = sw; }

// src/zfp.c:426-463
zfp_bool zfp_field_set_metadata(zfp_field* f, uint64 meta)
{
  if (meta >> ZFP_META_BITS)
    return zfp_false;
  f->type = (zfp_type)((meta & 0x3u) + 1);
  meta >>= 2;
  uint d = (uint)(meta & 0x3u) + 1;
  meta >>= 2;
  size_t n[4] = {0, 0, 0, 0};
  switch (d) {
    case 1:
      // the reference keeps 32 of the 48 extent bits here
      n[0] = (size_t)(meta & 0xffffffffull) + 1;
      break;
    case 2:
      n[0] = (size_t)(meta & 0xffffffull) + 1; meta >>= 24;
      n[1] = (size_t)(meta & 0xffffffull) + 1;
      break;
    case 3:
      for (int i = 0; i < 3; i++) { n[i] = (size_t)(meta & 0xffffull) + 1; meta >>= 16; }
      break;
    case 4:
      for (int i = 0; i < 4; i++) { n[i] = (size_t)(meta & 0xfffull) + 1; meta >>= 12; }
      break;
  }
  f->nx = n[0]; f->ny = n[1]; f->nz = n[2]; f->nw = n[3];
  f->sx = f->sy = f->sz = f->sw = 0;
  return zfp_true;
}

// ---------------------------------------------------------------------------
// configurations (src/zfp.c:467-532)
// ---------------------------------------------------------------------------

zfp_config zfp_config_none(void) { zfp_config c; std::memset(&c, 0, sizeof(c)); c.mode = zfp_mode_null; return c; }
zfp_config zfp_config_rate(double rate, zfp_bool align) { zfp_config c = zfp_config_none(); c.mode = zfp_mode_fixed_rate; c.arg.rate = align ? -rate : +rate; return c; }
zfp_config zfp_config_precision(uint p) { zfp_config c = zfp_config_none(); c.mode = zfp_mode_fixed_precision; c.arg.precision = p; return c; }
zfp_config zfp_config_accuracy(double t) { zfp_config c = zfp_config_none(); c.mode = zfp_mode_fixed_accuracy; c.arg.tolerance = t; return c; }
zfp_config zfp_config_reversible(void) { zfp_config c = zfp_config_none(); c.mode = zfp_mode_reversible; return c; }
zfp_config zfp_config_expert(uint minbits, uint maxbits, uint maxprec, int minexp)
{
  zfp_config c = zfp_config_none();
  c.mode = zfp_mode_expert;
  c.arg.expert.minbits = minbits;
  c.arg.expert.maxbits = maxbits;
  c.arg.expert.maxprec = maxprec;
  c.arg.expert.minexp = minexp;
  return c;
}

// ---------------------------------------------------------------------------
// compressed streams (src/zfp.c:534-873)
// ---------------------------------------------------------------------------

zfp_stream* zfp_stream_open(bitstream* stream)
{
  zfp_stream* z = (zfp_stream*)std::malloc(sizeof(zfp_stream));
  if (z) {
    z->stream = stream;
    z->minbits = ZFP_MIN_BITS;
    z->maxbits = ZFP_MAX_BITS;
    z->maxprec = ZFP_MAX_PREC;
    z->minexp = ZFP_MIN_EXP;
    z->exec.policy = zfp_exec_serial;
    z->exec.params = nullptr;
  }
  return z;
}

void zfp_stream_close(zfp_stream* z)
{
  if (z->exec.params)
    std::free(z->exec.params);
  std::free(z);
}

bitstream* zfp_stream_bit_stream(const zfp_stream* z) { return z->stream; }

// src/zfp.c:567-608
zfp_mode zfp_stream_compression_mode(const zfp_stream* z)
{
  if (z->minbits > z->maxbits || !(0 < z->maxprec && z->maxprec <= 64))
    return zfp_mode_null;
  if (z->minbits == ZFP_MIN_BITS && z->maxbits == ZFP_MAX_BITS && z->maxprec == ZFP_MAX_PREC &&
      z->minexp == ZFP_MIN_EXP)
    return zfp_mode_expert;
  if (z->minbits == z->maxbits && 1 <= z->maxbits && z->maxbits <= ZFP_MAX_BITS &&
      z->maxprec >= ZFP_MAX_PREC && z->minexp == ZFP_MIN_EXP)
    return zfp_mode_fixed_rate;
  if (z->minbits <= ZFP_MIN_BITS && z->maxbits >= ZFP_MAX_BITS && z->maxprec >= 1 &&
      z->minexp == ZFP_MIN_EXP)
    return zfp_mode_fixed_precision;
  if (z->minbits <= ZFP_MIN_BITS && z->maxbits >= ZFP_MAX_BITS && z->maxprec >= ZFP_MAX_PREC &&
      z->minexp >= ZFP_MIN_EXP)
    return zfp_mode_fixed_accuracy;
  if (z->minbits <= ZFP_MIN_BITS && z->maxbits >= ZFP_MAX_BITS && z->maxprec >= ZFP_MAX_PREC &&
      z->minexp < ZFP_MIN_EXP)
    return zfp_mode_reversible;
  return zfp_mode_expert;
}

double zfp_stream_rate(const zfp_stream* z, uint dims)
{
  return zfp_stream_compression_mode(z) == zfp_mode_fixed_rate ? (double)z->maxbits / (1u << (2 * dims)) : 0.0;
}

uint zfp_stream_precision(const zfp_stream* z)
{
  return zfp_stream_compression_mode(z) == zfp_mode_fixed_precision ? z->maxprec : 0;
}

double zfp_stream_accuracy(const zfp_stream* z)
{
  return zfp_stream_compression_mode(z) == zfp_mode_fixed_accuracy ? std::ldexp(1.0, z->minexp) : 0.0;
}

// 12-bit short or 64-bit long mode word (src/zfp.c:635-690)
uint64 zfp_stream_mode(const zfp_stream* z)
{
  switch (zfp_stream_compression_mode(z)) {
    case zfp_mode_fixed_rate:
      if (z->maxbits <= 2048)
        return z->maxbits - 1;
      break;
    case zfp_mode_fixed_precision:
      if (z->maxprec <= 128)
        return (z->maxprec - 1) + 2048;
      break;
    case zfp_mode_fixed_accuracy:
      if (z->minexp <= 843)
        return (uint64)(z->minexp - ZFP_MIN_EXP) + (2048 + 128 + 1);
      break;
    case zfp_mode_reversible:
      return 2048 + 128;
    default:
      break;
  }
  auto clampm1 = [](uint v, uint hi) -> uint64 { return (uint64)((v < 1 ? 1 : (v > hi ? hi : v)) - 1); };
  uint64 minbits = clampm1(z->minbits, 0x8000u);
  uint64 maxbits = clampm1(z->maxbits, 0x8000u);
  uint64 maxprec = clampm1(z->maxprec, 0x0080u);
  int e = z->minexp + 16495;
  uint64 minexp = (uint64)(e < 0 ? 0 : (e > 0x7fff ? 0x7fff : e));
  uint64 mode = 0;
  mode = (mode << 15) + minexp;
  mode = (mode << 7) + maxprec;
  mode = (mode << 15) + maxbits;
  mode = (mode << 15) + minbits;
  mode = (mode << 12) + 0xfffu;
  return mode;
}

void zfp_stream_params(const zfp_stream* z, uint* minbits, uint* maxbits, uint* maxprec, int* minexp)
{
  if (minbits) *minbits = z->minbits;
  if (maxbits) *maxbits = z->maxbits;
  if (maxprec) *maxprec = z->maxprec;
  if (minexp) *minexp = z->minexp;
}

size_t zfp_stream_compressed_size(const zfp_stream* z) { return stream_size(z->stream); }

// src/zfp.c:712-742
size_t zfp_stream_maximum_size(const zfp_stream* z, const zfp_field* f)
{
  bool rev = z->minexp < ZFP_MIN_EXP;
  uint d = zfp_field_dimensionality(f);
  size_t blocks = zfp_field_blocks(f);
  uint values = 1u << (2 * d);
  uint maxbits = 0;
  if (!d)
    return 0;
  switch (f->type) {
    case zfp_type_int32: maxbits += rev ? 5 : 0; break;
    case zfp_type_int64: maxbits += rev ? 6 : 0; break;
    case zfp_type_float: maxbits += rev ? 1 + 1 + 8 + 5 : 1 + 8; break;
    case zfp_type_double: maxbits += rev ? 1 + 1 + 11 + 6 : 1 + 11; break;
    default: return 0;
  }
  uint prec = zfp_field_precision(f);
  maxbits += values - 1 + values * (z->maxprec < prec ? z->maxprec : prec);
  if (maxbits > z->maxbits) maxbits = z->maxbits;
  if (maxbits < z->minbits) maxbits = z->minbits;
  return ((ZFP_HEADER_MAX_BITS + blocks * maxbits + stream_word_bits - 1) & ~(stream_word_bits - 1)) / 8;
}

void zfp_stream_set_bit_stream(zfp_stream* z, bitstream* s) { z->stream = s; }

void zfp_stream_set_reversible(zfp_stream* z)
{
  z->minbits = ZFP_MIN_BITS;
  z->maxbits = ZFP_MAX_BITS;
  z->maxprec = ZFP_MAX_PREC;
  z->minexp = ZFP_MIN_EXP - 1;
}

// src/zfp.c:760-784
double zfp_stream_set_rate(zfp_stream* z, double rate, zfp_type type, uint dims, zfp_bool align)
{
  uint n = 1u << (2 * dims);
  uint bits = (uint)std::floor(n * rate + 0.5);
  if (type == zfp_type_float && bits < 9) bits = 9;
  if (type == zfp_type_double && bits < 12) bits = 12;
  if (align) {
    bits += (uint)stream_word_bits - 1;
    bits &= ~(uint)(stream_word_bits - 1);
  }
  z->minbits = bits;
  z->maxbits = bits;
  z->maxprec = ZFP_MAX_PREC;
  z->minexp = ZFP_MIN_EXP;
  return (double)bits / n;
}

uint zfp_stream_set_precision(zfp_stream* z, uint precision)
{
  z->minbits = ZFP_MIN_BITS;
  z->maxbits = ZFP_MAX_BITS;
  z->maxprec = precision ? (precision < ZFP_MAX_PREC ? precision : ZFP_MAX_PREC) : ZFP_MAX_PREC;
  z->minexp = ZFP_MIN_EXP;
  return z->maxprec;
}

// src/zfp.c:797-811
double zfp_stream_set_accuracy(zfp_stream* z, double tolerance)
{
  int emin = ZFP_MIN_EXP;
  if (tolerance > 0) {
    std::frexp(tolerance, &emin);
    emin--;
  }
  z->minbits = ZFP_MIN_BITS;
  z->maxbits = ZFP_MAX_BITS;
  z->maxprec = ZFP_MAX_PREC;
  z->minexp = emin;
  return tolerance > 0 ? std::ldexp(1.0, emin) : 0;
}

// src/zfp.c:814-861
zfp_mode zfp_stream_set_mode(zfp_stream* z, uint64 mode)
{
  uint minbits, maxbits, maxprec;
  int minexp;
  if (mode <= ZFP_MODE_SHORT_MAX) {
    if (mode < 2048) {
      minbits = maxbits = (uint)mode + 1;
      maxprec = ZFP_MAX_PREC;
      minexp = ZFP_MIN_EXP;
    }
    else if (mode < 2048 + 128) {
      minbits = ZFP_MIN_BITS;
      maxbits = ZFP_MAX_BITS;
      maxprec = (uint)mode + 1 - 2048;
      minexp = ZFP_MIN_EXP;
    }
    else if (mode == 2048 + 128) {
      minbits = ZFP_MIN_BITS;
      maxbits = ZFP_MAX_BITS;
      maxprec = ZFP_MAX_PREC;
      minexp = ZFP_MIN_EXP - 1;
    }
    else {
      minbits = ZFP_MIN_BITS;
      maxbits = ZFP_MAX_BITS;
      maxprec = ZFP_MAX_PREC;
      minexp = (int)mode + ZFP_MIN_EXP - (2048 + 128 + 1);
    }
  }
  else {
    mode >>= 12; minbits = (uint)(mode & 0x7fffu) + 1;
    mode >>= 15; maxbits = (uint)(mode & 0x7fffu) + 1;
    mode >>= 15; maxprec = (uint)(mode & 0x007fu) + 1;
    mode >>= 7;  minexp = (int)(mode & 0x7fffu) - 16495;
  }
  if (!zfp_stream_set_params(z, minbits, maxbits, maxprec, minexp))
    return zfp_mode_null;
  return zfp_stream_compression_mode(z);
}

zfp_bool zfp_stream_set_params(zfp_stream* z, uint minbits, uint maxbits, uint maxprec, int minexp)
{
  if (minbits > maxbits || !(0 < maxprec && maxprec <= 64))
    return zfp_false;
  z->minbits = minbits;
  z->maxbits = maxbits;
  z->maxprec = maxprec;
  z->minexp = minexp;
  return zfp_true;
}

size_t zfp_stream_flush(zfp_stream* z) { return stream_flush(z->stream); }
size_t zfp_stream_align(zfp_stream* z) { return stream_align(z->stream); }
void zfp_stream_rewind(zfp_stream* z) { stream_rewind(z->stream); }

// ---------------------------------------------------------------------------
// execution policy (src/zfp.c:893-973)
// ---------------------------------------------------------------------------

zfp_exec_policy zfp_stream_execution(const zfp_stream* z) { return z->exec.policy; }

uint zfp_stream_omp_threads(const zfp_stream* z)
{
  return z->exec.policy == zfp_exec_omp ? ((zfp_exec_params_omp*)z->exec.params)->threads : 0u;
}

uint zfp_stream_omp_chunk_size(const zfp_stream* z)
{
  return z->exec.policy == zfp_exec_omp ? ((zfp_exec_params_omp*)z->exec.params)->chunk_size : 0u;
}

zfp_bool zfp_stream_set_execution(zfp_stream* z, zfp_exec_policy policy)
{
  switch (policy) {
    case zfp_exec_serial:
    case zfp_exec_hip:
      if (z->exec.policy != policy && z->exec.params) {
        std::free(z->exec.params);
        z->exec.params = nullptr;
      }
      break;
    case zfp_exec_omp:
      if (z->exec.policy != policy) {
        if (z->exec.params)
          std::free(z->exec.params);
        zfp_exec_params_omp* p = (zfp_exec_params_omp*)std::malloc(sizeof(zfp_exec_params_omp));
        if (!p)
          return zfp_false;
        p->threads = 0;
        p->chunk_size = 0;
        z->exec.params = p;
      }
      break;
    default:
      return zfp_false; // zfp_exec_cuda and unknown policies are not built
  }
  z->exec.policy = policy;
  return zfp_true;
}

zfp_bool zfp_stream_set_omp_threads(zfp_stream* z, uint threads)
{
  if (!zfp_stream_set_execution(z, zfp_exec_omp))
    return zfp_false;
  ((zfp_exec_params_omp*)z->exec.params)->threads = threads;
  return zfp_true;
}

zfp_bool zfp_stream_set_omp_chunk_size(zfp_stream* z, uint chunk_size)
{
  if (!zfp_stream_set_execution(z, zfp_exec_omp))
    return zfp_false;
  ((zfp_exec_params_omp*)z->exec.params)->chunk_size = chunk_size;
  return zfp_true;
}

// ---------------------------------------------------------------------------
// compression (src/zfp.c:1051-1180)
// ---------------------------------------------------------------------------

static bool type_ok(zfp_type t)
{
  return t == zfp_type_int32 || t == zfp_type_int64 || t == zfp_type_float || t == zfp_type_double;
}

size_t zfp_compress(zfp_stream* z, const zfp_field* f)
{
  if (!type_ok(f->type) || !zfp_field_dimensionality(f))
    return 0;
  if (z->exec.policy != zfp_exec_hip)
    return 0; // serial/OpenMP execution is provided by the reference libzfp
  return zfp_hip_compress(z, f);
}

size_t zfp_decompress(zfp_stream* z, zfp_field* f)
{
  if (!type_ok(f->type) || !zfp_field_dimensionality(f))
    return 0;
  if (z->exec.policy != zfp_exec_hip)
    return 0;
  return zfp_hip_decompress(z, f);
}

// ---------------------------------------------------------------------------
// headers (src/zfp.c:1182-1249)
// ---------------------------------------------------------------------------

size_t zfp_write_header(zfp_stream* z, const zfp_field* f, uint mask)
{
  size_t bits = 0;
  uint64 meta = 0;
  if (mask & ZFP_HEADER_META) {
    meta = zfp_field_metadata(f);
    if (meta == ZFP_META_NULL)
      return 0;
  }
  if (mask & ZFP_HEADER_MAGIC) {
    stream_write_bits(z->stream, 'z', 8);
    stream_write_bits(z->stream, 'f', 8);
    stream_write_bits(z->stream, 'p', 8);
    stream_write_bits(z->stream, zfp_codec_version, 8);
    bits += ZFP_MAGIC_BITS;
  }
  if (mask & ZFP_HEADER_META) {
    stream_write_bits(z->stream, meta, ZFP_META_BITS);
    bits += ZFP_META_BITS;
  }
  if (mask & ZFP_HEADER_MODE) {
    uint64 mode = zfp_stream_mode(z);
    uint size = mode > ZFP_MODE_SHORT_MAX ? ZFP_MODE_LONG_BITS : ZFP_MODE_SHORT_BITS;
    stream_write_bits(z->stream, mode, size);
    bits += size;
  }
  return bits;
}

size_t zfp_read_header(zfp_stream* z, zfp_field* f, uint mask)
{
  size_t bits = 0;
  if (mask & ZFP_HEADER_MAGIC) {
    if (stream_read_bits(z->stream, 8) != 'z' || stream_read_bits(z->stream, 8) != 'f' ||
        stream_read_bits(z->stream, 8) != 'p' || stream_read_bits(z->stream, 8) != zfp_codec_version)
      return 0;
    bits += ZFP_MAGIC_BITS;
  }
  if (mask & ZFP_HEADER_META) {
    uint64 meta = stream_read_bits(z->stream, ZFP_META_BITS);
    if (!zfp_field_set_metadata(f, meta))
      return 0;
    bits += ZFP_META_BITS;
  }
  if (mask & ZFP_HEADER_MODE) {
    uint64 mode = stream_read_bits(z->stream, ZFP_MODE_SHORT_BITS);
    bits += ZFP_MODE_SHORT_BITS;
    if (mode > ZFP_MODE_SHORT_MAX) {
      uint size = ZFP_MODE_LONG_BITS - ZFP_MODE_SHORT_BITS;
      mode += stream_read_bits(z->stream, size) << ZFP_MODE_SHORT_BITS;
      bits += size;
    }
    if (zfp_stream_set_mode(z, mode) == zfp_mode_null)
      return 0;
  }
  return bits;
}

} // extern "C"
