/*
 * zfp.h -- high-level C API of zfp with an MI355X execution policy.
 *
 * ABI-compatible with include/zfp.h of LLNL zfp 1.0.1 (codec 5) for the
 * compress/decompress path: the enums keep their values, the structs
 * zfp_stream, zfp_field, zfp_config and zfp_execution keep their layouts, and
 * every function declared here has the reference's name, arguments, meaning
 * and return values (citations are to src/zfp.c of the reference).
 *
 * Added: zfp_exec_hip = 3, the execution policy that runs the block codec on
 * an AMD Instinct MI355X (gfx950).  In this library zfp_compress() and
 * zfp_decompress() run only that policy; zfp_exec_serial (the default after
 * zfp_stream_open) and zfp_exec_omp are accepted by
 * zfp_stream_set_execution() but return 0 from zfp_compress/zfp_decompress,
 * the reference's "unsupported" result (src/zfp.c:1110-1113).  Serial and
 * OpenMP execution are served by the reference libzfp, into which the HIP
 * policy plugs through include/zfp_hip.h.
 */
#ifndef ZFP_H
#define ZFP_H

#include <stddef.h>
#include <stdint.h>
#include "zfp/bitstream.h"

/* version (include/zfp/version.h:5-14) */
#define ZFP_VERSION_MAJOR 1
#define ZFP_VERSION_MINOR 0
#define ZFP_VERSION_PATCH 1
#define ZFP_VERSION_TWEAK 0
#define ZFP_CODEC 5
#define ZFP_VERSION ((ZFP_VERSION_MAJOR << 12) + (ZFP_VERSION_MINOR << 8) + \
                     (ZFP_VERSION_PATCH << 4) + ZFP_VERSION_TWEAK)

/* default compression parameters (include/zfp.h:18-21) */
#define ZFP_MIN_BITS     1
#define ZFP_MAX_BITS 16658
#define ZFP_MAX_PREC    64
#define ZFP_MIN_EXP  -1074

/* header masks (include/zfp.h:24-28) */
#define ZFP_HEADER_NONE   0x0u
#define ZFP_HEADER_MAGIC  0x1u
#define ZFP_HEADER_META   0x2u
#define ZFP_HEADER_MODE   0x4u
#define ZFP_HEADER_FULL   0x7u

#define ZFP_META_NULL (UINT64_C(-1))

/* header field sizes (include/zfp.h:45-50) */
#define ZFP_MAGIC_BITS       32
#define ZFP_META_BITS        52
#define ZFP_MODE_SHORT_BITS  12
#define ZFP_MODE_LONG_BITS   64
#define ZFP_HEADER_MAX_BITS 148
#define ZFP_MODE_SHORT_MAX  ((1u << ZFP_MODE_SHORT_BITS) - 2)

#define ZFP_ROUND_FIRST (-1)
#define ZFP_ROUND_NEVER 0
#define ZFP_ROUND_LAST  1

typedef int32_t int32;
typedef int64_t int64;
typedef uint32_t uint32;

enum { zfp_false = 0, zfp_true = !zfp_false };
typedef int zfp_bool;

/* execution policy (include/zfp.h:68-72, plus zfp_exec_hip) */
typedef enum {
  zfp_exec_serial = 0,
  zfp_exec_omp    = 1,
  zfp_exec_cuda   = 2,
  zfp_exec_hip    = 3  /* MI355X (gfx950) block codec */
} zfp_exec_policy;

typedef struct {
  uint threads;
  uint chunk_size;
} zfp_exec_params_omp;

typedef struct {
  zfp_exec_policy policy;
  void* params;
} zfp_execution;

/* compressed stream (include/zfp.h:86-93) */
typedef struct {
  uint minbits;
  uint maxbits;
  uint maxprec;
  int minexp;
  bitstream* stream;
  zfp_execution exec;
} zfp_stream;

typedef enum {
  zfp_mode_null            = 0,
  zfp_mode_expert          = 1,
  zfp_mode_fixed_rate      = 2,
  zfp_mode_fixed_precision = 3,
  zfp_mode_fixed_accuracy  = 4,
  zfp_mode_reversible      = 5
} zfp_mode;

typedef struct {
  zfp_mode mode;
  union {
    double rate;
    uint precision;
    double tolerance;
    struct {
      uint minbits;
      uint maxbits;
      uint maxprec;
      int minexp;
    } expert;
  } arg;
} zfp_config;

typedef enum {
  zfp_type_none   = 0,
  zfp_type_int32  = 1,
  zfp_type_int64  = 2,
  zfp_type_float  = 3,
  zfp_type_double = 4
} zfp_type;

/* uncompressed array (include/zfp.h:131-136); data points at element 0 */
typedef struct {
  zfp_type type;
  size_t nx, ny, nz, nw;
  ptrdiff_t sx, sy, sz, sw;
  void* data;
} zfp_field;

#ifdef __cplusplus
extern "C" {
#endif

extern const uint zfp_codec_version;
extern const uint zfp_library_version;
extern const char* const zfp_version_string;

size_t zfp_type_size(zfp_type type);                                   /* zfp.c:89 */

/* compressed stream */
zfp_stream* zfp_stream_open(bitstream* stream);                         /* zfp.c:537 */
void zfp_stream_close(zfp_stream* stream);                              /* zfp.c:553 */
bitstream* zfp_stream_bit_stream(const zfp_stream* stream);             /* zfp.c:561 */
zfp_mode zfp_stream_compression_mode(const zfp_stream* stream);         /* zfp.c:567 */
double zfp_stream_rate(const zfp_stream* stream, uint dims);            /* zfp.c:611 */
uint zfp_stream_precision(const zfp_stream* stream);                    /* zfp.c:619 */
double zfp_stream_accuracy(const zfp_stream* stream);                   /* zfp.c:627 */
uint64 zfp_stream_mode(const zfp_stream* stream);                       /* zfp.c:635 */
void zfp_stream_params(const zfp_stream* stream, uint* minbits, uint* maxbits,
                       uint* maxprec, int* minexp);                     /* zfp.c:693 */
size_t zfp_stream_compressed_size(const zfp_stream* stream);            /* zfp.c:706 */
size_t zfp_stream_maximum_size(const zfp_stream* stream, const zfp_field* field); /* zfp.c:712 */
void zfp_stream_rewind(zfp_stream* stream);                             /* zfp.c:888 */
void zfp_stream_set_bit_stream(zfp_stream* stream, bitstream* bs);      /* zfp.c:745 */
void zfp_stream_set_reversible(zfp_stream* stream);                     /* zfp.c:751 */
double zfp_stream_set_rate(zfp_stream* stream, double rate, zfp_type type, uint dims,
                           zfp_bool align);                             /* zfp.c:760 */
uint zfp_stream_set_precision(zfp_stream* stream, uint precision);      /* zfp.c:787 */
double zfp_stream_set_accuracy(zfp_stream* stream, double tolerance);   /* zfp.c:797 */
zfp_mode zfp_stream_set_mode(zfp_stream* stream, uint64 mode);          /* zfp.c:814 */
zfp_bool zfp_stream_set_params(zfp_stream* stream, uint minbits, uint maxbits,
                               uint maxprec, int minexp);               /* zfp.c:864 */

/* execution policy */
zfp_exec_policy zfp_stream_execution(const zfp_stream* stream);         /* zfp.c:896 */
uint zfp_stream_omp_threads(const zfp_stream* stream);                  /* zfp.c:902 */
uint zfp_stream_omp_chunk_size(const zfp_stream* stream);               /* zfp.c:910 */
zfp_bool zfp_stream_set_execution(zfp_stream* stream, zfp_exec_policy policy); /* zfp.c:918 */
zfp_bool zfp_stream_set_omp_threads(zfp_stream* stream, uint threads);  /* zfp.c:958 */
zfp_bool zfp_stream_set_omp_chunk_size(zfp_stream* stream, uint chunk_size); /* zfp.c:967 */

/* configurations */
zfp_config zfp_config_none(void);                                       /* zfp.c:468 */
zfp_config zfp_config_rate(double rate, zfp_bool align);                /* zfp.c:476 */
zfp_config zfp_config_precision(uint precision);                        /* zfp.c:488 */
zfp_config zfp_config_accuracy(double tolerance);                       /* zfp.c:499 */
zfp_config zfp_config_reversible(void);                                 /* zfp.c:510 */
zfp_config zfp_config_expert(uint minbits, uint maxbits, uint maxprec, int minexp); /* zfp.c:518 */

/* fields */
zfp_field* zfp_field_alloc(void);                                       /* zfp.c:108 */
zfp_field* zfp_field_1d(void* pointer, zfp_type type, size_t nx);       /* zfp.c:121 */
zfp_field* zfp_field_2d(void* pointer, zfp_type type, size_t nx, size_t ny); /* zfp.c:133 */
zfp_field* zfp_field_3d(void* pointer, zfp_type type, size_t nx, size_t ny, size_t nz); /* zfp.c:146 */
zfp_field* zfp_field_4d(void* pointer, zfp_type type, size_t nx, size_t ny, size_t nz,
                        size_t nw);                                     /* zfp.c:160 */
void zfp_field_free(zfp_field* field);                                  /* zfp.c:175 */
void* zfp_field_pointer(const zfp_field* field);                        /* zfp.c:181 */
void* zfp_field_begin(const zfp_field* field);                          /* zfp.c:187 */
zfp_type zfp_field_type(const zfp_field* field);                        /* zfp.c:199 */
uint zfp_field_precision(const zfp_field* field);                       /* zfp.c:205 */
uint zfp_field_dimensionality(const zfp_field* field);                  /* zfp.c:211 */
size_t zfp_field_size(const zfp_field* field, size_t* size);            /* zfp.c:217 */
size_t zfp_field_size_bytes(const zfp_field* field);                    /* zfp.c:238 */
size_t zfp_field_blocks(const zfp_field* field);                        /* zfp.c:244 */
zfp_bool zfp_field_stride(const zfp_field* field, ptrdiff_t* stride);   /* zfp.c:260 */
zfp_bool zfp_field_is_contiguous(const zfp_field* field);               /* zfp.c:281 */
uint64 zfp_field_metadata(const zfp_field* field);                      /* zfp.c:287 */
void zfp_field_set_pointer(zfp_field* field, void* pointer);            /* zfp.c:333 */
zfp_type zfp_field_set_type(zfp_field* field, zfp_type type);           /* zfp.c:339 */
void zfp_field_set_size_1d(zfp_field* field, size_t nx);                /* zfp.c:354 */
void zfp_field_set_size_2d(zfp_field* field, size_t nx, size_t ny);     /* zfp.c:363 */
void zfp_field_set_size_3d(zfp_field* field, size_t nx, size_t ny, size_t nz); /* zfp.c:372 */
void zfp_field_set_size_4d(zfp_field* field, size_t nx, size_t ny, size_t nz, size_t nw); /* zfp.c:381 */
void zfp_field_set_stride_1d(zfp_field* field, ptrdiff_t sx);           /* zfp.c:390 */
void zfp_field_set_stride_2d(zfp_field* field, ptrdiff_t sx, ptrdiff_t sy); /* zfp.c:399 */
void zfp_field_set_stride_3d(zfp_field* field, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz); /* zfp.c:408 */
void zfp_field_set_stride_4d(zfp_field* field, ptrdiff_t sx, ptrdiff_t sy, ptrdiff_t sz,
                             ptrdiff_t sw);                             /* zfp.c:417 */
zfp_bool zfp_field_set_metadata(zfp_field* field, uint64 meta);         /* zfp.c:426 */

/* compression and decompression: nonzero return (cumulative stream bytes)
   on success, 0 when the execution policy does not support the request */
size_t zfp_compress(zfp_stream* stream, const zfp_field* field);        /* zfp.c:1052 */
size_t zfp_decompress(zfp_stream* stream, zfp_field* field);            /* zfp.c:1123 */
size_t zfp_write_header(zfp_stream* stream, const zfp_field* field, uint mask); /* zfp.c:1183 */
size_t zfp_read_header(zfp_stream* stream, zfp_field* field, uint mask);       /* zfp.c:1220 */

size_t zfp_stream_flush(zfp_stream* stream);                            /* zfp.c:876 */
size_t zfp_stream_align(zfp_stream* stream);                            /* zfp.c:882 */

#ifdef __cplusplus
}
#endif

#endif
