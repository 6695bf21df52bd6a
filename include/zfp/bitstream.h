/*
 * zfp/bitstream.h -- bit stream I/O over 64-bit words (least significant bit
 * first), ABI-compatible with include/zfp/bitstream.h of LLNL zfp 1.0.1.
 *
 * Bit i of a stream is bit (i mod 64) of little-endian word floor(i / 64)
 * (reference include/zfp/bitstream.inl:1-106).  The stream state is opaque to
 * callers; its layout (bits, buffer, ptr, begin, end) matches the reference's
 * struct bitstream (bitstream.inl:133-143) so that a stream opened by either
 * library can be handed to the other.
 */
#ifndef ZFP_BITSTREAM_H
#define ZFP_BITSTREAM_H

#include <stddef.h>
#include <stdint.h>

typedef unsigned int uint;
typedef uint64_t uint64;

/* opaque stream state */
typedef struct bitstream bitstream;

typedef uint64 bitstream_offset;          /* bit offset into a stream */
typedef bitstream_offset bitstream_size;  /* number of bits */
typedef size_t bitstream_count;           /* small number of bits (<= 64) */

#ifdef __cplusplus
extern "C" {
#endif

extern const size_t stream_word_bits; /* 64 */

bitstream* stream_open(void* buffer, size_t bytes);   /* bitstream.inl:441-455 */
void stream_close(bitstream* stream);                 /* :457-462 */
bitstream* stream_clone(const bitstream* stream);     /* :464-472 */
bitstream_count stream_alignment(void);               /* :173-177 */
void* stream_data(const bitstream* stream);           /* :180-184 */
size_t stream_size(const bitstream* stream);          /* :187-191 */
size_t stream_capacity(const bitstream* stream);      /* :194-198 */
size_t stream_stride_block(const bitstream* stream);  /* :201-210 */
ptrdiff_t stream_stride_delta(const bitstream* stream); /* :213-222 */
uint stream_read_bit(bitstream* stream);              /* :225-237 */
uint stream_write_bit(bitstream* stream, uint bit);   /* :240-250 */
uint64 stream_read_bits(bitstream* stream, bitstream_count n);               /* :253-285 */
uint64 stream_write_bits(bitstream* stream, uint64 value, bitstream_count n); /* :288-313 */
bitstream_offset stream_rtell(const bitstream* stream); /* :316-320 */
bitstream_offset stream_wtell(const bitstream* stream); /* :323-327 */
void stream_rewind(bitstream* stream);                  /* :330-336 */
void stream_rseek(bitstream* stream, bitstream_offset offset); /* :339-352 */
void stream_wseek(bitstream* stream, bitstream_offset offset); /* :355-370 */
void stream_skip(bitstream* stream, bitstream_size n);  /* :373-377 */
void stream_pad(bitstream* stream, bitstream_size n);   /* :380-389 */
bitstream_count stream_align(bitstream* stream);        /* :392-399 */
bitstream_count stream_flush(bitstream* stream);        /* :402-409 */
void stream_copy(bitstream* dst, bitstream* src, bitstream_size n); /* :412-425 */

#ifdef __cplusplus
}
#endif

#endif
