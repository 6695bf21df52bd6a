// bitstream.cpp -- bit stream I/O over 64-bit words, least significant bit
// first, with the semantics and state layout of the reference
// (include/zfp/bitstream.inl of LLNL zfp 1.0.1; line numbers cited per
// function).  Used by the header reader/writer and by callers of the API;
// the codec itself writes streams on the device.
#include <cstdlib>
#include <cstring>

#include "zfp/bitstream.h"

typedef uint64_t word;
static const size_t WBITS = 64;

// state layout of bitstream.inl:133-143 (non-strided build)
struct bitstream {
  size_t bits;  // buffered bits, 0 <= bits < 64
  word buffer;  // buffered bits (for writing: not yet in memory)
  word* ptr;    // next word to read or write
  word* begin;
  word* end;
};

extern "C" {

const size_t stream_word_bits = WBITS;

static word load_word(bitstream* s) { return *s->ptr++; }
static void store_word(bitstream* s, word v) { *s->ptr++ = v; }

bitstream_count stream_alignment(void) { return WBITS; }
void* stream_data(const bitstream* s) { return s->begin; }
size_t stream_size(const bitstream* s) { return (size_t)(s->ptr - s->begin) * sizeof(word); }
size_t stream_capacity(const bitstream* s) { return (size_t)(s->end - s->begin) * sizeof(word); }
size_t stream_stride_block(const bitstream* s) { (void)s; return 1; }
ptrdiff_t stream_stride_delta(const bitstream* s) { (void)s; return 0; }

// bitstream.inl:225-237
uint stream_read_bit(bitstream* s)
{
  if (!s->bits) {
    s->buffer = load_word(s);
    s->bits = WBITS;
  }
  s->bits--;
  uint b = (uint)(s->buffer & 1u);
  s->buffer >>= 1;
  return b;
}

// bitstream.inl:240-250
uint stream_write_bit(bitstream* s, uint bit)
{
  s->buffer += (word)bit << s->bits;
  if (++s->bits == WBITS) {
    store_word(s, s->buffer);
    s->buffer = 0;
    s->bits = 0;
  }
  return bit;
}

// bitstream.inl:253-285: read 0 <= n <= 64 bits
uint64 stream_read_bits(bitstream* s, bitstream_count n)
{
  uint64 v;
  if (n <= s->bits) {
    v = n ? (s->buffer & (n == 64 ? ~(word)0 : (((word)1 << n) - 1))) : 0;
    s->buffer = n >= 64 ? 0 : (s->buffer >> n);
    s->bits -= n;
    return v;
  }
  // take what is buffered, refill, take the rest
  size_t have = s->bits;
  v = s->buffer;
  word w = load_word(s);
  size_t take = n - have;                 // 1..64
  v |= have < 64 ? (w << have) : 0;       // bits beyond n are masked below
  if (n < 64)
    v &= ((word)1 << n) - 1;
  s->bits = WBITS - take;
  s->buffer = take >= 64 ? 0 : (w >> take);
  return v;
}

// bitstream.inl:288-313: write the low n bits of value, return value >> n
uint64 stream_write_bits(bitstream* s, uint64 value, bitstream_count n)
{
  word low = n >= 64 ? value : (value & (((word)1 << n) - 1));
  if (s->bits + n < WBITS) {
    s->buffer |= low << s->bits;
    s->bits += n;
  }
  else {
    size_t room = WBITS - s->bits;        // 1..64
    word full = s->buffer | (s->bits < 64 ? (low << s->bits) : 0);
    store_word(s, full);
    s->buffer = room >= 64 ? 0 : (low >> room);
    s->bits = n - room;
  }
  return n >= 64 ? 0 : (value >> n);
}

// bitstream.inl:316-327
bitstream_offset stream_rtell(const bitstream* s) { return (bitstream_offset)(s->ptr - s->begin) * WBITS - s->bits; }
bitstream_offset stream_wtell(const bitstream* s) { return (bitstream_offset)(s->ptr - s->begin) * WBITS + s->bits; }

// bitstream.inl:330-336
void stream_rewind(bitstream* s)
{
  s->ptr = s->begin;
  s->buffer = 0;
  s->bits = 0;
}

// bitstream.inl:339-352
void stream_rseek(bitstream* s, bitstream_offset offset)
{
  size_t r = (size_t)(offset % WBITS);
  s->ptr = s->begin + (size_t)(offset / WBITS);
  if (r) {
    s->buffer = load_word(s) >> r;
    s->bits = WBITS - r;
  }
  else {
    s->buffer = 0;
    s->bits = 0;
  }
}

// bitstream.inl:355-370
void stream_wseek(bitstream* s, bitstream_offset offset)
{
  size_t r = (size_t)(offset % WBITS);
  s->ptr = s->begin + (size_t)(offset / WBITS);
  if (r) {
    s->buffer = *s->ptr & (((word)1 << r) - 1);
    s->bits = r;
  }
  else {
    s->buffer = 0;
    s->bits = 0;
  }
}

// bitstream.inl:373-377
void stream_skip(bitstream* s, bitstream_size n) { stream_rseek(s, stream_rtell(s) + n); }

// bitstream.inl:380-389
void stream_pad(bitstream* s, bitstream_size n)
{
  bitstream_offset b = s->bits + n;
  while (b >= WBITS) {
    store_word(s, s->buffer);
    s->buffer = 0;
    b -= WBITS;
  }
  s->bits = (size_t)b;
}

// bitstream.inl:392-399
bitstream_count stream_align(bitstream* s)
{
  size_t b = s->bits;
  if (b)
    stream_skip(s, b);
  return b;
}

// bitstream.inl:402-409
bitstream_count stream_flush(bitstream* s)
{
  size_t b = (WBITS - s->bits) % WBITS;
  if (b)
    stream_pad(s, b);
  return b;
}

// bitstream.inl:412-425
void stream_copy(bitstream* dst, bitstream* src, bitstream_size n)
{
  while (n > WBITS) {
    stream_write_bits(dst, stream_read_bits(src, WBITS), WBITS);
    n -= WBITS;
  }
  if (n)
    stream_write_bits(dst, stream_read_bits(src, (bitstream_count)n), (bitstream_count)n);
}

// bitstream.inl:441-472
bitstream* stream_open(void* buffer, size_t bytes)
{
  bitstream* s = (bitstream*)std::malloc(sizeof(bitstream));
  if (s) {
    s->begin = (word*)buffer;
    s->end = s->begin + bytes / sizeof(word);
    stream_rewind(s);
  }
  return s;
}

void stream_close(bitstream* s) { std::free(s); }

bitstream* stream_clone(const bitstream* s)
{
  bitstream* c = (bitstream*)std::malloc(sizeof(bitstream));
  if (c)
    *c = *s;
  return c;
}

} // extern "C"
