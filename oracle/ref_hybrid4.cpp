// Checker (test infrastructure, never linked by the product): the
// reference's hybrid4 and hybrid8 block indexes (include/zfp/index.hpp) built
// from a sequence of block sizes, with its records and offsets exposed
// through a C ABI for the CPU tests.  Compiled against the reference's own
// headers where they lie (oracle/Makefile, target `ref`).
#include <cstdint>
#include <cstring>
#include "zfp.h"
#include "zfp/internal/array/exception.hpp"
#include "zfp/index.hpp"

namespace {
// read access to the reference's record array (a protected member)
struct raw_hybrid4 : zfp::index::hybrid4 {
  explicit raw_hybrid4(size_t n) : zfp::index::hybrid4(n) {}
  const void* records() const { return data; }
  static size_t record_size() { return sizeof(record); }
};
}

// Builds the reference's hybrid4 index of n blocks with the given bit sizes
// (sequential set_block_size, then flush), copies its ceil(n/4) records
// (record_size bytes each) to `records`, and writes every block offset and
// the range to offsets[0..n].  Returns 0, or 1 when the reference throws.
extern "C" int ref_hybrid4(const uint64_t* sizes, size_t n, void* records, uint64_t* offsets)
{
  try {
    raw_hybrid4 ix(n);
    for (size_t i = 0; i < n; i++)
      ix.set_block_size(i, (size_t)sizes[i]);
    ix.flush();
    std::memcpy(records, ix.records(), ((n + 3) / 4) * raw_hybrid4::record_size());
    for (size_t i = 0; i < n; i++)
      offsets[i] = ix.block_offset(i);
    offsets[n] = ix.range();
    return 0;
  }
  catch (...) {
    return 1;
  }
}

extern "C" size_t ref_hybrid4_record_size(void) { return raw_hybrid4::record_size(); }

namespace {
// read access to the reference's hybrid8 words (a protected member)
template <uint dims>
struct raw_hybrid8 : zfp::index::hybrid8<dims> {
  explicit raw_hybrid8(size_t n) : zfp::index::hybrid8<dims>(n) {}
  const void* words() const { return this->data; }
};

template <uint dims>
int build8(const uint64_t* sizes, size_t n, uint64_t* words, uint64_t* offsets)
{
  raw_hybrid8<dims> ix(n);
  for (size_t i = 0; i < n; i++)
    ix.set_block_size(i, (size_t)sizes[i]);
  ix.flush();
  std::memcpy(words, ix.words(), 2 * ((n + 7) / 8) * sizeof(uint64_t));
  for (size_t i = 0; i < n; i++)
    offsets[i] = ix.block_offset(i);
  offsets[n] = ix.range();
  return 0;
}
}

// The reference's hybrid8<dims> index (include/zfp/index.hpp:312-...) of n
// blocks: its 2 * ceil(n/8) words and every block offset and the range.
// Returns 0, 1 when the reference throws, 2 for an unsupported dims.
extern "C" int ref_hybrid8(unsigned dims, const uint64_t* sizes, size_t n, uint64_t* words, uint64_t* offsets)
{
  try {
    switch (dims) {
      case 1: return build8<1>(sizes, n, words, offsets);
      case 2: return build8<2>(sizes, n, words, offsets);
      case 3: return build8<3>(sizes, n, words, offsets);
      case 4: return build8<4>(sizes, n, words, offsets);
    }
    return 2;
  }
  catch (...) {
    return 1;
  }
}
