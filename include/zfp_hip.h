/*
 * zfp_hip.h -- C ABI of the MI355X (gfx950) zfp block codec.
 *
 * The high-level entry points take zfp's own structs (layouts of
 * include/zfp.h, LLNL zfp 1.0.1) and are what an execution policy of the
 * reference's zfp_compress()/zfp_decompress() dispatch table calls, the way
 * the CUDA policy calls cuda_compress()/cuda_decompress()
 * (src/cuda_zfp/cuZFP.h:9-10 via src/template/cudacompress.c:5-34).  The
 * device-level entry points take plain pointers and sizes only.
 *
 * All entry points are synchronous with respect to their results unless a
 * HIP stream argument says otherwise.  Errors are reported by return value;
 * nothing is printed.
 */
#ifndef ZFP_HIP_H
#define ZFP_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "zfp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes of the device-level entry points */
#define ZFP_HIP_OK            0
#define ZFP_HIP_EINVAL        1  /* unsupported or inconsistent arguments */
#define ZFP_HIP_ENODEV        2  /* no usable gfx950 device */
#define ZFP_HIP_ENOMEM        3  /* device allocation failed */
#define ZFP_HIP_ERUNTIME      4  /* a HIP call failed */
#define ZFP_HIP_ENOINDEX      5  /* variable-rate stream without a block index */

/*
 * Replaces cuda_compress() (src/cuda_zfp/cuZFP.h:9, src/cuda_zfp/cuZFP.cu:357-414).
 * Compresses `field` into the bit stream of `stream` starting at its current
 * write position, including bits still buffered in the stream state (e.g. a
 * header written by zfp_write_header).  Field data and stream buffer may be
 * host or device memory (detected per pointer).  All modes (fixed rate,
 * precision, accuracy, expert and reversible) are supported for 1D-4D
 * int32, int64, float and double fields with any strides.  On success the
 * stream is left flushed exactly as serial
 * zfp_compress() leaves it and the return value is the stream size in bytes;
 * on a request the policy declines, it returns 0 and leaves the stream
 * untouched.
 */
size_t zfp_hip_compress(zfp_stream* stream, const zfp_field* field);

/*
 * Replaces cuda_decompress() (src/cuda_zfp/cuZFP.h:10, cuZFP.cu:416-...).
 * Decompresses from the current read position and leaves the stream aligned
 * as serial zfp_decompress() does.  Returns the stream size in bytes
 * consumed so far (the value zfp_decompress() returns), or 0 when declined.
 * Variable-rate streams are decoded with the block index recorded when the
 * same buffer was compressed by zfp_hip_compress in this process, or with an
 * index imported by zfp_hip_index_import; without one, block offsets are
 * recovered by a sequential parse on the device.
 */
size_t zfp_hip_decompress(zfp_stream* stream, zfp_field* field);

/*
 * Device-level encode: no counterpart in the reference (its GPU policy only
 * exposes the zfp_stream entry above).  Compresses a device-resident field
 * (element 0 at d_data, extents n[0..3] with 0 for unused dimensions,
 * strides s[0..3] in elements with 0 meaning the default) with the codec
 * parameters (minbits, maxbits, maxprec, minexp) into the device-resident
 * word array d_words, starting at bit `offset`.  Bits of d_words below
 * `offset` are preserved; the stream is written up to the end of its last
 * (zero-padded) word.  *bits_out receives the number of bits written.
 * hip_stream is a hipStream_t (NULL: the default stream); the call returns
 * after the work is complete.
 */
int zfp_hip_encode(int type, const size_t n[4], const ptrdiff_t s[4], const void* d_data,
                   uint minbits, uint maxbits, uint maxprec, int minexp,
                   uint64_t* d_words, uint64_t offset, uint64_t* bits_out,
                   void* hip_stream);

/*
 * Allocation-free zfp_hip_encode: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * stream and result as zfp_hip_encode, with two differences:
 *   - the scratch of variable-rate modes (block lengths and offsets, scan
 *     tile sums) is an 80 MiB device array of the library, held by each GPU
 *     the library runs on until the process ends; fields of more than
 *     ZFP_HIP_NOALLOC_MAX_BLOCKS blocks are refused (ZFP_HIP_EINVAL);
 *   - no block index is recorded, so a later zfp_hip_decode of a
 *     variable-rate stream locates its blocks from the stream itself.
 * The call allocates no device memory (fixed-rate encoding uses no scratch
 * at all).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU race, so serialise them (the library lock does within
 * one process).
 */
#define ZFP_HIP_NOALLOC_MAX_BLOCKS 5000000u
int zfp_hip_encode_noalloc(int type, const size_t n[4], const ptrdiff_t s[4], const void* d_data,
                           uint minbits, uint maxbits, uint maxprec, int minexp,
                           uint64_t* d_words, uint64_t offset, uint64_t* bits_out,
                           void* hip_stream);

/* Device-level decode, the inverse of zfp_hip_encode.  d_words holds
   nwords readable words (the stream's own ceil((offset + bits) / 64) words
   suffice); no word at or past nwords is read.  *bits_out receives the
   number of bits consumed.  A variable-rate stream is decoded in parallel
   with the block index recorded when d_words was last compressed into (or
   one imported for it), after checking every block's length against that
   index; when they disagree, or there is no index, the block offsets are
   recovered from the stream itself.  The result is the same either way. */
int zfp_hip_decode(int type, const size_t n[4], const ptrdiff_t s[4], void* d_data,
                   uint minbits, uint maxbits, uint maxprec, int minexp,
                   const uint64_t* d_words, uint64_t nwords, uint64_t offset, uint64_t* bits_out,
                   void* hip_stream);

/*
 * Block codec for compressed arrays: replaces, for many blocks in one
 * launch, the per-block calls of zfp::codec::zfp_base::encode_block and
 * decode_block (include/zfp/codec/zfpcodec.hpp:146-190), which the
 * reference's compressed arrays make one cached block at a time.
 * d_blocks holds `count` whole blocks of 4^dims values, one after another
 * (block k at 4^dims * k values, x fastest within a block), in device
 * memory.  d_words/nwords: the device stream and its size in words.
 *   d_slots NULL: block k is block k of a stream that starts at bit
 *     `offset`: the stream zfp_hip_encode writes for the blocks as a field
 *     of extents 4 x ... x 4*count.  Any mode; variable-length streams have
 *     their block index recorded as by zfp_hip_encode.
 *   d_slots non-NULL (device array of count slot numbers): fixed rate with
 *     maxbits and offset multiples of 64 (the rates compressed arrays use);
 *     block k is written to / read from bits offset + d_slots[k] * maxbits,
 *     and no other word of the stream is touched.  A slot whose block would
 *     end past nwords fails the call (ZFP_HIP_EINVAL) before any block is
 *     coded.
 * *bits_out: bits written (consumed); with slots, count * maxbits.
 */
int zfp_hip_encode_blocks(int type, unsigned dims, size_t count, const void* d_blocks, const uint64_t* d_slots,
                          uint minbits, uint maxbits, uint maxprec, int minexp, uint64_t* d_words,
                          uint64_t nwords, uint64_t offset, uint64_t* bits_out, void* hip_stream);
int zfp_hip_decode_blocks(int type, unsigned dims, size_t count, void* d_blocks, const uint64_t* d_slots,
                          uint minbits, uint maxbits, uint maxprec, int minexp, const uint64_t* d_words,
                          uint64_t nwords, uint64_t offset, uint64_t* bits_out, void* hip_stream);

/*
 * The same block codec for compressed arrays held in host memory, the
 * batched replacement of zfp::codec::zfp_base::encode_block/decode_block
 * (include/zfp/codec/zfpcodec.hpp:146-190) used by the codec classes of
 * include/zfp/hipcodec.hpp.  blocks: `count` whole blocks of 4^dims values
 * in host memory, one after another; slots: host array of count slot
 * numbers; words/nwords: the array's host stream.  Fixed rate with maxbits
 * and offset multiples of 64 (the rates compressed arrays use): block k is
 * written to / read from bits offset + slots[k] * maxbits, and no other
 * word of the stream is touched.  A slot past nwords fails the call
 * (ZFP_HIP_EINVAL) before anything is written.  type: zfp_type (1-4).
 */
int zfp_hip_encode_blocks_host(int type, unsigned dims, size_t count, const void* blocks, const uint64_t* slots,
                               uint minbits, uint maxbits, uint maxprec, int minexp, uint64_t* words,
                               uint64_t nwords, uint64_t offset);
int zfp_hip_decode_blocks_host(int type, unsigned dims, size_t count, void* blocks, const uint64_t* slots,
                               uint minbits, uint maxbits, uint maxprec, int minexp, const uint64_t* words,
                               uint64_t nwords, uint64_t offset);

/* Number of per-block bit offsets recorded for the stream whose buffer
   begins at `stream_begin` (0 if none); and export/import of that index as
   64-bit absolute bit offsets, one per block plus the end offset, so that a
   variable-rate stream written by this library can be decoded in parallel in
   another process.  No counterpart in the reference. */
size_t zfp_hip_index_size(const void* stream_begin);
int zfp_hip_index_export(const void* stream_begin, uint64_t* offsets, size_t count);
int zfp_hip_index_import(const void* stream_begin, const uint64_t* offsets, size_t count);
/* Frees every recorded index and the codec's device scratch (block lengths,
   scan tiles, the staged bits of variable-rate compression); the next call
   that needs scratch allocates it again. */
void zfp_hip_index_clear(void);

/* The same index in the reference's hybrid4 layout
   (include/zfp/index.hpp:152-310, the index of its variable-rate compressed
   arrays): per chunk of 4 blocks a record {uint32 hi; uint16 lo[4]} (12
   bytes), block offset = (hi << 12) + lo[block % 4] bits from the first
   block; ceil(nblocks / 4) records, the last padded with zero-length
   blocks.  `start` is the bit position of the first block in the stream,
   `end_bits` the offset just past the last block (hybrid4's range()).  The
   conversions run on the host and need no device; they fail
   (ZFP_HIP_EINVAL) where the reference's index throws: a block longer than
   ZFP_MAX_BITS or an offset of 2^44 bits or more. */
int zfp_hip_index_export_hybrid4(const void* stream_begin, void* records, size_t nrecords, uint64_t* start,
                                 uint64_t* end_bits);
int zfp_hip_index_import_hybrid4(const void* stream_begin, uint64_t start, const void* records, size_t nblocks,
                                 uint64_t end_bits);
int zfp_hip_hybrid4_from_offsets(const uint64_t* offsets, size_t nblocks, void* records);
int zfp_hip_offsets_from_hybrid4(const void* records, size_t nblocks, uint64_t end_bits, uint64_t* offsets);

/* The index in the reference's hybrid8<dims> layout (include/zfp/index.hpp,
   class hybrid8): two 64-bit words per chunk of 8 blocks, 2 * ceil(nblocks
   / 8) words, as zfp::index::hybrid8<dims> stores them; blocks shorter than
   2^(8 + 2 (dims - 1)) bits.  start and end_bits as for hybrid4. */
int zfp_hip_index_export_hybrid8(const void* stream_begin, unsigned dims, uint64_t* words, size_t nwords,
                                 uint64_t* start, uint64_t* end_bits);
int zfp_hip_index_import_hybrid8(const void* stream_begin, unsigned dims, uint64_t start, const uint64_t* words,
                                 size_t nblocks, uint64_t end_bits);
int zfp_hip_hybrid8_from_offsets(unsigned dims, const uint64_t* offsets, size_t nblocks, uint64_t* words);
int zfp_hip_offsets_from_hybrid8(unsigned dims, const uint64_t* words, size_t nblocks, uint64_t end_bits,
                                 uint64_t* offsets);

/*
 * Batched codec: no counterpart in the reference, which compresses one field
 * per zfp_compress() call.  Many device-resident fields of mixed type
 * (int32, int64, float, double), dimensionality (1-4), size and mode
 * (fixed rate, precision, accuracy, expert, reversible) go through one
 * persistent kernel launch per pass (two when the batch mixes 1-3D lossy
 * jobs with 4D or reversible ones).  Reversible float and double jobs need
 * minbits <= 1.  The passes are recorded once into a HIP graph
 * and replayed by each run.  On the first run, the launch shape
 * (workgroup size, waves per CU, blocks per task) is chosen by timing a few
 * candidates; the choice is cached per batch shape.  Each job's stream is
 * the one zfp_hip_encode would write for it, bit for bit.
 */
typedef struct {
  int type;              /* zfp_type of the field */
  size_t n[4];           /* extents, 0 for unused dimensions */
  ptrdiff_t s[4];        /* strides in elements, 0 for the default */
  void* data;            /* device pointer to element 0 */
  uint minbits, maxbits, maxprec;
  int minexp;            /* codec parameters, as zfp_stream_set_params */
  uint64_t* stream;      /* device stream words */
  uint64_t offset;       /* bit offset of the field's first block */
  uint64_t nwords;       /* stream capacity in words (decompression bound) */
  uint64_t* index;       /* variable-rate jobs: device array of nblocks + 1
                            block offsets in bits relative to `offset`;
                            compression writes it when non-NULL,
                            decompression requires it */
} zfp_hip_job;

typedef struct zfp_hip_batch zfp_hip_batch;

#define ZFP_HIP_BATCH_COMPRESS   0
#define ZFP_HIP_BATCH_DECOMPRESS 1

/* Validates the jobs and builds the batch; NULL on failure (see
   zfp_hip_last_error).  The job array is copied; the buffers it points to
   must stay valid while the batch is used.  hip_stream: a hipStream_t
   (NULL: the default stream). */
zfp_hip_batch* zfp_hip_batch_create(const zfp_hip_job* jobs, size_t njobs, int direction, void* hip_stream);
/* Runs the batch and waits for it; bits[j] (when bits is non-NULL)
   receives the bits written (compression) or consumed (decompression) by
   job j.  Returns ZFP_HIP_OK or an error code. */
int zfp_hip_batch_run(zfp_hip_batch* batch, uint64_t* bits);
/* The launch shape in use (after the first run). */
int zfp_hip_batch_shape(const zfp_hip_batch* batch, unsigned* wg_size, unsigned* waves_per_cu,
                        unsigned* blocks_per_task);
/* Fixes the launch shape instead of tuning it (0 keeps tuning). */
int zfp_hip_batch_set_shape(zfp_hip_batch* batch, unsigned wg_size, unsigned waves_per_cu,
                            unsigned blocks_per_task);
void zfp_hip_batch_destroy(zfp_hip_batch* batch);

/*
 * Serving loop: no counterpart in the reference.  A scheduler thread admits
 * submitted compression and decompression requests in arrival order while
 * their declared device footprint fits the HBM budget, and gathers every
 * admitted request that is not yet running into one batch.  A batch is one
 * persistent launch per pass (the kernels of zfp_hip_batch_*).  Up to
 * `slots` batches run at once on their own HIP streams, so requests that
 * arrive while others are on the GPU start as soon as a slot and budget
 * free up.  Fields and streams are the caller's device buffers and stay
 * resident.  Each request's result is bit-identical to the same job run
 * alone.
 */
typedef struct zfp_hip_server zfp_hip_server;

/* hbm_budget: bytes that admitted requests may declare in total;
   max_batch_jobs: jobs per batch (0: 1024); slots: concurrent batches
   (0: 2).  NULL on failure. */
zfp_hip_server* zfp_hip_server_create(uint64_t hbm_budget, unsigned max_batch_jobs, unsigned slots);
/* Queues one job.  direction: ZFP_HIP_BATCH_COMPRESS or _DECOMPRESS.
   bytes: the job's device footprint for admission (field plus stream).
   hip_stream: the stream that produced the job's inputs; the job starts
   after the work queued on it so far.  *id receives the request id. */
int zfp_hip_server_submit(zfp_hip_server* server, const zfp_hip_job* job, int direction, uint64_t bytes,
                          void* hip_stream, uint64_t* id);
/* Waits for request `id`; *bits receives its bits written or consumed and
   *latency_ms the time from submission to completion.  Returns the
   request's status. */
int zfp_hip_server_wait(zfp_hip_server* server, uint64_t id, uint64_t* bits, double* latency_ms);
/* Counters since creation: requests done, batches launched, the largest
   number of requests in one batch and the most batches in flight at once. */
int zfp_hip_server_stats(zfp_hip_server* server, uint64_t* done, uint64_t* batches, uint64_t* max_batch,
                         uint64_t* max_inflight);
/* Finishes all queued work, then stops the scheduler. */
void zfp_hip_server_destroy(zfp_hip_server* server);

/* Library and device identification; 0 when a gfx950 device is usable. */
int zfp_hip_device_check(void);
const char* zfp_hip_version(void);

/* Why the last call on this thread declined or failed ("" after a success);
   the reference reports only a 0 return, this adds the reason. */
const char* zfp_hip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
