// batch.hip -- host side of the batched codec (include/zfp_hip.h): job
// checks, task lists, device buffers, the HIP graph that holds the passes,
// and the tuner that picks the persistent launch's shape by timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "batch_kernels.h"

namespace zfphip {

// Zeroes the stream words that blocks of the listed jobs OR into (all words
// of their ranges; the first word keeps the bits below the job's offset).
// Every thread of the grid takes a share of every job.  The jobs' word
// ranges are read in parallel into LDS, PREP_CHUNK jobs at a time, so that
// the walk over the jobs waits on LDS, not on two dependent global loads
// per job.
constexpr u32 PREP_CHUNK = 256;

__global__ void __launch_bounds__(256) k_batch_prepare(const BJob* __restrict__ jobs, const u32* __restrict__ pjobs,
                                                       u32 npjobs, const u64* __restrict__ offs)
{
  __shared__ u64* s_words[PREP_CHUNK];
  __shared__ u64 s_w0[PREP_CHUNK], s_w1[PREP_CHUNK];
  const u64 gt = (u64)blockIdx.x * blockDim.x + threadIdx.x;
  const u64 gn = (u64)gridDim.x * blockDim.x;
  for (u32 base = 0; base < npjobs; base += PREP_CHUNK) {
    const u32 cnt = umin(PREP_CHUNK, npjobs - base);
    for (u32 i = threadIdx.x; i < cnt; i += blockDim.x) {
      const BJob& J = jobs[pjobs[base + i]];
      const u64 bits =
          J.mode == BM_FIXED ? (u64)J.nblocks * J.z.maxbits : offs[J.vbase + J.nblocks] - offs[J.vbase];
      s_words[i] = J.words;
      s_w0[i] = J.offset; // bit offset; the word range follows below
      s_w1[i] = (J.offset + bits + 63) >> 6;
    }
    __syncthreads();
    for (u32 k = 0; k < cnt; k++) {
      u64* words = s_words[k];
      const u64 off = s_w0[k], w0 = off >> 6, w1 = s_w1[k];
      const u32 keep = (u32)(off & 63);
      for (u64 i = w0 + gt; i < w1; i += gn)
        words[i] = (i == w0 && keep) ? (words[i] & lowmask(keep)) : 0ull;
    }
    __syncthreads();
  }
}

// Bits written (compression) or consumed (decompression) per job; a
// variable-rate compression also completes the job's index with its total.
__global__ void k_batch_totals(const BJob* __restrict__ jobs, u32 njobs, const u64* __restrict__ offs,
                               u64* __restrict__ bits)
{
  const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= njobs)
    return;
  const BJob& J = jobs[j];
  u64 b;
  if (J.mode == BM_FIXED)
    b = (u64)J.nblocks * J.z.maxbits;
  else if (!J.dir) {
    b = offs[J.vbase + J.nblocks] - offs[J.vbase];
    if (J.index)
      J.index[J.nblocks] = b;
  }
  else
    b = J.index[J.nblocks];
  bits[j] = b;
}

int launch_batch_prepare(unsigned grid, hipStream_t st, const BJob* jobs, const u32* pjobs, u32 npjobs,
                         const u64* offs)
{
  hipLaunchKernelGGL(k_batch_prepare, dim3(grid), dim3(256), 0, st, jobs, pjobs, npjobs, offs);
  HIPCHK(hipGetLastError());
  return ZFP_HIP_OK;
}

int launch_batch_totals(hipStream_t st, const BJob* jobs, u32 njobs, const u64* offs, u64* bits)
{
  hipLaunchKernelGGL(k_batch_totals, dim3((njobs + 255) / 256), dim3(256), 0, st, jobs, njobs, offs, bits);
  HIPCHK(hipGetLastError());
  return ZFP_HIP_OK;
}

namespace {

struct Shape {
  unsigned wg = 0, wpc = 0, bpt = 0; // workgroup size, waves per CU, blocks per task
};

std::map<std::string, Shape> g_shape_cache;

template <typename T>
int dev_alloc(T** p, size_t count)
{
  *p = nullptr;
  if (!count)
    return ZFP_HIP_OK;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return record_error(ZFP_HIP_ENOMEM, "batch: device allocation failed");
  }
  return ZFP_HIP_OK;
}

} // namespace

struct Batch {
  int dir = 0;
  hipStream_t user = nullptr; // the caller's stream
  hipStream_t st = nullptr;   // private stream the graph runs on
  int ncu = 256;
  std::vector<BJob> jobs;
  std::vector<u32> order;     // job indices grouped by kind
  std::vector<u32> prep;      // jobs whose words are pre-zeroed
  u64 nvar = 0;               // blocks of variable-rate jobs
  BJob* d_jobs = nullptr;
  uint2* d_tasks = nullptr;
  size_t task_cap = 0;
  u32* d_prep = nullptr;
  u64* d_lens = nullptr;
  u64* d_offs = nullptr;
  u64* d_tmp = nullptr;
  u64* d_bits = nullptr;
  u64* h_bits = nullptr;      // pinned
  hipGraphExec_t exec = nullptr;
  Shape shape;
  bool fixed_shape = false;
  std::string sig;

  ~Batch()
  {
    if (exec)
      (void)hipGraphExecDestroy(exec);
    if (st)
      (void)hipStreamDestroy(st);
    for (void* p : {(void*)d_jobs, (void*)d_tasks, (void*)d_prep, (void*)d_lens, (void*)d_offs, (void*)d_tmp,
                    (void*)d_bits})
      if (p)
        (void)hipFree(p);
    if (h_bits)
      (void)hipHostFree(h_bits);
    (void)hipGetLastError();
  }

  // tasks of one shape, per family (0: 1-3D lossy, 1: 4D or reversible):
  // all jobs (ENC/DEC), and the variable-rate jobs (LEN)
  void tasks_for(unsigned bpt, std::vector<uint2>* all, std::vector<uint2>* var) const
  {
    for (int f = 0; f < 2; f++) {
      all[f].clear();
      var[f].clear();
    }
    for (u32 j : order) {
      const u32 fam = jobs[j].wide;
      for (u32 f = 0; f < jobs[j].nblocks; f += bpt) {
        all[fam].push_back(make_uint2(j, f));
        if (jobs[j].mode == BM_VAR)
          var[fam].push_back(make_uint2(j, f));
      }
    }
  }

  // records the passes for shape s into a graph (replacing the current one)
  int capture(const Shape& s)
  {
    std::vector<uint2> all[2], var[2];
    tasks_for(s.bpt, all, var);
    // d_tasks: all (fast, wide), then var (fast, wide)
    std::vector<uint2> list;
    for (auto* v : {&all[0], &all[1], &var[0], &var[1]})
      list.insert(list.end(), v->begin(), v->end());
    const size_t nt = list.size();
    const u32 na0 = (u32)all[0].size(), na1 = (u32)all[1].size();
    const u32 nv0 = (u32)var[0].size(), nv1 = (u32)var[1].size();
    if (nt > task_cap) {
      if (d_tasks)
        (void)hipFree(d_tasks);
      d_tasks = nullptr;
      task_cap = 0;
      int rc = dev_alloc(&d_tasks, nt);
      if (rc)
        return rc;
      task_cap = nt;
    }
    if (nt)
      HIPCHK(hipMemcpy(d_tasks, list.data(), nt * sizeof(uint2), hipMemcpyHostToDevice));
    const unsigned grid = std::max(1u, (unsigned)((u64)ncu * s.wpc * 64 / s.wg));
    const u32 njobs = (u32)jobs.size();

    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = ZFP_HIP_OK;
    if (dir == ZFP_HIP_BATCH_COMPRESS) {
      if (nvar) {
        rc = launch_batch_families<BK_LEN>(grid, s.wg, st, d_jobs, d_tasks + na0 + na1, nv0, nv1, s.bpt, d_lens,
                                           d_offs, 0);
        if (!rc)
          rc = exclusive_scan_tmp(d_lens, d_offs, nvar + 1, d_tmp, st);
      }
      if (!rc && !prep.empty())
        rc = launch_batch_prepare(ncu * 4, st, d_jobs, d_prep, (u32)prep.size(), d_offs);
      if (!rc)
        rc = launch_batch_families<BK_ENC>(grid, s.wg, st, d_jobs, d_tasks, na0, na1, s.bpt, d_lens, d_offs,
                                           batch_stage_stride(jobs.data(), jobs.size(), 0));
    }
    else {
      rc = launch_batch_families<BK_DEC>(grid, s.wg, st, d_jobs, d_tasks, na0, na1, s.bpt, d_lens, d_offs,
                                         batch_stage_stride(jobs.data(), jobs.size(), 1));
    }
    if (!rc)
      rc = launch_batch_totals(st, d_jobs, njobs, d_offs, d_bits);
    if (!rc && hipMemcpyAsync(h_bits, d_bits, njobs * sizeof(u64), hipMemcpyDeviceToHost, st) != hipSuccess)
      rc = record_error(ZFP_HIP_ERUNTIME, "batch: capture of the result copy failed");
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph)
        (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      return rc;
    }
    if (e != hipSuccess)
      return record_hip_error(e, "hipStreamEndCapture");
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess)
      return record_hip_error(e, "hipGraphInstantiate");
    if (exec)
      (void)hipGraphExecDestroy(exec);
    exec = x;
    shape = s;
    return ZFP_HIP_OK;
  }

  int launch_and_wait()
  {
    HIPCHK(hipGraphLaunch(exec, st));
    HIPCHK(hipStreamSynchronize(st));
    return ZFP_HIP_OK;
  }

  // time the candidate shapes and keep the fastest
  int tune()
  {
    static const Shape cand[] = {{256, 8, 64}, {256, 8, 256}, {256, 4, 64}, {128, 8, 64}, {64, 8, 64},
                                 {256, 8, 1024}};
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    float best = 1e30f;
    Shape bs = cand[0];
    int rc = ZFP_HIP_OK;
    for (const Shape& s : cand) {
      rc = capture(s);
      if (rc)
        break;
      rc = launch_and_wait(); // warm-up
      if (rc)
        break;
      float t = 1e30f;
      for (int k = 0; k < 3 && !rc; k++) {
        if (hipEventRecord(e0, st) != hipSuccess || hipGraphLaunch(exec, st) != hipSuccess ||
            hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
          rc = record_error(ZFP_HIP_ERUNTIME, "batch: timing launch failed");
          break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t = std::min(t, ms);
      }
      if (!rc && t < best) {
        best = t;
        bs = s;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
      return rc;
    g_shape_cache[sig] = bs;
    return capture(bs);
  }
};

int job_to_device(const zfp_hip_job& in, int dir, BJob& J, u64& vbase)
{
  memset(&J, 0, sizeof(J));
  Desc d;
  int rc = make_desc(in.type, in.n, in.s, d);
  if (rc)
    return rc;
  // reversible mode: minexp below ZFP_MIN_EXP (zfp_stream_set_reversible,
  // src/zfp.c:750-757)
  const bool rev = in.minexp < -1074;
  if (in.minbits > in.maxbits || !in.maxbits || !in.maxprec || in.maxprec > 64)
    return record_error(ZFP_HIP_EINVAL, "batch: invalid codec parameters");
  // serial zfp reads an all-zero reversible float block as minbits bits
  // although one bit was written (revdecodef.c:24-59), so with minbits > 1
  // such a stream decodes only block after block, never from an index
  if (rev && in.minbits > 1 && (in.type == zfp_type_float || in.type == zfp_type_double))
    return record_error(ZFP_HIP_EINVAL, "batch: reversible float jobs need minbits <= 1");
  if (!in.data || !in.stream || !is_device_ptr(in.data) || !is_device_ptr(in.stream))
    return record_error(ZFP_HIP_EINVAL, "batch: field and stream must be device memory");
  J.g = d.g;
  J.z = Params{in.minbits, in.maxbits, in.maxprec, in.minexp};
  J.data = in.data;
  J.words = in.stream;
  J.offset = in.offset;
  J.index = in.index;
  J.nblocks = d.nblocks;
  J.type = (u32)in.type;
  J.dims = (u32)d.dims;
  J.rev = rev ? 1u : 0u;
  J.wide = rev || d.dims == 4 ? 1u : 0u;
  J.kind = job_kind(J.type, J.dims, rev);
  J.mode = !rev && in.minbits == in.maxbits ? BM_FIXED : BM_VAR;
  J.dir = dir == ZFP_HIP_BATCH_DECOMPRESS ? 1u : 0u;
  if (J.mode == BM_VAR && dir == ZFP_HIP_BATCH_COMPRESS) {
    J.vbase = vbase;
    vbase += d.nblocks;
  }
  if (dir == ZFP_HIP_BATCH_DECOMPRESS) {
    if (J.mode == BM_VAR && !in.index)
      return record_error(ZFP_HIP_ENOINDEX, "batch: variable-rate decompression needs the job's index");
    if (J.mode == BM_VAR && !in.nwords)
      return record_error(ZFP_HIP_EINVAL, "batch: variable-rate decompression needs the stream's word count");
    // the reader never loads past nwords; for fixed rate the stream's extent
    // is known, which bounds a zero capacity too
    const u64 need = (in.offset + (u64)d.nblocks * in.maxbits + 63) / 64;
    J.nwords = in.nwords ? in.nwords : need;
  }
  // fixed rate, word-aligned blocks in stream order: staged through LDS
  bool order = true;
  for (int i = 0; i < d.dims; i++)
    order = order && d.g.lax[i] == (u32)i;
  J.stage = J.mode == BM_FIXED && !J.wide && order && !(in.offset & 63) && !(in.maxbits & 63) &&
                    in.maxbits / 64 <= BATCH_STAGE_WORDS && (!J.dir || J.nwords > 0)
                ? 1u
                : 0u;
  return ZFP_HIP_OK;
}

namespace {

std::string signature(const Batch& b)
{
  // direction, and per kind of either family the job count and log2 of
  // its block total
  constexpr int NK = 12 + 20;
  u64 blocks[NK] = {0};
  unsigned count[NK] = {0};
  bool var = false;
  for (const BJob& J : b.jobs) {
    const u32 k = J.wide ? 12 + J.kind : J.kind;
    blocks[k] += J.nblocks;
    count[k]++;
    var = var || J.mode == BM_VAR;
  }
  std::string s = b.dir ? "D" : "C";
  s += var ? "v" : "f";
  char buf[32];
  for (int k = 0; k < NK; k++) {
    int lg = 0;
    while ((blocks[k] >> lg) > 1)
      lg++;
    snprintf(buf, sizeof(buf), "|%u:%d", count[k] ? (unsigned)(31 - __builtin_clz(count[k])) : 99u, blocks[k] ? lg : -1);
    s += buf;
  }
  return s;
}

int create(const zfp_hip_job* in, size_t njobs, int dir, void* user_stream, Batch** out)
{
  *out = nullptr;
  int rc = device_check();
  if (rc)
    return rc;
  if (!in || !njobs || njobs > 0xffffffffull || (dir != ZFP_HIP_BATCH_COMPRESS && dir != ZFP_HIP_BATCH_DECOMPRESS))
    return record_error(ZFP_HIP_EINVAL, "batch: no jobs or unknown direction");
  Batch* b = new Batch();
  b->dir = dir;
  b->user = (hipStream_t)user_stream;
  b->jobs.resize(njobs);
  u64 vbase = 0;
  for (size_t j = 0; j < njobs && !rc; j++)
    rc = job_to_device(in[j], dir, b->jobs[j], vbase);
  if (rc) {
    delete b;
    return rc;
  }
  b->nvar = dir == ZFP_HIP_BATCH_COMPRESS ? vbase : 0;
  for (u32 j = 0; j < (u32)njobs; j++) {
    b->order.push_back(j);
    const BJob& J = b->jobs[j];
    if (dir == ZFP_HIP_BATCH_COMPRESS && (J.mode == BM_VAR || (J.offset & 63) || (J.z.maxbits & 63)))
      b->prep.push_back(j);
  }
  std::stable_sort(b->order.begin(), b->order.end(), [b](u32 x, u32 y) {
    const BJob &a = b->jobs[x], &c = b->jobs[y];
    return a.wide != c.wide ? a.wide < c.wide : a.kind < c.kind;
  });
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    b->ncu = prop.multiProcessorCount;
  (void)hipGetLastError();
  if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    b->st = nullptr;
    delete b;
    return record_error(ZFP_HIP_ERUNTIME, "batch: stream creation failed");
  }
  rc = dev_alloc(&b->d_jobs, njobs);
  if (!rc)
    rc = dev_alloc(&b->d_bits, njobs);
  if (!rc && !b->prep.empty())
    rc = dev_alloc(&b->d_prep, b->prep.size());
  if (!rc && b->nvar) {
    rc = dev_alloc(&b->d_lens, b->nvar + 1);
    if (!rc)
      rc = dev_alloc(&b->d_offs, b->nvar + 1);
    if (!rc)
      rc = dev_alloc(&b->d_tmp, scan_tmp_words(b->nvar + 1));
  }
  if (!rc && hipHostMalloc((void**)&b->h_bits, njobs * sizeof(u64), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    b->h_bits = nullptr;
    rc = record_error(ZFP_HIP_ENOMEM, "batch: pinned allocation failed");
  }
  if (!rc && hipMemcpy(b->d_jobs, b->jobs.data(), njobs * sizeof(BJob), hipMemcpyHostToDevice) != hipSuccess)
    rc = record_error(ZFP_HIP_ERUNTIME, "batch: job upload failed");
  if (!rc && b->d_prep &&
      hipMemcpy(b->d_prep, b->prep.data(), b->prep.size() * sizeof(u32), hipMemcpyHostToDevice) != hipSuccess)
    rc = record_error(ZFP_HIP_ERUNTIME, "batch: job list upload failed");
  // the scan reads one length past the last block: keep it zero
  if (!rc && b->nvar && hipMemset(b->d_lens + b->nvar, 0, sizeof(u64)) != hipSuccess)
    rc = record_error(ZFP_HIP_ERUNTIME, "batch: initialisation failed");
  if (rc) {
    (void)hipGetLastError();
    delete b;
    return rc;
  }
  b->sig = signature(*b);
  *out = b;
  return ZFP_HIP_OK;
}

int run(Batch* b, uint64_t* bits)
{
  int rc = device_check();
  if (rc)
    return rc;
  // work queued on the caller's stream (e.g. the producer of the fields)
  // completes before the batch reads or writes the buffers
  HIPCHK(hipStreamSynchronize(b->user));
  if (!b->exec) {
    auto it = g_shape_cache.find(b->sig);
    if (b->fixed_shape)
      rc = b->capture(b->shape);
    else if (it != g_shape_cache.end())
      rc = b->capture(it->second);
    else
      rc = b->tune();
    if (rc)
      return rc;
  }
  rc = b->launch_and_wait();
  if (rc)
    return rc;
  if (bits)
    memcpy(bits, b->h_bits, b->jobs.size() * sizeof(u64));
  return ZFP_HIP_OK;
}

} // namespace
} // namespace zfphip

using namespace zfphip;

extern "C" zfp_hip_batch* zfp_hip_batch_create(const zfp_hip_job* jobs, size_t njobs, int direction,
                                               void* hip_stream)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  Batch* b = nullptr;
  if (create(jobs, njobs, direction, hip_stream, &b))
    return nullptr;
  record_error(ZFP_HIP_OK, "");
  return reinterpret_cast<zfp_hip_batch*>(b);
}

extern "C" int zfp_hip_batch_run(zfp_hip_batch* batch, uint64_t* bits)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!batch)
    return record_error(ZFP_HIP_EINVAL, "batch: null batch");
  int rc = run(reinterpret_cast<Batch*>(batch), bits);
  if (!rc)
    record_error(ZFP_HIP_OK, "");
  return rc;
}

extern "C" int zfp_hip_batch_shape(const zfp_hip_batch* batch, unsigned* wg_size, unsigned* waves_per_cu,
                                   unsigned* blocks_per_task)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!batch)
    return record_error(ZFP_HIP_EINVAL, "batch: null batch");
  const Batch* b = reinterpret_cast<const Batch*>(batch);
  if (wg_size)
    *wg_size = b->shape.wg;
  if (waves_per_cu)
    *waves_per_cu = b->shape.wpc;
  if (blocks_per_task)
    *blocks_per_task = b->shape.bpt;
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_batch_set_shape(zfp_hip_batch* batch, unsigned wg_size, unsigned waves_per_cu,
                                       unsigned blocks_per_task)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!batch)
    return record_error(ZFP_HIP_EINVAL, "batch: null batch");
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!wg_size) {
    b->fixed_shape = false;
    return ZFP_HIP_OK;
  }
  if ((wg_size != 64 && wg_size != 128 && wg_size != 256) || !waves_per_cu || waves_per_cu > 32 ||
      !blocks_per_task || blocks_per_task % 64)
    return record_error(ZFP_HIP_EINVAL, "batch: shape needs wg 64/128/256, 1-32 waves per CU, blocks a multiple of 64");
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  b->fixed_shape = true;
  b->shape = Shape{wg_size, waves_per_cu, blocks_per_task};
  return ZFP_HIP_OK;
}

extern "C" void zfp_hip_batch_destroy(zfp_hip_batch* batch)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  delete reinterpret_cast<Batch*>(batch);
}
