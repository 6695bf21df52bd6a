// server.hip -- the serving loop of include/zfp_hip.h: requests are admitted
// in arrival order while their declared footprint fits the HBM budget, every
// admitted request that is not yet running joins the next batch, and up to
// `slots` batches run at once, each on its own HIP stream.  Requests that
// touch the same memory keep their submission order: one that reads or
// writes what an earlier one writes (or writes what it reads) waits for it,
// except a decompression behind a compression in the same batch, which a
// batch runs in that order anyway.  A batch runs the
// persistent kernels of the batched codec (batch_kernels.h) with a launch
// shape fixed at creation.  A scheduler thread owns the streams; clients
// only queue requests and wait for them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "batch_kernels.h"

namespace zfphip {

struct Server;

namespace {

using Clock = std::chrono::steady_clock;

// byte range [lo, hi) of device memory
struct Span {
  uintptr_t lo = 0, hi = 0;
  bool meets(const Span& o) const { return lo < o.hi && o.lo < hi; }
};

struct Request {
  u64 id = 0;
  int dir = 0;
  u64 bytes = 0;
  BJob job{};
  hipEvent_t ready = nullptr; // recorded on the caller's stream at submission
  Clock::time_point t_submit, t_done;
  u64 bits = 0;
  int status = -1;            // -1 while queued or running
  Span rd[2], wr[2];          // memory the request reads and writes
  u64 cap = 0;                // the job's stream capacity in words (0: not given)
};

// The field's bytes: element 0 at data, any strides (batch_kernels.h BJob)
Span field_span(const BJob& J)
{
  const i64 es = J.type == 1 || J.type == 3 ? 4 : 8;
  i64 lo = 0, hi = 0;
  for (int d = 0; d < 4; d++) {
    const i64 e = (i64)(J.g.n[d] ? J.g.n[d] - 1 : 0) * J.g.s[d];
    (e < 0 ? lo : hi) += e;
  }
  const uintptr_t b = (uintptr_t)J.data;
  return Span{b + (uintptr_t)(lo * es), b + (uintptr_t)((hi + 1) * es)};
}

// What a request reads and writes: compression reads the field and writes
// the stream words its blocks can reach (and the index); decompression
// reads the stream's readable words (and the index) and writes the field.
void request_spans(Request& r)
{
  const BJob& J = r.job;
  const u64 size = 1ull << (2 * J.dims), P = J.type == 1 || J.type == 3 ? 32 : 64;
  // block header bits by type, lossy and reversible (int32, int64, float,
  // double)
  static const u64 hdr[2][4] = {{0, 0, 9, 12}, {5, 6, 15, 19}};
  u64 words;
  if (r.dir == ZFP_HIP_BATCH_DECOMPRESS)
    words = J.nwords;
  else {
    // longest block: header, group tests and maxprec planes, within the
    // budget (zfp_stream_maximum_size, src/zfp.c:711-742); no further than
    // the capacity the caller gave
    const u64 most = hdr[J.rev][J.type - 1] + size - 1 + size * std::min<u64>(J.z.maxprec, P);
    const u64 per = std::max<u64>(J.z.minbits, std::min<u64>(J.z.maxbits, most));
    words = (J.offset + (u64)J.nblocks * per + 63) / 64;
    if (r.cap)
      words = std::min<u64>(words, r.cap);
  }
  const uintptr_t w = (uintptr_t)J.words;
  const Span stream{w, w + words * 8};
  const Span index = J.index ? Span{(uintptr_t)J.index, (uintptr_t)(J.index + J.nblocks + 1)} : Span{};
  const Span field = field_span(J);
  if (r.dir == ZFP_HIP_BATCH_DECOMPRESS) {
    r.rd[0] = stream;
    r.rd[1] = index;
    r.wr[0] = field;
  }
  else {
    r.rd[0] = field;
    r.wr[0] = stream;
    r.wr[1] = index;
  }
}

bool meets(const Span* a, const Span* b)
{
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      if (a[i].hi > a[i].lo && b[j].hi > b[j].lo && a[i].meets(b[j]))
        return true;
  return false;
}

// whether request b, submitted after a, has to wait for a: one writes what
// the other reads or writes
bool depends(const Request& a, const Request& b)
{
  return meets(a.wr, b.rd) || meets(a.rd, b.wr) || meets(a.wr, b.wr);
}

// whether b may still run in the same batch as a (a earlier): a batch runs
// every compression before any decompression, so a decompression that
// depends only on earlier compressions keeps their order
bool same_batch_ok(const Request& a, const Request& b)
{
  return !depends(a, b) || (a.dir == ZFP_HIP_BATCH_COMPRESS && b.dir == ZFP_HIP_BATCH_DECOMPRESS);
}

// grow-only device and pinned host buffers
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n)
  {
    if (n <= cap)
      return ZFP_HIP_OK;
    if (p)
      (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(n, (size_t)64);
    if (hipMalloc((void**)&p, want * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return ZFP_HIP_ENOMEM;
    }
    cap = want;
    return ZFP_HIP_OK;
  }
  void release()
  {
    if (p)
      (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

template <typename T>
struct PinBuf {
  T* p = nullptr;
  size_t cap = 0;
  int ensure(size_t n)
  {
    if (n <= cap)
      return ZFP_HIP_OK;
    if (p)
      (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(n, (size_t)64);
    if (hipHostMalloc((void**)&p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return ZFP_HIP_ENOMEM;
    }
    cap = want;
    return ZFP_HIP_OK;
  }
  void release()
  {
    if (p)
      (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Slot {
  hipStream_t st = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  Server* server = nullptr;
  std::vector<Request*> reqs;
  PinBuf<BJob> h_jobs;
  DevBuf<BJob> d_jobs;
  PinBuf<uint2> h_tasks;
  DevBuf<uint2> d_tasks;
  PinBuf<u32> h_prep;
  DevBuf<u32> d_prep;
  DevBuf<u64> d_lens, d_offs, d_tmp, d_bits;
  PinBuf<u64> h_bits;
};

// launch shape of the serving batches: 128-lane workgroups, 8 waves per CU,
// 64 blocks per task
constexpr unsigned SRV_WG = 128, SRV_WPC = 8, SRV_BPT = 64;

} // namespace

struct Server {
  u64 budget = 0, used = 0;
  unsigned max_jobs = 1024;
  int ncu = 256;
  std::vector<Slot> slots;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> pending;
  std::unordered_map<u64, Request*> reqs;
  u64 next_id = 1;
  bool stop = false;
  std::thread worker;
  u64 n_done = 0, n_batches = 0, max_batch = 0, max_inflight = 0;

  // Builds and queues one batch on slot s (scheduler thread, lock not held).
  int launch(Slot& s, const std::vector<Request*>& batch)
  {
    const size_t n = batch.size();
    // task lists per family (0: 1-3D lossy, 1: 4D or reversible)
    std::vector<uint2> comp_all[2], comp_var[2], dec_all[2];
    std::vector<u32> prep;
    u64 vbase = 0;
    int rc = s.h_jobs.ensure(n);
    if (!rc)
      rc = s.d_jobs.ensure(n);
    if (!rc)
      rc = s.h_bits.ensure(n);
    if (!rc)
      rc = s.d_bits.ensure(n);
    if (rc)
      return rc;
    for (size_t i = 0; i < n; i++) {
      BJob J = batch[i]->job;
      const bool comp = batch[i]->dir == ZFP_HIP_BATCH_COMPRESS;
      if (comp && J.mode == BM_VAR) {
        J.vbase = vbase;
        vbase += J.nblocks;
      }
      if (comp && (J.mode == BM_VAR || (J.offset & 63) || (J.z.maxbits & 63)))
        prep.push_back((u32)i);
      for (u32 f = 0; f < J.nblocks; f += SRV_BPT) {
        const uint2 t = make_uint2((u32)i, f);
        if (!comp)
          dec_all[J.wide].push_back(t);
        else {
          comp_all[J.wide].push_back(t);
          if (J.mode == BM_VAR)
            comp_var[J.wide].push_back(t);
        }
      }
      s.h_jobs.p[i] = J;
    }
    // d_tasks: comp_all, comp_var, dec_all, each fast then wide
    std::vector<uint2> list;
    for (auto* v : {&comp_all[0], &comp_all[1], &comp_var[0], &comp_var[1], &dec_all[0], &dec_all[1]})
      list.insert(list.end(), v->begin(), v->end());
    const size_t nt = list.size();
    const u32 nc0 = (u32)comp_all[0].size(), nc1 = (u32)comp_all[1].size();
    const u32 nv0 = (u32)comp_var[0].size(), nv1 = (u32)comp_var[1].size();
    const u32 nd0 = (u32)dec_all[0].size(), nd1 = (u32)dec_all[1].size();
    rc = s.h_tasks.ensure(nt);
    if (!rc)
      rc = s.d_tasks.ensure(nt);
    if (!rc && !prep.empty()) {
      rc = s.h_prep.ensure(prep.size());
      if (!rc)
        rc = s.d_prep.ensure(prep.size());
    }
    if (!rc && vbase) {
      rc = s.d_lens.ensure(vbase + 1);
      if (!rc)
        rc = s.d_offs.ensure(vbase + 1);
      if (!rc)
        rc = s.d_tmp.ensure(scan_tmp_words(vbase + 1));
    }
    if (rc)
      return rc;
    std::copy(list.begin(), list.end(), s.h_tasks.p);
    std::copy(prep.begin(), prep.end(), s.h_prep.p);
    const uint2* t_comp = s.d_tasks.p;
    const uint2* t_var = t_comp + nc0 + nc1;
    const uint2* t_dec = t_var + nv0 + nv1;
    const unsigned grid = std::max(1u, (unsigned)((u64)ncu * SRV_WPC * 64 / SRV_WG));

    for (Request* r : batch)
      if (r->ready)
        HIPCHK(hipStreamWaitEvent(s.st, r->ready, 0));
    HIPCHK(hipMemcpyAsync(s.d_jobs.p, s.h_jobs.p, n * sizeof(BJob), hipMemcpyHostToDevice, s.st));
    if (nt)
      HIPCHK(hipMemcpyAsync(s.d_tasks.p, s.h_tasks.p, nt * sizeof(uint2), hipMemcpyHostToDevice, s.st));
    if (!prep.empty())
      HIPCHK(hipMemcpyAsync(s.d_prep.p, s.h_prep.p, prep.size() * sizeof(u32), hipMemcpyHostToDevice, s.st));
    if (vbase) {
      HIPCHK(hipMemsetAsync(s.d_lens.p + vbase, 0, sizeof(u64), s.st));
      rc = launch_batch_families<BK_LEN>(grid, SRV_WG, s.st, s.d_jobs.p, t_var, nv0, nv1, SRV_BPT, s.d_lens.p,
                                         s.d_offs.p, 0);
      if (!rc)
        rc = exclusive_scan_tmp(s.d_lens.p, s.d_offs.p, vbase + 1, s.d_tmp.p, s.st);
    }
    if (!rc && !prep.empty())
      rc = launch_batch_prepare(ncu * 4, s.st, s.d_jobs.p, s.d_prep.p, (u32)prep.size(), s.d_offs.p);
    if (!rc)
      rc = launch_batch_families<BK_ENC>(grid, SRV_WG, s.st, s.d_jobs.p, t_comp, nc0, nc1, SRV_BPT, s.d_lens.p,
                                         s.d_offs.p, batch_stage_stride(s.h_jobs.p, n, 0));
    if (!rc)
      rc = launch_batch_families<BK_DEC>(grid, SRV_WG, s.st, s.d_jobs.p, t_dec, nd0, nd1, SRV_BPT, s.d_lens.p,
                                         s.d_offs.p, batch_stage_stride(s.h_jobs.p, n, 1));
    if (!rc)
      rc = launch_batch_totals(s.st, s.d_jobs.p, (u32)n, s.d_offs.p, s.d_bits.p);
    if (rc)
      return rc;
    HIPCHK(hipMemcpyAsync(s.h_bits.p, s.d_bits.p, n * sizeof(u64), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipEventRecord(s.done, s.st));
    HIPCHK(hipLaunchHostFunc(s.st, &Server::batch_done, &s));
    return ZFP_HIP_OK;
  }

  // runs on a runtime thread once slot s's batch (and the copy of its bit
  // counts) is complete: wakes the scheduler (taking the lock first, so
  // that the wake cannot fall between its check and its wait)
  static void batch_done(void* arg)
  {
    Server* sv = static_cast<Slot*>(arg)->server;
    {
      std::lock_guard<std::mutex> lk(sv->m);
    }
    sv->cv_work.notify_all();
  }

  // completes the requests of slot s (lock held)
  void finish(Slot& s, int status)
  {
    const auto now = Clock::now();
    for (size_t i = 0; i < s.reqs.size(); i++) {
      Request* r = s.reqs[i];
      r->bits = status ? 0 : s.h_bits.p[i];
      r->status = status;
      r->t_done = now;
      used -= r->bytes;
      n_done++;
    }
    s.reqs.clear();
    s.busy = false;
    cv_done.notify_all();
  }

  // whether r may join `batch` now: no conflict with a request in flight,
  // with an earlier request left waiting, or (but for compress-then-
  // decompress) with one already in the batch
  bool admissible(const Request& r, const std::vector<Request*>& batch, const std::vector<Request*>& held) const
  {
    for (const Slot& t : slots)
      if (t.busy)
        for (const Request* a : t.reqs)
          if (depends(*a, r))
            return false;
    for (const Request* a : held)
      if (depends(*a, r))
        return false;
    for (const Request* a : batch)
      if (!same_batch_ok(*a, r))
        return false;
    return true;
  }

  void loop()
  {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      // completions (the streams' host functions wake this loop; the event
      // tells, and also catches a batch whose stream failed)
      for (Slot& s : slots) {
        if (!s.busy)
          continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipSuccess)
          finish(s, ZFP_HIP_OK);
        else if (q != hipSuccess && q != hipErrorNotReady) {
          (void)hipGetLastError();
          finish(s, ZFP_HIP_ERUNTIME);
        }
      }
      // admission and launches, in arrival order: a request that does not
      // fit the budget stops the admission (it waits at the head); one that
      // touches memory of a request in flight or of an earlier waiting one
      // waits, and later requests pass it
      for (Slot& s : slots) {
        if (s.busy || pending.empty())
          continue;
        std::vector<Request*> batch, held;
        for (auto it = pending.begin(); it != pending.end() && batch.size() < max_jobs;) {
          Request* r = *it;
          if (used + r->bytes > budget)
            break;
          if (!admissible(*r, batch, held)) {
            held.push_back(r);
            ++it;
            continue;
          }
          batch.push_back(r);
          used += r->bytes;
          it = pending.erase(it);
        }
        if (batch.empty())
          break;
        s.reqs = batch;
        s.busy = true;
        lk.unlock();
        const int rc = launch(s, batch);
        lk.lock();
        if (rc) {
          (void)hipGetLastError();
          finish(s, rc);
          continue;
        }
        n_batches++;
        max_batch = std::max<u64>(max_batch, batch.size());
        u64 inflight = 0;
        for (const Slot& t : slots)
          inflight += t.busy ? 1 : 0;
        max_inflight = std::max(max_inflight, inflight);
      }
      bool busy = false;
      for (const Slot& s : slots)
        busy = busy || s.busy;
      if (stop && pending.empty() && !busy)
        break;
      // woken by submissions and by batch ends; the timed wait bounds the
      // lag of a host function that runs late (or never, on a failed stream)
      if (busy)
        cv_work.wait_for(lk, std::chrono::microseconds(100));
      else
        cv_work.wait(lk, [this] { return stop || !pending.empty(); });
    }
  }

  ~Server()
  {
    {
      std::lock_guard<std::mutex> lk(m);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable())
      worker.join();
    for (Slot& s : slots) {
      if (s.st)
        (void)hipStreamSynchronize(s.st);
      s.h_jobs.release();
      s.d_jobs.release();
      s.h_tasks.release();
      s.d_tasks.release();
      s.h_prep.release();
      s.d_prep.release();
      s.d_lens.release();
      s.d_offs.release();
      s.d_tmp.release();
      s.d_bits.release();
      s.h_bits.release();
      if (s.done)
        (void)hipEventDestroy(s.done);
      if (s.st)
        (void)hipStreamDestroy(s.st);
    }
    for (auto& kv : reqs) {
      if (kv.second->ready)
        (void)hipEventDestroy(kv.second->ready);
      delete kv.second;
    }
    (void)hipGetLastError();
  }
};

} // namespace zfphip

using namespace zfphip;

extern "C" zfp_hip_server* zfp_hip_server_create(uint64_t hbm_budget, unsigned max_batch_jobs, unsigned slots)
{
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (device_check())
      return nullptr;
  }
  if (!hbm_budget) {
    record_error(ZFP_HIP_EINVAL, "server: zero HBM budget");
    return nullptr;
  }
  Server* sv = new Server();
  sv->budget = hbm_budget;
  sv->max_jobs = max_batch_jobs ? max_batch_jobs : 1024;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    sv->ncu = prop.multiProcessorCount;
  (void)hipGetLastError();
  sv->slots.resize(slots ? slots : 2);
  for (Slot& s : sv->slots) {
    s.server = sv;
    if (hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      delete sv;
      record_error(ZFP_HIP_ERUNTIME, "server: stream or event creation failed");
      return nullptr;
    }
  }
  sv->worker = std::thread([sv] { sv->loop(); });
  record_error(ZFP_HIP_OK, "");
  return reinterpret_cast<zfp_hip_server*>(sv);
}

extern "C" int zfp_hip_server_submit(zfp_hip_server* server, const zfp_hip_job* job, int direction,
                                     uint64_t bytes, void* hip_stream, uint64_t* id)
{
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv || !job || !id || (direction != ZFP_HIP_BATCH_COMPRESS && direction != ZFP_HIP_BATCH_DECOMPRESS))
    return record_error(ZFP_HIP_EINVAL, "server: invalid request");
  Request* r = new Request();
  u64 vb = 0;
  int rc;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    rc = job_to_device(*job, direction, r->job, vb);
  }
  if (!rc && bytes > sv->budget)
    rc = record_error(ZFP_HIP_EINVAL, "server: request larger than the HBM budget");
  if (!rc && (hipEventCreateWithFlags(&r->ready, hipEventDisableTiming) != hipSuccess ||
              hipEventRecord(r->ready, (hipStream_t)hip_stream) != hipSuccess)) {
    (void)hipGetLastError();
    rc = record_error(ZFP_HIP_ERUNTIME, "server: cannot order the request after its inputs");
  }
  if (rc) {
    if (r->ready)
      (void)hipEventDestroy(r->ready);
    delete r;
    return rc;
  }
  r->dir = direction;
  r->bytes = bytes;
  r->cap = job->nwords;
  request_spans(*r);
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(sv->m);
    r->id = sv->next_id++;
    sv->reqs[r->id] = r;
    sv->pending.push_back(r);
  }
  sv->cv_work.notify_one();
  *id = r->id;
  return ZFP_HIP_OK;
}

extern "C" int zfp_hip_server_wait(zfp_hip_server* server, uint64_t id, uint64_t* bits, double* latency_ms)
{
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv)
    return record_error(ZFP_HIP_EINVAL, "server: null server");
  std::unique_lock<std::mutex> lk(sv->m);
  auto it = sv->reqs.find(id);
  if (it == sv->reqs.end())
    return record_error(ZFP_HIP_EINVAL, "server: unknown request id");
  Request* r = it->second;
  sv->cv_done.wait(lk, [r] { return r->status >= 0; });
  if (bits)
    *bits = r->bits;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  const int status = r->status;
  sv->reqs.erase(it);
  lk.unlock();
  if (r->ready)
    (void)hipEventDestroy(r->ready);
  delete r;
  return status;
}

extern "C" int zfp_hip_server_stats(zfp_hip_server* server, uint64_t* done, uint64_t* batches, uint64_t* max_batch,
                                    uint64_t* max_inflight)
{
  Server* sv = reinterpret_cast<Server*>(server);
  if (!sv)
    return record_error(ZFP_HIP_EINVAL, "server: null server");
  std::lock_guard<std::mutex> lk(sv->m);
  if (done)
    *done = sv->n_done;
  if (batches)
    *batches = sv->n_batches;
  if (max_batch)
    *max_batch = sv->max_batch;
  if (max_inflight)
    *max_inflight = sv->max_inflight;
  return ZFP_HIP_OK;
}

extern "C" void zfp_hip_server_destroy(zfp_hip_server* server)
{
  delete reinterpret_cast<Server*>(server);
}
