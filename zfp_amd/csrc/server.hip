// server.hip -- the serving loop of include/zfp_hip.h: requests are admitted
// in arrival order while their declared footprint fits the HBM budget, every
// admitted request that is not yet running joins the next batch, and up to
// `slots` batches run at once, each on its own HIP stream.  A batch runs the
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
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  u64 id = 0;
  int dir = 0;
  u64 bytes = 0;
  BJob job{};
  hipEvent_t ready = nullptr; // recorded on the caller's stream at submission
  Clock::time_point t_submit, t_done;
  u64 bits = 0;
  int status = -1;            // -1 while queued or running
};

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
    std::vector<uint2> comp_all, comp_var, dec_all;
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
          dec_all.push_back(t);
        else {
          comp_all.push_back(t);
          if (J.mode == BM_VAR)
            comp_var.push_back(t);
        }
      }
      s.h_jobs.p[i] = J;
    }
    const size_t nt = comp_all.size() + comp_var.size() + dec_all.size();
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
    std::copy(comp_all.begin(), comp_all.end(), s.h_tasks.p);
    std::copy(comp_var.begin(), comp_var.end(), s.h_tasks.p + comp_all.size());
    std::copy(dec_all.begin(), dec_all.end(), s.h_tasks.p + comp_all.size() + comp_var.size());
    std::copy(prep.begin(), prep.end(), s.h_prep.p);
    const uint2* t_comp = s.d_tasks.p;
    const uint2* t_var = t_comp + comp_all.size();
    const uint2* t_dec = t_var + comp_var.size();
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
      rc = launch_batch_pass<BK_LEN>(grid, SRV_WG, s.st, s.d_jobs.p, t_var, (u32)comp_var.size(), SRV_BPT,
                                     s.d_lens.p, s.d_offs.p);
      if (!rc)
        rc = exclusive_scan_tmp(s.d_lens.p, s.d_offs.p, vbase + 1, s.d_tmp.p, s.st);
    }
    if (!rc && !prep.empty())
      rc = launch_batch_prepare(ncu * 4, s.st, s.d_jobs.p, s.d_prep.p, (u32)prep.size(), s.d_offs.p);
    if (!rc && !comp_all.empty())
      rc = launch_batch_pass<BK_ENC>(grid, SRV_WG, s.st, s.d_jobs.p, t_comp, (u32)comp_all.size(), SRV_BPT,
                                     s.d_lens.p, s.d_offs.p);
    if (!rc && !dec_all.empty())
      rc = launch_batch_pass<BK_DEC>(grid, SRV_WG, s.st, s.d_jobs.p, t_dec, (u32)dec_all.size(), SRV_BPT,
                                     s.d_lens.p, s.d_offs.p);
    if (!rc)
      rc = launch_batch_totals(s.st, s.d_jobs.p, (u32)n, s.d_offs.p, s.d_bits.p);
    if (rc)
      return rc;
    HIPCHK(hipMemcpyAsync(s.h_bits.p, s.d_bits.p, n * sizeof(u64), hipMemcpyDeviceToHost, s.st));
    HIPCHK(hipEventRecord(s.done, s.st));
    return ZFP_HIP_OK;
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

  void loop()
  {
    std::unique_lock<std::mutex> lk(m);
    for (;;) {
      // completions
      for (Slot& s : slots) {
        if (!s.busy)
          continue;
        const hipError_t q = hipEventQuery(s.done);
        if (q == hipSuccess)
          finish(s, ZFP_HIP_OK);
        else if (q != hipErrorNotReady) {
          (void)hipGetLastError();
          finish(s, ZFP_HIP_ERUNTIME);
        }
      }
      // admission and launches: the queue head waits until it fits
      for (Slot& s : slots) {
        if (s.busy || pending.empty())
          continue;
        std::vector<Request*> batch;
        while (!pending.empty() && batch.size() < max_jobs && used + pending.front()->bytes <= budget) {
          batch.push_back(pending.front());
          used += pending.front()->bytes;
          pending.pop_front();
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
      if (busy)
        cv_work.wait_for(lk, std::chrono::microseconds(20));
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
