// Serving loop over a resident index: continuous batching.
//
// Requests (one query each) wait in a FIFO.  max_inflight worker threads,
// each with its own HIP stream, pinned staging and device buffers, loop:
// take every pending request that fits (max_batch queries, and the worker's
// share of the scratch budget), upload the batch, answer it with
// Index::knn_device on the worker's stream, copy the results back and
// complete the requests.  Scratch is charged by k: a batch runs at the
// smallest power-of-two k class that covers its requests' k, every query of
// it is charged the path's scratch at that class, and the batch's device
// buffers are sized for the queries its share admits at that class.  A request that arrives while
// batches are on the device is admitted by the next worker that frees up,
// so with two workers one batch is uploaded and started while the other is
// still running.  Each request is answered exactly as it would be alone:
// the result of a query does not depend on the batch it ran in.
//
// The reference answers a query file in a loop, one query at a time
// (hercules_file_buffer_manager / query answering in hercules.c); this
// loop has no single counterpart there.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hkg_engine.h"

namespace hkg {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  std::vector<float> q;
  uint32_t k = 0;
  Clock::time_point t0;
  int state = 0;  // 0 pending, 1 on the device, 2 done, 3 failed
  hkg_status rc = HKG_OK;
  std::vector<uint64_t> ids;
  std::vector<float> d2;
  double lat_us = 0.0;
};

}  // namespace

struct Server {
  Index *ix = nullptr;
  hkg_server_opts opt{};
  uint64_t batch_scratch = 0;  // scratch share of one batch
  // k class of a request: the smallest power of two >= k, at most max_k
  uint32_t k_class(uint32_t k) const
  {
    uint32_t c = 1;
    while (c < k) c <<= 1;
    return std::min(c, opt.max_k);
  }
  // scratch one query needs on the path at a k class
  uint64_t per_query(uint32_t kc) const { return ix->scratch_per_query((hkg_path)opt.path, kc); }
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::unordered_map<uint64_t, Request> reqs;
  std::deque<uint64_t> pending;
  uint64_t next_ticket = 1;
  bool stopping = false;
  std::vector<std::thread> workers;
  hkg_status start_rc = HKG_OK;

  void worker(int wi);
};

namespace {

struct Slot {
  hipStream_t st = nullptr;
  float *hQ = nullptr, *dQ = nullptr, *hd2 = nullptr, *dd2 = nullptr;
  uint32_t *hk = nullptr, *dk = nullptr;
  uint64_t *hids = nullptr, *dids = nullptr;
  ~Slot()
  {
    for (void *p : {(void *)hQ, (void *)hd2, (void *)hk, (void *)hids})
      if (p) (void)hipHostFree(p);
    for (void *p : {(void *)dQ, (void *)dd2, (void *)dk, (void *)dids})
      if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
  }
  bool alloc(uint32_t B, uint32_t L, uint32_t K)
  {
    return hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
           hipHostMalloc((void **)&hQ, (size_t)B * L * 4, hipHostMallocDefault) == hipSuccess &&
           hipHostMalloc((void **)&hk, (size_t)B * 4, hipHostMallocDefault) == hipSuccess &&
           hipHostMalloc((void **)&hids, (size_t)B * K * 8, hipHostMallocDefault) == hipSuccess &&
           hipHostMalloc((void **)&hd2, (size_t)B * K * 4, hipHostMallocDefault) == hipSuccess &&
           hipMalloc((void **)&dQ, (size_t)B * L * 4) == hipSuccess &&
           hipMalloc((void **)&dk, (size_t)B * 4) == hipSuccess &&
           hipMalloc((void **)&dids, (size_t)B * K * 8) == hipSuccess &&
           hipMalloc((void **)&dd2, (size_t)B * K * 4) == hipSuccess;
  }
};

}  // namespace

void Server::worker(int wi)
{
  // HKG_SERVER_TRACE: one line per batch on stderr (size, k class, time,
  // overflow re-runs, buffer sets made and dropped)
  const bool trace = getenv("HKG_SERVER_TRACE") != nullptr;
  const uint32_t L = ix->dev.L, B = opt.max_batch, K = opt.max_k;
  if (hipSetDevice(ix->dev.device) != hipSuccess) return;
  Slot s;
  const bool ok = s.alloc(B, L, K);
  {
    std::lock_guard<std::mutex> lock(mu);
    if (!ok) {
      (void)hipGetLastError();
      start_rc = HKG_ENOMEM;
      stopping = true;
      cv_work.notify_all();
      cv_done.notify_all();
      return;
    }
  }
  hkg_query_opts qo;
  hkg_default_query_opts(&qo);
  qo.path = (hkg_path)opt.path;
  qo.batch = B;
  qo.bf16_bound_pass = opt.bf16_bound_pass;
  std::vector<Request *> batch;
  batch.reserve(B);
  for (;;) {
    batch.clear();
    uint32_t kc = 1;  // the batch's k class
    {
      std::unique_lock<std::mutex> lock(mu);
      cv_work.wait(lock, [&] { return stopping || !pending.empty(); });
      if (pending.empty()) break;  // stopping and drained
      // admit what fits: max_batch queries, each charged the scratch of the
      // batch's k class, within this batch's scratch share; on the BF16
      // path in whole query groups when more than one group waits (a
      // partial group leaves filter blocks idle)
      uint32_t bmax = B;
      if (opt.path == HKG_PATH_BF16 && pending.size() < B) {
        const uint32_t gq = bf16_group_queries(ix->dev.L, ix->bf16_qt());
        if (pending.size() > gq) bmax = (uint32_t)(pending.size() / gq * gq);
      }
      while (!pending.empty() && batch.size() < bmax) {
        Request &r = reqs[pending.front()];
        const uint32_t kc2 = std::max(kc, k_class(r.k));
        if (!batch.empty() && (batch.size() + 1) * per_query(kc2) > batch_scratch) break;
        pending.pop_front();
        r.state = 1;
        batch.push_back(&r);
        kc = kc2;
      }
    }
    // requests on the device are not touched by the client (their state
    // is 1), and unordered_map never moves its elements
    const uint32_t n = (uint32_t)batch.size();
    // a batch that fits the share at the largest class runs at it: the sets
    // the server prepared at start then serve every such batch (a set per
    // class would be allocated, and under memory pressure dropped and
    // re-allocated, by the batches of each class).  Only a batch that the
    // share admits at its own class alone runs at that class.
    if ((uint64_t)n * per_query(K) <= batch_scratch) kc = K;
    // buffer sets sized for what the share admits at this class (batches of
    // fewer queries reuse them)
    call_settings().min_buffer_batch =
        (uint32_t)std::max<uint64_t>(n, std::min<uint64_t>(B, batch_scratch / per_query(kc)));
    for (uint32_t i = 0; i < n; ++i) {
      const Request &r = *batch[i];
      memcpy(s.hQ + (size_t)i * L, r.q.data(), (size_t)L * 4);
      s.hk[i] = r.k;
    }
    hkg_status rc = HKG_OK;
    if (hipMemcpyAsync(s.dQ, s.hQ, (size_t)n * L * 4, hipMemcpyHostToDevice, s.st) != hipSuccess ||
        hipMemcpyAsync(s.dk, s.hk, (size_t)n * 4, hipMemcpyHostToDevice, s.st) != hipSuccess) {
      rc = HKG_EDEVICE;
    }
    // results at the class's stride (kc ranks per query)
    const Clock::time_point t_run = Clock::now();
    uint64_t sets0 = 0, evict0 = 0;
    hkg_query_stats stt{};
    if (trace) {
      std::lock_guard<std::mutex> lock(ix->mu_bufs);
      sets0 = ix->sets_allocated;
      evict0 = ix->sets_evicted;
    }
    if (rc == HKG_OK) rc = ix->knn_device(s.dQ, n, s.dk, kc, &qo, s.dids, s.dd2, s.st, trace ? &stt : nullptr);
    if (rc == HKG_OK &&
        (hipMemcpyAsync(s.hids, s.dids, (size_t)n * kc * 8, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
         hipMemcpyAsync(s.hd2, s.dd2, (size_t)n * kc * 4, hipMemcpyDeviceToHost, s.st) != hipSuccess ||
         hipStreamSynchronize(s.st) != hipSuccess)) {
      rc = HKG_EDEVICE;
    }
    const Clock::time_point now = Clock::now();
    if (trace) {
      uint64_t sets1, evict1;
      {
        std::lock_guard<std::mutex> lock(ix->mu_bufs);
        sets1 = ix->sets_allocated;
        evict1 = ix->sets_evicted;
      }
      fprintf(stderr, "hkg server: worker %d batch %u k class %u: %.2f ms (device %.2f ms), %u overflow re-runs, %llu buffer sets made, %llu dropped\n",
              wi, n, kc, std::chrono::duration<double, std::milli>(now - t_run).count(), stt.device_ms,
              stt.overflow_retries, (unsigned long long)(sets1 - sets0), (unsigned long long)(evict1 - evict0));
    }
    std::lock_guard<std::mutex> lock(mu);
    for (uint32_t i = 0; i < n; ++i) {
      Request &r = *batch[i];
      r.rc = rc;
      if (rc == HKG_OK) {
        r.ids.assign(s.hids + (size_t)i * kc, s.hids + (size_t)i * kc + r.k);
        r.d2.assign(s.hd2 + (size_t)i * kc, s.hd2 + (size_t)i * kc + r.k);
      }
      r.q.clear();
      r.q.shrink_to_fit();
      r.lat_us = std::chrono::duration<double, std::micro>(now - r.t0).count();
      r.state = rc == HKG_OK ? 2 : 3;
    }
    cv_done.notify_all();
  }
}

}  // namespace hkg

using hkg::Server;

extern "C" {

void hkg_default_server_opts(hkg_server_opts *so)
{
  if (!so) return;
  memset(so, 0, sizeof(*so));
  so->max_batch = 1024;
  so->scratch_bytes = 16ull << 30;
  so->queue_capacity = 65536;
  so->path = HKG_PATH_BATCHED;
  so->max_inflight = 2;
  so->max_k = hkg::kMaxK;
}

hkg_status hkg_server_start(hkg_index *index, const hkg_server_opts *so, hkg_server **out)
{
  if (!index || !out) return HKG_EINVAL;
  *out = nullptr;
  hkg_server_opts o;
  hkg_default_server_opts(&o);
  if (so) o = *so;
  if (o.max_batch == 0) o.max_batch = 1024;
  if (o.scratch_bytes == 0) o.scratch_bytes = 16ull << 30;
  if (o.queue_capacity == 0) o.queue_capacity = 65536;
  if (o.max_inflight == 0) o.max_inflight = 2;
  if (o.max_k == 0) o.max_k = hkg::kMaxK;
  if (o.max_batch > 16384 || o.max_inflight > 8 || o.max_k > (uint32_t)hkg::kMaxK ||
      (o.path != HKG_PATH_AUTO && o.path != HKG_PATH_BATCHED && o.path != HKG_PATH_INDEX &&
       o.path != HKG_PATH_SCAN && o.path != HKG_PATH_BF16)) {
    hkg::set_error("server: max_batch <= 16384, max_inflight <= 8, max_k <= 256 and a valid path");
    return HKG_EINVAL;
  }
  if (o.path == HKG_PATH_AUTO) o.path = HKG_PATH_BATCHED;
  std::unique_ptr<Server> s(new Server());
  s->ix = reinterpret_cast<hkg::Index *>(index);
  s->opt = o;
  s->batch_scratch = o.scratch_bytes / o.max_inflight;
  if (s->per_query(o.max_k) > s->batch_scratch) {
    hkg::set_error("server: the scratch budget does not hold one query per batch");
    return HKG_EINVAL;
  }
  if (o.path == HKG_PATH_BF16) {
    hkg_status rc = s->ix->make_bf16();
    if (rc != HKG_OK) return rc;
  }
  {
    // the workers' buffer sets for a full batch at the largest k class,
    // allocated now rather than by the first requests (a few GB each)
    const uint32_t bq = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(o.max_batch, s->batch_scratch / s->per_query(o.max_k)));
    hkg_status rc = s->ix->prepare_buffers((hkg_path)o.path, bq, o.max_k, o.max_inflight);
    if (rc != HKG_OK) return rc;
  }
  for (uint32_t i = 0; i < o.max_inflight; ++i) s->workers.emplace_back(&Server::worker, s.get(), (int)i);
  *out = reinterpret_cast<hkg_server *>(s.release());
  return HKG_OK;
}

hkg_status hkg_submit(hkg_server *sv, const float *q, uint32_t k, uint64_t *ticket)
{
  Server *s = reinterpret_cast<Server *>(sv);
  if (!s || !q || !ticket) return HKG_EINVAL;
  if (k == 0 || k > s->opt.max_k) {
    hkg::set_error("submit: k must be in [1, max_k]");
    return HKG_EINVAL;
  }
  const uint32_t L = s->ix->dev.L;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->stopping) {
    hkg::set_error(s->start_rc != HKG_OK ? "server failed to start its workers" : "server is stopping");
    return s->start_rc != HKG_OK ? s->start_rc : HKG_EINVAL;
  }
  if (s->pending.size() >= s->opt.queue_capacity) {
    hkg::set_error("submit: queue full");
    return HKG_EBUSY;
  }
  const uint64_t t = s->next_ticket++;
  hkg::Request &r = s->reqs[t];
  r.q.assign(q, q + L);
  r.k = k;
  r.t0 = hkg::Clock::now();
  s->pending.push_back(t);
  *ticket = t;
  s->cv_work.notify_one();
  return HKG_OK;
}

// copy a finished request out and forget it
static hkg_status take(Server *s, uint64_t ticket, uint64_t *ids, float *d2, double *lat, bool block)
{
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) {
    hkg::set_error("unknown ticket");
    return HKG_EINVAL;
  }
  if (block)
    s->cv_done.wait(lock, [&] { return it->second.state >= 2 || (s->stopping && s->start_rc != HKG_OK); });
  hkg::Request &r = it->second;
  if (r.state < 2) return s->start_rc != HKG_OK ? s->start_rc : HKG_ENOTREADY;
  const hkg_status rc = r.rc;
  if (rc == HKG_OK) {
    if (ids) memcpy(ids, r.ids.data(), r.ids.size() * 8);
    if (d2) memcpy(d2, r.d2.data(), r.d2.size() * 4);
  }
  if (lat) *lat = r.lat_us;
  s->reqs.erase(it);
  return rc;
}

hkg_status hkg_poll(hkg_server *sv, uint64_t ticket, uint64_t *ids, float *d2, double *latency_us)
{
  Server *s = reinterpret_cast<Server *>(sv);
  if (!s) return HKG_EINVAL;
  return take(s, ticket, ids, d2, latency_us, false);
}

hkg_status hkg_wait(hkg_server *sv, uint64_t ticket, uint64_t *ids, float *d2, double *latency_us)
{
  Server *s = reinterpret_cast<Server *>(sv);
  if (!s) return HKG_EINVAL;
  return take(s, ticket, ids, d2, latency_us, true);
}

void hkg_server_stop(hkg_server *sv)
{
  Server *s = reinterpret_cast<Server *>(sv);
  if (!s) return;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stopping = true;
  }
  s->cv_work.notify_all();
  for (std::thread &t : s->workers) t.join();  // pending requests are answered first
  delete s;
}

}  // extern "C"
