// rz_server.hip -- the serving loop: admission by a device-memory budget,
// continuous batching over several streams, per-request latency.
//
// Included by rz_runtime.hip (it drives run_batch on the server's contexts).
//
// n_streams worker threads each own one context (one HIP stream, default
// priority).  Submitted requests wait in a FIFO queue.  An idle worker admits
// queued requests in arrival order, up to max_batch of them, while the
// budget holds them (a request larger than the whole budget is failed), runs
// them as one batch on its context and waits for it.  Batches of different
// workers overlap on the device, and one batch's host round trips (the eager
// pipeline reads a few totals back) do not hold up the others, so requests
// arriving while others run are admitted as soon as a worker and memory free
// up.  Outputs stay on the device until the caller releases them; released
// outputs are kept in a cache of size classes for later requests.
//
// Memory: the budget holds the outputs (admitted and cached) plus every
// worker context's workspace as measured after its last batch, plus, while a
// batch runs, the growth its workspace estimate may need beyond that.
// Admission only does this bookkeeping under the server's lock; the worker
// allocates fresh outputs and frees evicted cache entries afterwards, with
// stream-ordered allocation on its own stream (no device-wide
// synchronisation, and no worker waits on another's allocation).
#pragma once

#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  uint64_t ticket = 0;
  rz_job job{};
  size_t out_bytes = 0, ws_bytes = 0;
  int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
  bool admitted = false;  // its output bytes are held in the budget until released
  rz_status status = RZ_OK;
  std::string error;
  Clock::time_point t_submit, t_start, t_done;
};

// Workspace a request may need (run_batch's buffers): per vertex its edge or
// segment record and band copies (64 B); per item its box, statistics,
// descriptor, scan entries and, per tile its box covers, a bin pair with its
// sort copies, gathered descriptor and heavy-tile masks (4 x 2 KiB at most);
// per output pixel at most 2 B of wide-pre-pass masks and the tile tables.
size_t workspace_estimate(const rz_job& j) {
  const rz_device_geometry* g = j.geometry;
  const size_t nl = j.layer_dev ? j.n_layers : 1u;
  const size_t px = (size_t)j.info.nrows * (size_t)j.info.ncols * nl;
  const size_t tw = tile_w((int)rz_dtype_size(j.dtype));
  const size_t tiles = ((size_t)j.info.ncols + tw - 1) / tw * (((size_t)j.info.nrows + TILE_H - 1) / TILE_H) * nl;
  // an item's box meets about 4 tiles (its tile rectangle, most items being
  // smaller than a tile); pairs beyond that are covered by the per-pixel term
  const size_t pairs = std::min<size_t>((size_t)g->I * 4, tiles * 64);
  return (size_t)g->V * 64 + (size_t)g->I * 160 + pairs * (16 + 64 + 8 + 8 * 1024) + 2 * px + tiles * 40 +
         (1u << 20);
}

}  // namespace

struct rz_server {
  int device = 0;
  uint64_t budget = 0;
  int max_batch = 64;
  std::vector<rz_context*> ctx;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  std::unordered_multimap<size_t, void*> free_bufs;  // released outputs, by size
  uint64_t next_ticket = 1;
  uint64_t out_used = 0;        // outputs of admitted, unreleased requests
  uint64_t cached = 0;          // bytes in free_bufs
  std::vector<uint64_t> ws;     // each context's workspace, measured after its last batch
  std::vector<uint64_t> grow;   // each running batch's reserved workspace growth
  uint64_t n_batches = 0, n_done = 0, peak_used = 0, running = 0;
  uint64_t phase_sum[3] = {0, 0, 0}, phase_max[3] = {0, 0, 0};  // alloc, run, wait (us)

  uint64_t used() const {
    uint64_t u = out_used + cached;
    for (size_t k = 0; k < ws.size(); ++k) u += ws[k] + grow[k];
    return u;
  }
  bool stop = false;
  std::vector<std::thread> workers;

  // Output buffers come in size classes (four per power of two) so released
  // outputs are reused by later requests of similar size: fresh allocations
  // and the eviction of cached ones stay rare.
  static size_t size_class_bytes(size_t bytes) {
    size_t c = 1u << 20;
    while (c < bytes) c <<= 1;  // next power of two >= 1 MiB
    const size_t q = c / 8;     // quarters of the octave [c / 2, c)
    return c <= (1u << 20) ? c : ((bytes + q - 1) / q) * q;
  }

  // What a worker must do outside the lock before its batch runs.
  struct Admission {
    std::vector<uint64_t> tickets;
    std::vector<uint64_t> fresh;   // tickets whose output is allocated fresh
    std::vector<void*> evict;      // cached outputs to free
  };

  // Admits queued requests in arrival order while the budget holds them
  // (lock held; bookkeeping only).
  Admission admit(int slot) {
    Admission a;
    uint64_t est = 0;
    while (!queue.empty() && (int)a.tickets.size() < max_batch) {
      Request& r = reqs[queue.front()];
      const uint64_t est_r = est + r.ws_bytes;
      const uint64_t g = est_r > ws[slot] ? est_r - ws[slot] : 0;  // this slot's growth with r
      if (r.out_bytes + r.ws_bytes > budget) {  // can never fit
        r.state = 3;
        r.status = RZ_CAPACITY_ERROR;
        r.error = "request exceeds the server's memory budget";
        r.t_done = Clock::now();
        queue.pop_front();
        cv_done.notify_all();
        continue;
      }
      // a cached output of r's size class is reused; otherwise cache
      // entries are evicted while that makes r fit
      auto hit = free_bufs.find(r.out_bytes);
      const uint64_t fresh_bytes = hit == free_bufs.end() ? r.out_bytes : 0;
      const uint64_t base = used() - grow[slot];
      while (fresh_bytes && base + fresh_bytes + g > budget && !free_bufs.empty()) {
        auto victim = free_bufs.begin();
        cached -= victim->first;
        a.evict.push_back(victim->second);
        free_bufs.erase(victim);
      }
      if (used() - grow[slot] + fresh_bytes + g > budget) break;
      if (hit != free_bufs.end()) {
        r.job.out_dev = hit->second;
        cached -= hit->first;
        free_bufs.erase(hit);
      } else {
        r.job.out_dev = nullptr;
        a.fresh.push_back(r.ticket);
      }
      out_used += r.out_bytes;
      r.admitted = true;
      grow[slot] = g;
      est = est_r;
      peak_used = std::max(peak_used, used());
      r.state = 1;
      r.t_start = Clock::now();
      a.tickets.push_back(r.ticket);
      queue.pop_front();
    }
    return a;
  }

  // One worker per context: admit a batch, run it, wait for it on this
  // thread (the host round trips of one batch do not hold up the others),
  // complete its requests.
  void work(int slot) {
    (void)hipSetDevice(device);
    rz_context* c = ctx[slot];
    hipEvent_t ev = nullptr;
    const bool have_ev = hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      if (stop) {  // shutting down: fail what is still queued (running batches finish)
        for (uint64_t t : queue) {
          Request& r = reqs[t];
          r.state = 3;
          r.status = RZ_RUNTIME_ERROR;
          r.error = "server shut down before the request ran";
          r.t_done = Clock::now();
        }
        queue.clear();
        cv_done.notify_all();
        break;
      }
      if (queue.empty()) {
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        continue;
      }
      Admission adm = admit(slot);
      std::vector<uint64_t>& batch = adm.tickets;
      if (batch.empty()) {  // the budget is full: wait for releases
        cv_work.wait_for(lk, std::chrono::microseconds(200));
        continue;
      }
      std::vector<uint64_t> fresh_bytes;
      for (uint64_t t : adm.fresh) fresh_bytes.push_back(reqs[t].out_bytes);
      ++running;
      lk.unlock();
      // evictions and fresh outputs, stream-ordered on this worker's stream
      const auto t_alloc = Clock::now();
      rz_status st = RZ_OK;
      for (void* p : adm.evict)
        if (hipFreeAsync(p, c->stream) != hipSuccess) st = rz::fail(RZ_DEVICE_ERROR, "freeing a cached output failed");
      std::vector<void*> fresh_ptr(adm.fresh.size(), nullptr);
      for (size_t k = 0; k < adm.fresh.size() && st == RZ_OK; ++k)
        if (hipMallocAsync(&fresh_ptr[k], std::max<uint64_t>(fresh_bytes[k], 1), c->stream) != hipSuccess)
          st = rz::fail(RZ_CAPACITY_ERROR, "device memory exhausted allocating an output");
      std::vector<rz_job> jobs;
      jobs.reserve(batch.size());
      {
        std::lock_guard<std::mutex> g(mu);
        for (size_t k = 0; k < adm.fresh.size(); ++k) reqs[adm.fresh[k]].job.out_dev = fresh_ptr[k];
        for (uint64_t t : batch) jobs.push_back(reqs[t].job);
      }
      const auto t_run = Clock::now();
      if (st == RZ_OK) st = run_batch(c, jobs.data(), (int64_t)jobs.size());
      const auto t_wait = Clock::now();
      if (st == RZ_OK && !(have_ev && hipEventRecord(ev, c->stream) == hipSuccess &&
                           hipEventSynchronize(ev) == hipSuccess))
        st = rz::fail(RZ_DEVICE_ERROR, "waiting for a batch failed");
      if (st == RZ_OK) st = check_status_word(c);
      const std::string msg = st ? rz_last_error() : std::string();
      const uint64_t ws_now = rz_context_workspace_bytes(c);
      const auto now = Clock::now();
      const uint64_t us[3] = {(uint64_t)(ms_between(t_alloc, t_run) * 1e3), (uint64_t)(ms_between(t_run, t_wait) * 1e3),
                              (uint64_t)(ms_between(t_wait, now) * 1e3)};
      lk.lock();
      for (int k = 0; k < 3; ++k) {
        phase_sum[k] += us[k];
        phase_max[k] = std::max(phase_max[k], us[k]);
      }
      ws[slot] = ws_now;  // the measured workspace replaces the reservation
      grow[slot] = 0;
      peak_used = std::max(peak_used, used());
      for (uint64_t t : batch) {
        Request& r = reqs[t];
        r.state = st == RZ_OK ? 2 : 3;
        r.status = st;
        r.error = msg;
        r.t_done = now;
        ++n_done;
      }
      --running;
      ++n_batches;
      cv_done.notify_all();
      cv_work.notify_all();  // the reservation returned to the budget
    }
    lk.unlock();
    if (have_ev) (void)hipEventDestroy(ev);
  }
};

extern "C" rz_status rz_server_create(int device, uint64_t memory_budget, int n_streams,
                                      int max_batch, rz_server** out) {
  if (!out || n_streams < 1 || n_streams > 16 || max_batch < 1)
    return rz::fail(RZ_VALUE_ERROR, "rz_server_create: n_streams must be 1..16 and max_batch >= 1");
  auto* s = new (std::nothrow) rz_server();
  if (!s) return rz::fail(RZ_CAPACITY_ERROR, "out of host memory");
  s->device = device;
  s->budget = memory_budget;
  s->max_batch = max_batch;
  for (int k = 0; k < n_streams; ++k) {
    rz_context* c = nullptr;
    rz_status st = rz_context_create(device, &c);
    if (st) {
      for (rz_context* x : s->ctx) rz_context_destroy(x);
      delete s;
      return st;
    }
    s->ctx.push_back(c);
  }
  s->ws.assign((size_t)n_streams, 0);
  s->grow.assign((size_t)n_streams, 0);
  // Outputs and workspaces come from the device's default memory pool in
  // stream order.  A pool returns freed memory to the system at every
  // synchronisation by default, so the next large allocation maps it again
  // (a stall of many milliseconds under load): keep up to the budget
  // reserved in the pool instead.
  hipMemPool_t pool = nullptr;
  if (hipSetDevice(device) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t thr = 0;
    (void)hipMemPoolGetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    if (thr < memory_budget) {
      thr = memory_budget;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
  }
  for (int k = 0; k < n_streams; ++k) s->workers.emplace_back([s, k] { s->work(k); });
  *out = s;
  return RZ_OK;
}

extern "C" rz_status rz_server_destroy(rz_server* s) {
  if (!s) return RZ_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (std::thread& w : s->workers)
    if (w.joinable()) w.join();
  (void)hipSetDevice(s->device);
  for (auto& kv : s->reqs)
    if (kv.second.job.out_dev) (void)hipFree(kv.second.job.out_dev);
  for (auto& kv : s->free_bufs) (void)hipFree(kv.second);
  for (rz_context* c : s->ctx) rz_context_destroy(c);
  delete s;
  return RZ_OK;
}

extern "C" rz_status rz_server_submit(rz_server* s, const rz_job* job, uint64_t* ticket) {
  if (!s || !job || !ticket) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_server_submit");
  rz_status st = check_job_common(*job, /*need_out=*/false);
  if (st) return st;
  Request r;
  r.job = *job;
  r.job.out_dev = nullptr;
  const uint32_t nl = job->layer_dev ? job->n_layers : 1u;
  r.out_bytes = rz_server::size_class_bytes((size_t)nl * job->info.nrows * job->info.ncols *
                                            rz_dtype_size(job->dtype));
  r.ws_bytes = workspace_estimate(*job);
  r.t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r.ticket = s->next_ticket++;
    *ticket = r.ticket;
    s->queue.push_back(r.ticket);
    s->reqs.emplace(r.ticket, std::move(r));
  }
  s->cv_work.notify_one();
  return RZ_OK;
}

extern "C" rz_status rz_server_wait(rz_server* s, uint64_t ticket, int timeout_ms, void** out_dev,
                                    double* latency_ms) {
  if (!s) return rz::fail(RZ_VALUE_ERROR, "NULL server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return rz::fail(RZ_VALUE_ERROR, "unknown ticket");
  auto done = [&] { return it->second.state >= 2; };
  if (timeout_ms < 0) s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
    return rz::fail(RZ_RUNTIME_ERROR, "timed out waiting for the request");
  const Request& r = it->second;
  if (r.state == 3) return rz::fail(r.status, r.error);
  if (out_dev) *out_dev = r.job.out_dev;
  if (latency_ms) *latency_ms = ms_between(r.t_submit, r.t_done);
  return RZ_OK;
}

extern "C" rz_status rz_server_release(rz_server* s, uint64_t ticket) {
  if (!s) return rz::fail(RZ_VALUE_ERROR, "NULL server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return rz::fail(RZ_VALUE_ERROR, "unknown ticket");
  if (it->second.state < 2) return rz::fail(RZ_VALUE_ERROR, "request still running");
  Request& r = it->second;
  if (r.admitted) s->out_used -= r.out_bytes;
  if (r.job.out_dev) {
    s->free_bufs.emplace(r.out_bytes, r.job.out_dev);
    s->cached += r.out_bytes;
  }
  s->reqs.erase(it);
  s->cv_work.notify_one();
  return RZ_OK;
}

extern "C" rz_status rz_server_get_stats(rz_server* s, rz_server_stats* out) {
  if (!s || !out) return rz::fail(RZ_VALUE_ERROR, "NULL argument to rz_server_get_stats");
  std::lock_guard<std::mutex> lk(s->mu);
  out->queued = s->queue.size();
  out->in_flight_batches = s->running;
  out->batches = s->n_batches;
  out->completed = s->n_done;
  out->memory_used = s->used();
  out->memory_peak = s->peak_used;
  out->memory_budget = s->budget;
  out->alloc_us_sum = s->phase_sum[0];
  out->run_us_sum = s->phase_sum[1];
  out->wait_us_sum = s->phase_sum[2];
  out->alloc_us_max = s->phase_max[0];
  out->run_us_max = s->phase_max[1];
  out->wait_us_max = s->phase_max[2];
  return RZ_OK;
}
