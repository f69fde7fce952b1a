// rz_server.hip -- the serving loop: admission by a device-memory budget,
// continuous batching over several streams, per-request latency.
//
// Included by rz_runtime.hip (it drives run_batch on the server's contexts).
//
// n_streams worker threads each own one context (one HIP stream, default
// priority).  Submitted requests wait in a FIFO queue.  An idle worker admits
// queued requests in arrival order while their output rasters and estimated
// workspace fit the memory budget (a request larger than the whole budget is
// failed), up to max_batch of them, allocates their outputs, runs them as one
// batch on its context and waits for it.  Batches of different workers
// overlap on the device, and one batch's host round trips (the eager
// pipeline reads a few totals back) do not hold up the others, so requests
// arriving while others run are admitted as soon as a worker and memory free
// up.  Outputs stay on the device until the caller releases them, which
// returns their memory to the budget.
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
  rz_status status = RZ_OK;
  std::string error;
  Clock::time_point t_submit, t_start, t_done;
};

// Device memory a request may need besides its output: edge records and
// coordinates per vertex, descriptors and bins per item, and the wide
// pre-pass masks (capped at 2 bytes per output pixel).
size_t workspace_estimate(const rz_job& j) {
  const rz_device_geometry* g = j.geometry;
  const size_t px = (size_t)j.info.nrows * (size_t)j.info.ncols * (j.layer_dev ? j.n_layers : 1u);
  return (size_t)g->V * 64 + (size_t)g->I * 512 + 2 * px;
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
  uint64_t used = 0;       // outputs and workspace of admitted requests
  uint64_t cached = 0;     // bytes in free_bufs
  uint64_t n_batches = 0, n_done = 0, peak_used = 0, running = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  // Output buffers come in size classes (four per power of two) so released
  // outputs are reused by later requests of similar size: fresh allocations,
  // and the hipFree of cached ones (which synchronises the device), stay rare.
  static size_t size_class_bytes(size_t bytes) {
    size_t c = 1u << 20;
    while (c < bytes) c <<= 1;  // next power of two >= 1 MiB
    const size_t q = c / 8;     // quarters of the octave [c / 2, c)
    return c <= (1u << 20) ? c : ((bytes + q - 1) / q) * q;
  }

  void* take_buffer(size_t bytes) {
    auto it = free_bufs.find(bytes);
    if (it != free_bufs.end()) {
      void* p = it->second;
      free_bufs.erase(it);
      cached -= bytes;
      return p;
    }
    // make room in the cache for the budget before a fresh allocation
    while (cached > 0 && used + cached + bytes > budget) {
      auto victim = free_bufs.begin();
      cached -= victim->first;
      (void)hipFree(victim->second);
      free_bufs.erase(victim);
    }
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 1)) != hipSuccess) return nullptr;
    return p;
  }

  // Admits queued requests in arrival order while they fit (lock held).
  std::vector<uint64_t> admit() {
    std::vector<uint64_t> batch;
    while (!queue.empty() && (int)batch.size() < max_batch) {
      Request& r = reqs[queue.front()];
      const uint64_t need = r.out_bytes + r.ws_bytes;
      if (need > budget) {  // can never fit
        r.state = 3;
        r.status = RZ_CAPACITY_ERROR;
        r.error = "request exceeds the server's memory budget";
        r.t_done = Clock::now();
        queue.pop_front();
        cv_done.notify_all();
        continue;
      }
      if (used + need > budget) break;
      void* out = take_buffer(r.out_bytes);
      if (!out) break;  // device memory exhausted: wait for releases
      r.job.out_dev = out;
      used += need;
      peak_used = std::max(peak_used, used);
      r.state = 1;
      r.t_start = Clock::now();
      batch.push_back(r.ticket);
      queue.pop_front();
    }
    return batch;
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
      std::vector<uint64_t> batch = admit();
      if (batch.empty()) {  // the budget is full: wait for releases
        cv_work.wait_for(lk, std::chrono::microseconds(200));
        continue;
      }
      std::vector<rz_job> jobs;
      jobs.reserve(batch.size());
      for (uint64_t t : batch) jobs.push_back(reqs[t].job);
      ++running;
      lk.unlock();
      rz_status st = run_batch(c, jobs.data(), (int64_t)jobs.size());
      if (st == RZ_OK && !(have_ev && hipEventRecord(ev, c->stream) == hipSuccess &&
                           hipEventSynchronize(ev) == hipSuccess))
        st = rz::fail(RZ_DEVICE_ERROR, "waiting for a batch failed");
      if (st == RZ_OK) st = check_status_word(c);
      const std::string msg = st ? rz_last_error() : std::string();
      const auto now = Clock::now();
      lk.lock();
      for (uint64_t t : batch) {
        Request& r = reqs[t];
        r.state = st == RZ_OK ? 2 : 3;
        r.status = st;
        r.error = msg;
        r.t_done = now;
        used -= r.ws_bytes;  // the workspace estimate is released; the output stays
        r.ws_bytes = 0;
        ++n_done;
      }
      --running;
      ++n_batches;
      cv_done.notify_all();
      cv_work.notify_all();  // workspace returned to the budget
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
  if (r.job.out_dev) {
    s->free_bufs.emplace(r.out_bytes, r.job.out_dev);
    s->cached += r.out_bytes;
    s->used -= r.out_bytes;
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
  out->memory_used = s->used;
  out->memory_peak = s->peak_used;
  out->memory_budget = s->budget;
  return RZ_OK;
}
