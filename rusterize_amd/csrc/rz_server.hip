// rz_server.hip -- the serving loop: admission by a device-memory budget,
// continuous batching over several streams, per-request latency.
//
// Included by rz_runtime.hip (it drives run_batch on the server's contexts).
//
// A worker thread owns n_streams contexts (one HIP stream each, default
// priority).  Submitted requests wait in a FIFO queue.  Whenever a context is
// idle, the worker admits queued requests in arrival order while their output
// rasters and estimated workspace fit the memory budget (a request larger
// than the whole budget is failed), up to max_batch of them, allocates their
// outputs and runs them as one batch on that context.  Batches on different
// contexts overlap on the device, so requests arriving while others run are
// admitted as soon as a context and memory free up.  Completion is observed
// through one event per batch.  Outputs stay on the device until the caller
// releases them, which returns their memory to the budget.
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

struct InFlight {
  int slot = 0;
  hipEvent_t done = nullptr;
  std::vector<uint64_t> tickets;
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
  std::vector<bool> busy;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> queue;
  std::unordered_map<uint64_t, Request> reqs;
  std::vector<InFlight> inflight;
  std::unordered_multimap<size_t, void*> free_bufs;  // released outputs, by size
  uint64_t next_ticket = 1;
  uint64_t used = 0;       // outputs and workspace of admitted requests
  uint64_t cached = 0;     // bytes in free_bufs
  uint64_t n_batches = 0, n_done = 0, peak_used = 0;
  bool stop = false;
  std::thread worker;

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

  // Completes the batches whose events have fired.
  void reap() {
    for (size_t k = 0; k < inflight.size();) {
      InFlight& f = inflight[k];
      if (hipEventQuery(f.done) != hipSuccess) {
        ++k;
        continue;
      }
      const auto now = Clock::now();
      rz_status st = check_status_word(ctx[f.slot]);
      for (uint64_t t : f.tickets) {
        Request& r = reqs[t];
        r.state = st == RZ_OK ? 2 : 3;
        r.status = st;
        if (st != RZ_OK) r.error = rz_last_error();
        r.t_done = now;
        used -= r.ws_bytes;  // the workspace estimate is released; the output stays
        r.ws_bytes = 0;
        ++n_done;
      }
      busy[f.slot] = false;
      (void)hipEventDestroy(f.done);
      inflight.erase(inflight.begin() + (long)k);
      cv_done.notify_all();
    }
  }

  void loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
      reap();
      int slot = -1;
      for (size_t k = 0; k < busy.size(); ++k)
        if (!busy[k]) {
          slot = (int)k;
          break;
        }
      if (stop) {  // shutting down: fail what is still queued, drain what runs
        for (uint64_t t : queue) {
          Request& r = reqs[t];
          r.state = 3;
          r.status = RZ_RUNTIME_ERROR;
          r.error = "server shut down before the request ran";
          r.t_done = Clock::now();
        }
        queue.clear();
        cv_done.notify_all();
        if (inflight.empty()) break;
      }
      if (slot < 0 || queue.empty()) {
        // nothing to launch: wait for a submission or poll the running batches
        if (inflight.empty()) cv_work.wait(lk, [&] { return stop || !queue.empty(); });
        else cv_work.wait_for(lk, std::chrono::microseconds(50));
        continue;
      }
      // admit queued requests in arrival order while they fit
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
      if (batch.empty()) {
        cv_work.wait_for(lk, std::chrono::microseconds(50));
        continue;
      }
      std::vector<rz_job> jobs;
      jobs.reserve(batch.size());
      for (uint64_t t : batch) jobs.push_back(reqs[t].job);
      busy[slot] = true;
      rz_context* c = ctx[slot];
      lk.unlock();
      rz_status st = run_batch(c, jobs.data(), (int64_t)jobs.size());
      hipEvent_t ev = nullptr;
      if (st == RZ_OK && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
          hipEventRecord(ev, c->stream) == hipSuccess) {
        lk.lock();
        inflight.push_back(InFlight{slot, ev, batch});
        ++n_batches;
      } else {
        const std::string msg = st ? rz_last_error() : "event creation failed";
        lk.lock();
        for (uint64_t t : batch) {
          Request& r = reqs[t];
          r.state = 3;
          r.status = st ? st : RZ_DEVICE_ERROR;
          r.error = msg;
          r.t_done = Clock::now();
          used -= r.ws_bytes;
          r.ws_bytes = 0;
        }
        busy[slot] = false;
        cv_done.notify_all();
      }
    }
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
    s->busy.push_back(false);
  }
  s->worker = std::thread([s] { s->loop(); });
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
  if (s->worker.joinable()) s->worker.join();
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
  r.out_bytes = (size_t)nl * job->info.nrows * job->info.ncols * rz_dtype_size(job->dtype);
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
  out->in_flight_batches = s->inflight.size();
  out->batches = s->n_batches;
  out->completed = s->n_done;
  out->memory_used = s->used;
  out->memory_peak = s->peak_used;
  out->memory_budget = s->budget;
  return RZ_OK;
}
