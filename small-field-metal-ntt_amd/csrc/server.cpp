// Serving loop: continuous batching of NTT / coset-LDE jobs (no counterpart in
// the reference, which runs one synchronous call at a time).
//
//  - submit() validates and enqueues (FIFO), bounded by max_queue;
//  - max_inflight worker threads each own one HIP stream; a free worker takes
//    the queue head once its device-memory footprint fits the budget
//    (admission by HBM footprint), plus queued jobs of the same shape
//    (continuous batching): NTT jobs (any field) of one (field, dir, log_n)
//    run as one multi-job launch per pass, LDE jobs of one shape run back to
//    back on the worker's stream with a single completion wait;
//  - host-memory jobs are staged through the worker's device buffers;
//    device-memory jobs (device_ptrs = 1) run in place in HBM;
//  - every job records submit / start / end times on the host steady clock.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "runtime.hpp"
#include "tables.hpp"

using sfn::Context;

extern uint64_t sfn_now_ns();

namespace {

struct Job {
  uint64_t ticket = 0;
  sfn_request req{};
  size_t in_bytes = 0, out_bytes = 0;
  uint64_t submit_ns = 0, start_ns = 0, end_ns = 0;
  uint32_t group_size = 1;
  sfn_status status = SFN_OK;
  bool done = false;

  size_t footprint() const { return req.device_ptrs ? 0 : in_bytes + out_bytes; }
  bool same_shape(const Job &o) const {
    return req.kind == o.req.kind && req.field == o.req.field && req.dir == o.req.dir &&
           req.log_n == o.req.log_n && req.added_bits == o.req.added_bits && req.shift_raw == o.req.shift_raw &&
           req.device_ptrs == o.req.device_ptrs;
  }
};

}  // namespace

struct sfn_server {
  Context *ctx = nullptr;
  sfn_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job *> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  uint64_t next_ticket = 1;
  size_t reserved = 0;  // device bytes held by running host-memory jobs
  uint32_t running = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  void worker_main(int idx);
  sfn_status run_group(std::vector<Job *> &g, hipStream_t s, sfn::DevBuf &bin, sfn::DevBuf &bout,
                       sfn::DevBuf &segs);
};

namespace {

sfn_status grow(Context *c, sfn::DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes) return SFN_OK;
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b.ptr = nullptr;
    return SFN_ERR_BUFFER_ALLOC_FAILED;
  }
  b.bytes = bytes;
  (void)c;
  return SFN_OK;
}

}  // namespace

// Runs one group (same shape) on stream s and waits for it.
sfn_status sfn_server::run_group(std::vector<Job *> &g, hipStream_t s, sfn::DevBuf &bin, sfn::DevBuf &bout,
                                 sfn::DevBuf &segs) {
  Context &c = *ctx;
  const sfn_request &r0 = g[0]->req;
  hipError_t e;
  // device pointers of every member
  std::vector<const uint32_t *> din(g.size());
  std::vector<uint32_t *> dout(g.size());
  if (r0.device_ptrs) {
    for (size_t i = 0; i < g.size(); ++i) {
      din[i] = g[i]->req.in;
      dout[i] = g[i]->req.out;
    }
  } else {
    size_t tin = 0, tout = 0;
    for (Job *j : g) {
      tin += j->in_bytes;
      tout += j->out_bytes;
    }
    sfn_status st = grow(&c, bin, tin);
    if (!st) st = grow(&c, bout, tout);
    if (st) return st;
    size_t oi = 0, oo = 0;
    for (size_t i = 0; i < g.size(); ++i) {
      din[i] = (const uint32_t *)((char *)bin.ptr + oi);
      dout[i] = (uint32_t *)((char *)bout.ptr + oo);
      e = hipMemcpyAsync((void *)din[i], g[i]->req.in, g[i]->in_bytes, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
      oi += g[i]->in_bytes;
      oo += g[i]->out_bytes;
    }
  }
  if (r0.kind == SFN_JOB_LDE) {
    for (size_t i = 0; i < g.size(); ++i) {
      const sfn_request &r = g[i]->req;
      sfn_status st = c.run_lde(r.log_n, r.batch, r.added_bits, r.shift_raw, SFN_LAYOUT_SOA, din[i], dout[i], s);
      if (st) return st;
    }
  } else if (g.size() == 1 || r0.log_n == 0) {
    for (size_t i = 0; i < g.size(); ++i) {
      const sfn_request &r = g[i]->req;
      sfn_status st = c.run_ntt(r.field, SFN_VARIANT_AUTO, r.dir, r.log_n, r.batch, SFN_LAYOUT_SOA, din[i], dout[i], s);
      if (st) return st;
    }
  } else {
    // one multi-job launch per pass
    std::vector<sfn::MultiJob> mj;
    for (size_t i = 0; i < g.size(); ++i) {
      const sfn_request &r = g[i]->req;
      mj.push_back(sfn::MultiJob{r.field, r.dir, r.log_n, r.batch, din[i], dout[i]});
    }
    std::vector<sfn::MultiLaunch> ml;
    sfn_status st = sfn::build_multi(&c, mj, ml);
    if (st) return st;
    size_t total = 0;
    for (auto &m : ml) total += m.segs.size() * sizeof(sfn::SegDesc);
    if (grow(&c, segs, total)) return SFN_ERR_BUFFER_ALLOC_FAILED;
    std::vector<sfn::SegDesc> flat;
    for (auto &m : ml) flat.insert(flat.end(), m.segs.begin(), m.segs.end());
    e = hipMemcpyAsync(segs.ptr, flat.data(), total, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
    size_t off = 0;
    for (auto &m : ml) {
      e = sfn::launch_multi_any(m, (const sfn::SegDesc *)segs.ptr + off, sfn::multi_grid(&c, m), s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
      off += m.segs.size();
    }
  }
  if (!r0.device_ptrs) {
    for (size_t i = 0; i < g.size(); ++i) {
      e = hipMemcpyAsync(g[i]->req.out, dout[i], g[i]->out_bytes, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
    }
  }
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? SFN_OK : SFN_ERR_GPU_EXECUTION;
}

void sfn_server::worker_main(int idx) {
  (void)idx;
  (void)hipSetDevice(ctx->device);
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  sfn::DevBuf bin, bout, segs;
  for (;;) {
    std::vector<Job *> group;
    size_t foot = 0;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_work.wait(lk, [&] {
        return stop || (!queue.empty() && reserved + queue.front()->footprint() <= cfg.mem_budget_bytes);
      });
      if (stop && queue.empty()) break;
      if (stop && !queue.empty() && reserved + queue.front()->footprint() > cfg.mem_budget_bytes) break;
      Job *head = queue.front();
      queue.pop_front();
      group.push_back(head);
      foot = head->footprint();
      uint64_t cols = head->req.batch;
      // continuous batching: take queued jobs of the same shape, small ones only
      const size_t small = 16u << 20;
      if (cfg.max_group > 1 && head->in_bytes + head->out_bytes <= small) {
        for (auto it = queue.begin(); it != queue.end() && group.size() < cfg.max_group;) {
          Job *j = *it;
          if (j->same_shape(*head) && cols + j->req.batch <= cfg.max_group_cols &&
              reserved + foot + j->footprint() <= cfg.mem_budget_bytes) {
            group.push_back(j);
            foot += j->footprint();
            cols += j->req.batch;
            it = queue.erase(it);
          } else {
            ++it;
          }
        }
      }
      reserved += foot;
      ++running;
      const uint64_t t = sfn_now_ns();
      for (Job *j : group) {
        j->start_ns = t;
        j->group_size = (uint32_t)group.size();
      }
    }
    sfn_status st = s ? run_group(group, s, bin, bout, segs) : SFN_ERR_GPU_EXECUTION;
    const uint64_t t = sfn_now_ns();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Job *j : group) {
        j->end_ns = t;
        j->status = st;
        j->done = true;
      }
      reserved -= foot;
      --running;
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
  if (s) (void)hipStreamSynchronize(s);
  for (sfn::DevBuf *b : {&bin, &bout, &segs})
    if (b->ptr) (void)hipFree(b->ptr);
  if (s) (void)hipStreamDestroy(s);
}

extern "C" {

uint64_t sfn_now(void) { return sfn_now_ns(); }

sfn_status sfn_host_alloc(sfn_ctx *ctx, size_t bytes, void **ptr) {
  if (!ctx || !ptr) return SFN_ERR_INVALID_ARGUMENT;
  *ptr = nullptr;
  if (bytes == 0) return SFN_ERR_BUFFER_ALLOC_FAILED;
  (void)hipSetDevice(ctx->impl.device);
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    return ctx->impl.fail(SFN_ERR_BUFFER_ALLOC_FAILED, "hipHostMalloc failed");
  }
  return SFN_OK;
}

sfn_status sfn_host_free(sfn_ctx *ctx, void *ptr) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  if (ptr) (void)hipHostFree(ptr);
  return SFN_OK;
}

sfn_status sfn_server_create(sfn_ctx *cx, const sfn_server_config *cfg, sfn_server **out) {
  if (!cx || !out) return SFN_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  auto srv = std::make_unique<sfn_server>();
  srv->ctx = &cx->impl;
  if (cfg) srv->cfg = *cfg;
  if (!srv->cfg.mem_budget_bytes) srv->cfg.mem_budget_bytes = 8ull << 30;
  if (!srv->cfg.max_inflight) srv->cfg.max_inflight = 4;
  if (srv->cfg.max_inflight > 32) srv->cfg.max_inflight = 32;
  if (!srv->cfg.max_queue) srv->cfg.max_queue = 4096;
  if (!srv->cfg.max_group) srv->cfg.max_group = 64;
  if (!srv->cfg.max_group_cols) srv->cfg.max_group_cols = 512;
  for (uint32_t i = 0; i < srv->cfg.max_inflight; ++i)
    srv->workers.emplace_back(&sfn_server::worker_main, srv.get(), (int)i);
  *out = srv.release();
  return SFN_OK;
}

sfn_status sfn_server_submit(sfn_server *srv, const sfn_request *req, uint64_t *ticket) {
  if (!srv || !req || !ticket) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = *srv->ctx;
  sfn_status st;
  size_t in_b, out_b;
  if (req->kind == SFN_JOB_LDE) {
    if (req->field != SFN_FIELD_BABYBEAR) return SFN_ERR_UNSUPPORTED;
    if (req->batch == 0 || req->log_n + req->added_bits + 1 > 27) return SFN_ERR_INVALID_SIZE;
    in_b = ((size_t)req->batch << req->log_n) * 4;
    out_b = ((size_t)req->batch << (req->log_n + req->added_bits)) * 4;
  } else if (req->kind == SFN_JOB_NTT) {
    if ((req->field != SFN_FIELD_M31 && req->field != SFN_FIELD_BABYBEAR && req->field != SFN_FIELD_KOALABEAR) ||
        (req->dir != SFN_FORWARD && req->dir != SFN_INVERSE))
      return SFN_ERR_INVALID_ARGUMENT;
    const uint32_t max_log = req->field == SFN_FIELD_M31 ? 30u : sfn::host::monty_field(req->field).max_log;
    if (req->batch == 0 || req->log_n > max_log) return SFN_ERR_INVALID_SIZE;
    in_b = out_b = ((size_t)req->batch << req->log_n) * 4;
  } else {
    return SFN_ERR_INVALID_ARGUMENT;
  }
  (void)st;
  (void)c;
  if (!req->in || !req->out) return SFN_ERR_INVALID_ARGUMENT;
  // an LDE writes column b at b*N while later columns are still to be read
  if (req->kind == SFN_JOB_LDE && sfn::ranges_overlap(req->in, in_b, req->out, out_b)) return SFN_ERR_INVALID_ARGUMENT;
  auto job = std::make_unique<Job>();
  job->req = *req;
  job->in_bytes = in_b;
  job->out_bytes = out_b;
  if (job->footprint() > srv->cfg.mem_budget_bytes) return SFN_ERR_BUFFER_ALLOC_FAILED;  // can never be admitted
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return SFN_ERR_INVALID_ARGUMENT;
  if (srv->queue.size() >= srv->cfg.max_queue) return SFN_ERR_BUSY;
  job->ticket = srv->next_ticket++;
  job->submit_ns = sfn_now_ns();
  *ticket = job->ticket;
  srv->queue.push_back(job.get());
  srv->jobs.emplace(job->ticket, std::move(job));
  srv->cv_work.notify_one();
  return SFN_OK;
}

static void fill_timing(const Job &j, sfn_job_timing *t) {
  if (!t) return;
  t->submit_ns = j.submit_ns;
  t->start_ns = j.start_ns;
  t->end_ns = j.end_ns;
  t->group_size = j.group_size;
  t->status = j.status;
}

sfn_status sfn_server_poll(sfn_server *srv, uint64_t ticket, sfn_job_timing *timing) {
  if (!srv) return SFN_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(ticket);
  if (it == srv->jobs.end()) return SFN_ERR_INVALID_ARGUMENT;
  if (!it->second->done) return SFN_ERR_NOT_READY;
  fill_timing(*it->second, timing);
  const sfn_status st = it->second->status;
  srv->jobs.erase(it);
  return st;
}

sfn_status sfn_server_wait(sfn_server *srv, uint64_t ticket, sfn_job_timing *timing) {
  if (!srv) return SFN_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->jobs.find(ticket);
  if (it == srv->jobs.end()) return SFN_ERR_INVALID_ARGUMENT;
  Job *j = it->second.get();
  srv->cv_done.wait(lk, [&] { return j->done; });
  fill_timing(*j, timing);
  const sfn_status st = j->status;
  srv->jobs.erase(ticket);
  return st;
}

sfn_status sfn_server_drain(sfn_server *srv) {
  if (!srv) return SFN_ERR_INVALID_ARGUMENT;
  std::unique_lock<std::mutex> lk(srv->mu);
  srv->cv_done.wait(lk, [&] { return srv->queue.empty() && srv->running == 0; });
  return SFN_OK;
}

void sfn_server_destroy(sfn_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  for (auto &t : srv->workers) t.join();
  delete srv;
}

}  // extern "C"
