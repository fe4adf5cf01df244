// Serving loop: continuous batching of NTT / coset-LDE jobs (no counterpart in
// the reference, which runs one synchronous call at a time).
//
//  - submit() validates and enqueues (FIFO), bounded by max_queue;
//  - max_inflight worker threads each own one HIP stream.  A free worker takes
//    the oldest queued job it may run, plus the queued jobs it can fuse with
//    it (continuous batching); an idle worker polls for work for kPollNs
//    before it sleeps.  Large device-memory jobs (more than kSmallJobElems
//    output elements) (nearly) fill the device: one runs at a time, and a
//    second group (at most kLargeDevGroups) starts only while more large ones
//    wait; host-memory groups, bound by their PCIe copies, run one at a time
//    (a large one alone, in overlapped column chunks); worker 0 runs small
//    jobs only when there are several workers, so small device-memory jobs
//    never queue behind large ones, but while a group of kDeferElems or more
//    runs they wait up to kSmallDeferNs (such a job beside small ones runs
//    slower than alone, and these jobs set the tail latency).  Fusion:
//      NTT jobs of any field, direction and size run together as one
//      many-job launch per pass;
//      LDE jobs of one (log_n, added_bits, shift) run as one batched LDE over
//      their columns, laid out back to back in the worker's staging buffers
//      (host-memory jobs are uploaded there; small device-memory jobs are
//      gathered there and scattered back by one segmented copy each, which
//      saves launches where launches dominate);
//  - admission by HBM footprint: every worker's staging buffers count against
//    mem_budget_bytes for as long as it holds them; a job whose staging does
//    not fit waits (or, for device-memory jobs, runs unfused in place);
//  - every job records submit / start / end times on the host steady clock;
//  - destroy fails the jobs still queued with SFN_ERR_BUSY, lets the running
//    ones finish, and returns once no thread is inside sfn_server_wait.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
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

// jobs up to this many output elements are "small" (the dedicated worker's lane)
constexpr uint64_t kSmallJobElems = 1ull << 22;
// large device-memory groups running at once: two, so that one whose grid
// does not fill the 256 CUs (a few hundred workgroups) leaves the rest to the
// other, while small jobs still find free streams
constexpr uint32_t kLargeDevGroups = 2;
// a small device-memory job waits at most this long for running device-memory
// groups of at least kDeferElems output elements (whose run of milliseconds
// sets the tail latency) before it runs beside them
constexpr uint64_t kSmallDeferNs = 2000000;
constexpr uint64_t kDeferElems = 1ull << 26;
// SFNTT_SERVER_LARGE_DEV=<n> and SFNTT_SERVER_SMALL_DEFER_US=<us> override the
// two limits above (A/B measurement)
uint32_t env_limit(const char *name, uint32_t dflt) {
  const char *e = getenv(name);
  if (!e || !*e) return dflt;
  const long v = strtol(e, nullptr, 10);
  return v >= 1 && v <= 32 ? (uint32_t)v : dflt;
}
// one fused group moves at most this many elements (bounds its latency)
constexpr uint64_t kGroupElems = 1ull << 27;
// an idle worker polls this long for new work before it sleeps: waking a
// sleeping thread costs tens of microseconds, as much as a small job
constexpr uint64_t kPollNs = 100000;

struct Job {
  uint64_t ticket = 0;
  sfn_request req{};
  size_t in_bytes = 0, out_bytes = 0;
  uint64_t submit_ns = 0, start_ns = 0, end_ns = 0;
  uint32_t group_size = 1;
  sfn_status status = SFN_OK;
  bool done = false;

  uint64_t out_elems() const { return out_bytes / 4; }
  bool small() const { return out_elems() <= kSmallJobElems; }
  // LDE jobs fuse when they share the transform; NTT jobs fuse with any NTT job
  bool fuses_with(const Job &o) const {
    if (req.kind != o.req.kind || req.device_ptrs != o.req.device_ptrs) return false;
    if (req.kind == SFN_JOB_NTT) return req.log_n > 0 && o.req.log_n > 0;
    return req.log_n == o.req.log_n && req.added_bits == o.req.added_bits && req.shift_raw == o.req.shift_raw;
  }
};

// A worker's device buffers: staging (in, out) for its current group, and the
// descriptor arrays of its many-job launches and segmented copies.
struct Staging {
  sfn::DevBuf in, out, desc;
  size_t held() const { return in.bytes + out.bytes; }
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
  size_t reserved = 0;  // device bytes of every worker's staging buffers (held or being allocated)
  uint32_t running = 0;
  uint32_t large_dev = 0;   // large device-memory groups running (at most max_large_dev)
  uint32_t huge_dev = 0;    // device-memory groups of at least kDeferElems output elements running
  uint32_t max_large_dev = kLargeDevGroups;
  uint64_t small_defer_ns = kSmallDeferNs;
  uint32_t host_groups = 0;  // host-memory groups running (at most one)
  // host-memory groups run one at a time, so they share one staging pair and
  // one set of transfer streams: it grows to the largest group once and is
  // reused, where per-worker staging would be reallocated as groups move
  // between workers (a hipFree waits for the whole device)
  Staging host_stg;
  size_t host_held = 0;  // host_stg's share of `reserved`
  sfn::ChunkStreams host_cs;
  bool starved = false;        // a host-memory job waits for staging budget
  std::atomic<uint64_t> epoch{0};  // bumped by every submit and every finished group
  uint32_t waiting = 0;        // threads inside sfn_server_wait
  bool stop = false;
  std::vector<std::thread> workers;

  void worker_main(int idx);
  // picks the group under mu; returns false when there is nothing to run
  bool pick(int idx, const Staging &stg, std::vector<Job *> &group, size_t &need_in, size_t &need_out, bool &staged);
  sfn_status run_group(std::vector<Job *> &g, bool staged, hipStream_t s, Staging &st, sfn::ChunkStreams &cs);
};

namespace {

sfn_status grow(sfn::DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes) return SFN_OK;
  if (b.ptr) (void)hipFree(b.ptr);
  b = sfn::DevBuf{};
  if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b.ptr = nullptr;
    return SFN_ERR_BUFFER_ALLOC_FAILED;
  }
  b.bytes = bytes;
  return SFN_OK;
}

void release(sfn::DevBuf &b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b = sfn::DevBuf{};
}

}  // namespace

// Under mu.  The head is the oldest queued job this worker may run (worker 0
// takes small jobs only while other workers exist).  The group is the head
// plus the queued jobs that fuse with it, in arrival order, within max_group,
// max_group_cols (LDE columns) and kGroupElems.  Host-memory groups always
// stage; a device-memory group of two or more LDE jobs stages when the budget
// allows, else it runs in place one job after another.
bool sfn_server::pick(int idx, const Staging &stg, std::vector<Job *> &group, size_t &need_in, size_t &need_out,
                      bool &staged) {
  const bool small_lane = idx == 0 && cfg.max_inflight > 1;
  // the oldest job this worker may start: small ones always; a large one when
  // this is not the small lane and the large-job limits allow it
  // a second large device-memory group starts only while another large one
  // waits behind it (a backlog): alone, a large job runs as fast as on an idle
  // device, as in the plain path
  uint32_t large_dev_queued = 0;
  for (const Job *j : queue) large_dev_queued += j->req.device_ptrs && !j->small();
  const uint64_t now = sfn_now_ns();
  auto may_start = [&](const Job *j) {
    // host-memory groups run one at a time: they are bound by PCIe copies,
    // and two at once share the link, so each takes longer than alone
    if (!j->req.device_ptrs) return host_groups == 0 && (j->small() || !small_lane);
    // a small job beside a running large one takes CUs from it: it waits
    // (for the large group to finish, or small_defer_ns since its submit, seen
    // at the next submit or finished group) unless none runs
    if (j->small()) return huge_dev == 0 || now - j->submit_ns >= small_defer_ns;
    if (small_lane) return false;
    return large_dev == 0 || (large_dev < max_large_dev && large_dev_queued >= 2);
  };
  auto head_it = queue.begin();
  while (head_it != queue.end() && !may_start(*head_it)) ++head_it;
  if (head_it == queue.end()) return false;
  Job *head = *head_it;
  const bool host = !head->req.device_ptrs;
  const Staging &use = host ? host_stg : stg;
  const size_t others = reserved - (host ? host_held : stg.held());  // every other staging
  // staging grows to at least (in, out) and never shrinks while held
  auto fits = [&](size_t in, size_t out) {
    return others + std::max(use.in.bytes, in) + std::max(use.out.bytes, out) <= cfg.mem_budget_bytes;
  };
  if (host && !fits(head->in_bytes, head->out_bytes)) {
    // wait for staging to free up: idle workers give theirs back
    if (!starved) cv_work.notify_all();
    starved = true;
    return false;
  }
  starved = false;
  group.assign(1, head);
  need_in = head->in_bytes;
  need_out = head->out_bytes;
  uint64_t elems = head->out_elems(), cols = head->req.batch;
  // a large host-memory job runs alone, in column chunks whose copies overlap
  // its kernels; fused, its upload, kernels and download would run in turn
  const uint32_t max_group = host && !head->small() ? 1u : cfg.max_group;
  for (auto it = std::next(head_it); it != queue.end() && group.size() < max_group; ++it) {
    Job *j = *it;
    if (!j->fuses_with(*head) || j->small() != head->small()) continue;
    // device-memory LDE jobs are gathered into staging only while launches dominate
    if (j->req.kind == SFN_JOB_LDE && !host && !head->small()) continue;
    if (elems + j->out_elems() > kGroupElems) continue;
    if (j->req.kind == SFN_JOB_LDE && cols + j->req.batch > cfg.max_group_cols) continue;
    if (j->req.kind == SFN_JOB_LDE && !host && !fits(need_in + j->in_bytes, need_out + j->out_bytes)) continue;
    if (host && !fits(need_in + j->in_bytes, need_out + j->out_bytes)) continue;
    group.push_back(j);
    need_in += j->in_bytes;
    need_out += j->out_bytes;
    elems += j->out_elems();
    cols += j->req.batch;
  }
  staged = host || (head->req.kind == SFN_JOB_LDE && group.size() > 1);
  for (Job *j : group) queue.erase(std::find(queue.begin(), queue.end(), j));
  return true;
}

// Runs one group on stream s and waits for it.
sfn_status sfn_server::run_group(std::vector<Job *> &g, bool staged, hipStream_t s, Staging &stg,
                                 sfn::ChunkStreams &xfer) {
  Context &c = *ctx;
  sfn::TablePins pins(&c);
  const sfn_request &r0 = g[0]->req;
  const bool host = !r0.device_ptrs;
  hipError_t e;
  if (host && g.size() == 1 && !g[0]->small()) {
    // one large host-memory job: column chunks on the worker's transfer
    // streams, so its upload, kernels and download overlap (as the plain
    // host-pointer calls do)
    const uint64_t in_rows = 1ull << r0.log_n;
    const uint64_t out_rows = r0.kind == SFN_JOB_LDE ? in_rows << r0.added_bits : in_rows;
    const uint32_t q = sfn::chunk_columns(g[0]->in_bytes, r0.batch, SFN_LAYOUT_SOA);
    if (q) {
      const sfn::ChunkRun run = [&](uint32_t cols, const uint32_t *di, uint32_t *dout, hipStream_t x) {
        return r0.kind == SFN_JOB_LDE
                   ? c.run_lde(r0.log_n, cols, r0.added_bits, r0.shift_raw, SFN_LAYOUT_SOA, di, dout, x)
                   : c.run_ntt(r0.field, SFN_VARIANT_AUTO, r0.dir, r0.log_n, cols, SFN_LAYOUT_SOA, di, dout, x);
      };
      return sfn::run_chunked(c, xfer, s, r0.in, in_rows, r0.out, out_rows, r0.batch, SFN_LAYOUT_SOA, q,
                              (uint32_t *)stg.in.ptr, (uint32_t *)stg.out.ptr, run, nullptr);
    }
  }
  // where every member's input and output live during the transform
  std::vector<const uint32_t *> din(g.size());
  std::vector<uint32_t *> dout(g.size());
  std::vector<sfn::CopySeg> gather, scatter;
  uint64_t gather_total = 0, scatter_total = 0;
  size_t oi = 0, oo = 0;
  for (size_t i = 0; i < g.size(); ++i) {
    if (!staged) {
      din[i] = g[i]->req.in;
      dout[i] = g[i]->req.out;
      continue;
    }
    din[i] = (const uint32_t *)((char *)stg.in.ptr + oi);
    dout[i] = (uint32_t *)((char *)stg.out.ptr + oo);
    if (host) {
      e = hipMemcpyAsync((void *)din[i], g[i]->req.in, g[i]->in_bytes, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
    } else {
      gather.push_back(sfn::CopySeg{g[i]->req.in, (uint32_t *)din[i], gather_total});
      gather_total += g[i]->in_bytes / 4;
      scatter.push_back(sfn::CopySeg{dout[i], g[i]->req.out, scatter_total});
      scatter_total += g[i]->out_bytes / 4;
    }
    oi += g[i]->in_bytes;
    oo += g[i]->out_bytes;
  }
  // device descriptor arrays of this group: [gather][scatter][many-job segments]
  std::vector<sfn::MultiLaunch> ml;
  const bool multi = r0.kind == SFN_JOB_NTT && g.size() > 1;
  if (multi) {
    std::vector<sfn::MultiJob> mj;
    for (size_t i = 0; i < g.size(); ++i) {
      const sfn_request &r = g[i]->req;
      mj.push_back(sfn::MultiJob{r.field, r.dir, r.log_n, r.batch, din[i], dout[i]});
    }
    sfn_status st = sfn::build_multi(&c, mj, ml);
    if (st) return st;
  }
  const size_t cs = sizeof(sfn::CopySeg);
  size_t nseg = 0;
  for (auto &m : ml) nseg += m.segs.size();
  std::vector<char> desc((gather.size() + scatter.size()) * cs + nseg * sizeof(sfn::SegDesc));
  if (!gather.empty()) memcpy(desc.data(), gather.data(), gather.size() * cs);
  if (!scatter.empty()) memcpy(desc.data() + gather.size() * cs, scatter.data(), scatter.size() * cs);
  size_t off = (gather.size() + scatter.size()) * cs;
  for (auto &m : ml) {
    memcpy(desc.data() + off, m.segs.data(), m.segs.size() * sizeof(sfn::SegDesc));
    off += m.segs.size() * sizeof(sfn::SegDesc);
  }
  if (!desc.empty()) {
    if (grow(stg.desc, desc.size())) return SFN_ERR_BUFFER_ALLOC_FAILED;
    e = hipMemcpyAsync(stg.desc.ptr, desc.data(), desc.size(), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
  }
  char *dd = (char *)stg.desc.ptr;
  if (!gather.empty()) {
    e = sfn::launch_copy_segments((const sfn::CopySeg *)dd, (uint32_t)gather.size(), gather_total, s);
    if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
  }
  if (r0.kind == SFN_JOB_LDE) {
    if (staged) {  // the members' columns are back to back: one batched LDE
      uint32_t cols = 0;
      for (Job *j : g) cols += j->req.batch;
      sfn_status st = c.run_lde(r0.log_n, cols, r0.added_bits, r0.shift_raw, SFN_LAYOUT_SOA, din[0], dout[0], s);
      if (st) return st;
    } else {
      for (size_t i = 0; i < g.size(); ++i) {
        const sfn_request &r = g[i]->req;
        sfn_status st = c.run_lde(r.log_n, r.batch, r.added_bits, r.shift_raw, SFN_LAYOUT_SOA, din[i], dout[i], s);
        if (st) return st;
      }
    }
  } else if (!multi) {
    const sfn_request &r = g[0]->req;
    sfn_status st = c.run_ntt(r.field, SFN_VARIANT_AUTO, r.dir, r.log_n, r.batch, SFN_LAYOUT_SOA, din[0], dout[0], s);
    if (st) return st;
  } else {
    size_t o = (gather.size() + scatter.size()) * cs;
    for (auto &m : ml) {
      e = sfn::launch_multi_any(m, (const sfn::SegDesc *)(dd + o), sfn::multi_grid(&c, m), s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
      o += m.segs.size() * sizeof(sfn::SegDesc);
    }
  }
  if (!scatter.empty()) {
    e = sfn::launch_copy_segments((const sfn::CopySeg *)(dd + gather.size() * cs), (uint32_t)scatter.size(),
                                  scatter_total, s);
    if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
  }
  if (host) {
    for (size_t i = 0; i < g.size(); ++i) {
      e = hipMemcpyAsync(g[i]->req.out, dout[i], g[i]->out_bytes, hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return SFN_ERR_GPU_EXECUTION;
    }
  }
  e = hipStreamSynchronize(s);
  return e == hipSuccess ? SFN_OK : SFN_ERR_GPU_EXECUTION;
}

void sfn_server::worker_main(int idx) {
  (void)hipSetDevice(ctx->device);
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
  Staging stg;
  sfn::ChunkStreams cs;  // large host-memory jobs run in column chunks on these
  size_t held = 0;  // this worker's share of `reserved` (its staging, allocated or about to be)
  auto drop_staging = [&] {
    release(stg.in);
    release(stg.out);
    release(stg.desc);
    std::lock_guard<std::mutex> lk(mu);
    reserved -= held;
    held = 0;
  };
  for (;;) {
    std::vector<Job *> group;
    size_t need_in = 0, need_out = 0;
    bool staged = false, shed = false;
    bool large = false;       // the group holds a large device-memory slot
    bool huge = false;        // the group counts in huge_dev
    bool host_group = false;  // the group holds the host-memory slot
    {
      std::unique_lock<std::mutex> lk(mu);
      auto ready = [&] {
        if (stop) return true;
        if (pick(idx, stg, group, need_in, need_out, staged)) return true;
        // an idle worker holding staging gives it back while a job waits for
        // budget (keeping it otherwise: a hipFree waits for the whole device)
        shed = held > 0 && starved;
        return shed;
      };
      if (!ready()) {
        const uint64_t seen = epoch.load(std::memory_order_acquire);
        lk.unlock();
        for (const uint64_t until = sfn_now_ns() + kPollNs;
             epoch.load(std::memory_order_acquire) == seen && sfn_now_ns() < until;)
          std::this_thread::yield();
        lk.lock();
        cv_work.wait(lk, ready);
      }
      if (group.empty() && !shed) break;  // stopping
      if (!group.empty()) {
        host_group = !group[0]->req.device_ptrs;
        if (staged) {
          Staging &u = host_group ? host_stg : stg;
          size_t &h = host_group ? host_held : held;
          const size_t grown = std::max(u.in.bytes, need_in) + std::max(u.out.bytes, need_out);
          reserved += grown - h;
          h = grown;
        }
        ++running;
        large = group[0]->req.device_ptrs && !group[0]->small();
        large_dev += large;
        uint64_t elems = 0;
        for (const Job *j : group) elems += j->out_elems();
        huge = large && elems >= kDeferElems;
        huge_dev += huge;
        host_groups += host_group;
        const uint64_t t = sfn_now_ns();
        for (Job *j : group) {
          j->start_ns = t;
          j->group_size = (uint32_t)group.size();
        }
      }
    }
    if (shed) {
      drop_staging();
      cv_work.notify_all();
      continue;
    }
    sfn_status st = SFN_ERR_GPU_EXECUTION;
    Staging &u = host_group ? host_stg : stg;
    sfn::ChunkStreams &ucs = host_group ? host_cs : cs;
    if (s) {
      st = SFN_OK;
      if (staged && (grow(u.in, need_in) || grow(u.out, need_out))) st = SFN_ERR_BUFFER_ALLOC_FAILED;
      if (!st) st = run_group(group, staged, s, u, ucs);
      if (st) {
        (void)hipStreamSynchronize(s);
        for (auto x : ucs.s)
          if (x) (void)hipStreamSynchronize(x);
      }
    }
    if (st) {  // after a failure the staging in use starts again from nothing
      if (host_group) {
        release(host_stg.in);
        release(host_stg.out);
        release(host_stg.desc);
        std::lock_guard<std::mutex> lk(mu);
        reserved -= host_held;
        host_held = 0;
      } else {
        drop_staging();
      }
    }
    const uint64_t t = sfn_now_ns();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Job *j : group) {
        j->end_ns = t;
        j->status = st;
        j->done = true;
      }
      large_dev -= large;
      huge_dev -= huge;
      host_groups -= host_group;
      --running;
    }
    epoch.fetch_add(1, std::memory_order_release);
    cv_done.notify_all();
    cv_work.notify_all();
  }
  if (s) (void)hipStreamSynchronize(s);
  for (auto x : cs.s)
    if (x) (void)hipStreamSynchronize(x);
  drop_staging();
  cs.release();
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
  // default budget: a quarter of the device's memory (72 GB on a 288 GB MI355X)
  if (!srv->cfg.mem_budget_bytes)
    srv->cfg.mem_budget_bytes = std::max<size_t>(8ull << 30, cx->impl.props.totalGlobalMem / 4);
  if (!srv->cfg.max_inflight) srv->cfg.max_inflight = 6;
  if (srv->cfg.max_inflight > 32) srv->cfg.max_inflight = 32;
  if (!srv->cfg.max_queue) srv->cfg.max_queue = 4096;
  if (!srv->cfg.max_group) srv->cfg.max_group = 128;
  if (!srv->cfg.max_group_cols) srv->cfg.max_group_cols = 512;
  srv->max_large_dev = env_limit("SFNTT_SERVER_LARGE_DEV", kLargeDevGroups);
  if (const char *e = getenv("SFNTT_SERVER_SMALL_DEFER_US")) srv->small_defer_ns = strtoull(e, nullptr, 10) * 1000;
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
  // a host-memory job is staged whole: larger than the budget, it can never run
  if (!req->device_ptrs && in_b + out_b > srv->cfg.mem_budget_bytes) return SFN_ERR_BUFFER_ALLOC_FAILED;
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stop) return SFN_ERR_INVALID_ARGUMENT;
  if (srv->queue.size() >= srv->cfg.max_queue) return SFN_ERR_BUSY;
  job->ticket = srv->next_ticket++;
  job->submit_ns = sfn_now_ns();
  *ticket = job->ticket;
  srv->queue.push_back(job.get());
  srv->jobs.emplace(job->ticket, std::move(job));
  srv->epoch.fetch_add(1, std::memory_order_release);
  // every sleeping worker rechecks: one woken alone may be unable to run the
  // job (the small-job lane, a busy host-memory slot) and sleep again, and
  // the job would then wait for the next submit or finished group
  srv->cv_work.notify_all();
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
  ++srv->waiting;
  srv->cv_done.wait(lk, [&] { return j->done; });
  --srv->waiting;
  fill_timing(*j, timing);
  const sfn_status st = j->status;
  srv->jobs.erase(ticket);
  if (srv->stop) srv->cv_done.notify_all();  // destroy may be waiting for the last waiter
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
    // jobs that have not started will not run: their waiters get SFN_ERR_BUSY
    const uint64_t t = sfn_now_ns();
    for (Job *j : srv->queue) {
      j->status = SFN_ERR_BUSY;
      j->start_ns = j->end_ns = t;
      j->done = true;
    }
    srv->queue.clear();
  }
  srv->cv_work.notify_all();
  srv->cv_done.notify_all();
  for (auto &t : srv->workers) t.join();
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    srv->cv_done.wait(lk, [&] { return srv->waiting == 0; });
  }
  // the workers are gone: the host-memory groups' shared staging and streams
  (void)hipSetDevice(srv->ctx->device);
  for (auto x : srv->host_cs.s)
    if (x) (void)hipStreamSynchronize(x);
  release(srv->host_stg.in);
  release(srv->host_stg.out);
  release(srv->host_stg.desc);
  srv->host_cs.release();
  delete srv;
}

}  // extern "C"
