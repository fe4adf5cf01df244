// Batched plans (many transforms of uneven sizes and both fields in one
// launch per pass, replayed from a hipGraph) and the schedule autotuner.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <sstream>

#include "runtime.hpp"

using sfn::Context;
using sfn::PassArgs;
using sfn::PassDesc;
using sfn::Schedule;
using sfn::SegDesc;

struct sfn_plan {
  Context *ctx = nullptr;
  struct Launch {
    sfn::MultiLaunch m;
    void *d_segs = nullptr;
    uint32_t grid = 0;
  };
  std::vector<Launch> launches;
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t capture_stream = nullptr;
  uint64_t total_tiles = 0;
  std::vector<sfn::TableKey> pinned;  // cached tables the descriptors point to

  ~sfn_plan() {
    if (ctx && !pinned.empty()) ctx->unpin(pinned);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (capture_stream) (void)hipStreamDestroy(capture_stream);
    for (auto &l : launches)
      if (l.d_segs) (void)hipFree(l.d_segs);
  }

  hipError_t issue(hipStream_t s) {
    for (auto &l : launches) {
      hipError_t e = sfn::launch_multi_any(l.m, (const SegDesc *)l.d_segs, l.grid, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

extern "C" {

sfn_status sfn_plan_create(sfn_ctx *cx, const sfn_job_desc *jobs, uint32_t count, int use_graph, sfn_plan **out) {
  if (!cx || !out || (!jobs && count)) return SFN_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  Context &c = cx->impl;
  (void)hipSetDevice(c.device);
  auto plan = std::make_unique<sfn_plan>();
  plan->ctx = &c;
  std::vector<sfn::MultiJob> mj;
  for (uint32_t j = 0; j < count; ++j) {
    const sfn_job_desc &jd = jobs[j];
    sfn_status st = sfn::validate_ntt(&c, jd.field, SFN_VARIANT_AUTO, jd.dir, jd.log_n, jd.batch, SFN_LAYOUT_SOA);
    if (st) return st;
    if (!jd.d_in || !jd.d_out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null job buffer");
    if (jd.log_n == 0) {
      if (jd.d_in != jd.d_out) return c.fail(SFN_ERR_UNSUPPORTED, "out-of-place log_n == 0 job in a plan");
      continue;  // identity in place
    }
    mj.push_back(sfn::MultiJob{jd.field, jd.dir, jd.log_n, jd.batch, jd.d_in, jd.d_out});
  }
  std::vector<sfn::MultiLaunch> ml;
  sfn::TablePins pins(&c);  // the plan keeps every table its descriptors name
  sfn_status bst = sfn::build_multi(&c, mj, ml);
  if (bst) return bst;
  plan->pinned = pins.take();
  for (auto &m : ml) {
    sfn_plan::Launch L;
    L.m = std::move(m);
    L.grid = sfn::multi_grid(&c, L.m);
    const size_t bytes = L.m.segs.size() * sizeof(SegDesc);
    if (bytes) {
      hipError_t e = hipMalloc(&L.d_segs, bytes);
      if (e != hipSuccess) return c.hip_fail(e, "plan descriptors");
      e = hipMemcpy(L.d_segs, L.m.segs.data(), bytes, hipMemcpyHostToDevice);
      if (e != hipSuccess) return c.hip_fail(e, "plan descriptors");
    }
    plan->total_tiles += L.m.tiles;
    plan->launches.push_back(std::move(L));
  }
  if (use_graph) {
    hipError_t e = hipStreamCreateWithFlags(&plan->capture_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return c.hip_fail(e, "stream");
    e = hipStreamBeginCapture(plan->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return c.hip_fail(e, "graph capture");
    hipError_t le = plan->issue(plan->capture_stream);
    e = hipStreamEndCapture(plan->capture_stream, &plan->graph);
    if (le != hipSuccess) return c.hip_fail(le, "graph capture launch");
    if (e != hipSuccess) return c.hip_fail(e, "graph capture");
    e = hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return c.hip_fail(e, "graph instantiate");
    plan->use_graph = true;
  }
  *out = plan.release();
  return SFN_OK;
}

sfn_status sfn_plan_launch(sfn_plan *plan, void *stream) {
  if (!plan) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = *plan->ctx;
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  hipError_t e = plan->use_graph ? hipGraphLaunch(plan->exec, s) : plan->issue(s);
  if (e != hipSuccess) return c.hip_fail(e, "plan launch");
  return SFN_OK;
}

sfn_status sfn_plan_info(sfn_plan *plan, uint32_t *launches, uint32_t *tiles) {
  if (!plan) return SFN_ERR_INVALID_ARGUMENT;
  if (launches) *launches = (uint32_t)plan->launches.size();
  if (tiles) *tiles = (uint32_t)std::min<uint64_t>(plan->total_tiles, 0xFFFFFFFFu);
  return SFN_OK;
}

void sfn_plan_destroy(sfn_plan *plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->ctx->device);
  (void)hipDeviceSynchronize();
  delete plan;
}

// ------------------------------------------------------------------ autotuner
// Candidates: largest LDS sub-transform (pass split), register radix and lanes
// per workgroup (and the LDE middle pass's lanes).  Each is timed with a
// hipEvent pair over a few launches on scratch buffers; the fastest is kept.
sfn_status sfn_autotune(sfn_ctx *cx, sfn_field field, uint32_t log_n, uint32_t batch, sfn_layout layout,
                        uint32_t *best_log_tile) {
  if (!cx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = cx->impl;
  const bool lde = (int)field == SFN_TUNE_LDE;  // the coset LDE at 2x
  sfn_status st = lde ? sfn::validate_lde(&c, log_n, batch, 1, layout)
                      : sfn::validate_ntt(&c, field, SFN_VARIANT_AUTO, SFN_FORWARD, log_n, batch, layout);
  if (st) return st;
  if (log_n < 2) {
    if (best_log_tile) *best_log_tile = log_n;
    return SFN_OK;
  }
  (void)hipSetDevice(c.device);
  const uint64_t ni = (uint64_t)batch << log_n, no = lde ? ni * 2 : ni;
  if ((st = c.ensure(c.scratch_in, ni * 4))) return st;
  if ((st = c.ensure(c.scratch_out, no * 4))) return st;
  (void)hipMemsetAsync(c.scratch_in.ptr, 0, ni * 4, c.stream);
  std::vector<Schedule> cands;
  for (uint32_t mlt = 6; mlt <= 11; ++mlt) {
    if ((log_n + mlt - 1) / mlt > 3) continue;
    if (lde && (log_n > 2 * mlt)) continue;
    for (uint32_t logr : {3u, 4u, 5u})
      for (uint32_t lw : {2u, 3u, 4u, 5u})
        for (uint32_t mlw : {2u, 3u, 4u, 5u}) {
          if (!lde && mlw != lw) continue;  // the middle-pass width only matters to the LDE
          Schedule s;
          s.max_log_t = mlt;
          s.logr = logr;
          s.log_w = lw;
          s.mid_log_w = mlw;
          if (mlt + std::max(lw, mlw) > 15) continue;
          cands.push_back(s);
        }
  }
  float best = 1e30f;
  Schedule bs = c.schedule_for(lde ? SFN_TUNE_LDE : (int)field, log_n, batch, layout);
  const uint32_t *din = (const uint32_t *)c.scratch_in.ptr;
  uint32_t *dout = (uint32_t *)c.scratch_out.ptr;
  for (const Schedule &s : cands) {
    auto run = [&]() {
      return lde ? c.run_lde(log_n, batch, 1, 0, layout, din, dout, c.stream, &s)
                 : c.run_ntt(field, SFN_VARIANT_AUTO, SFN_FORWARD, log_n, batch, layout, din, dout, c.stream, &s);
    };
    if (run() != SFN_OK) {  // candidate not launchable (LDS / lane limits): skip it
      (void)hipGetLastError();
      continue;
    }
    (void)hipEventRecord(c.ev0, c.stream);
    const int reps = ni >= (1u << 24) ? 3 : 10;
    bool ok = true;
    for (int r = 0; r < reps; ++r) ok = ok && run() == SFN_OK;
    (void)hipEventRecord(c.ev1, c.stream);
    if (hipEventSynchronize(c.ev1) != hipSuccess || !ok) continue;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c.ev0, c.ev1);
    if (ms < best) {
      best = ms;
      bs = s;
    }
  }
  {
    std::lock_guard<std::mutex> lk(c.tune_mu);
    c.tuned[sfn::TuneKey{lde ? SFN_TUNE_LDE : (int)field, log_n, batch, (int)layout}] = bs;
  }
  c.timing_pending = false;  // the event pair now brackets a timing run, not a caller's launch
  if (best_log_tile) *best_log_tile = bs.max_log_t;
  c.clear_error();
  return SFN_OK;
}

sfn_status sfn_tuning_save(sfn_ctx *cx, char *buf, size_t cap, size_t *len) {
  if (!cx) return SFN_ERR_INVALID_ARGUMENT;
  std::ostringstream o;
  o << "[";
  bool first = true;
  std::lock_guard<std::mutex> lk(cx->impl.tune_mu);
  for (auto &kv : cx->impl.tuned) {
    if (!first) o << ",";
    first = false;
    o << "{\"field\":" << kv.first.field << ",\"log_n\":" << kv.first.log_n << ",\"batch\":" << kv.first.batch
      << ",\"layout\":" << kv.first.layout << ",\"max_log_t\":" << kv.second.max_log_t
      << ",\"logr\":" << kv.second.logr << ",\"log_w\":" << kv.second.log_w
      << ",\"mid_log_w\":" << kv.second.mid_log_w << "}";
  }
  o << "]";
  const std::string s = o.str();
  if (len) *len = s.size() + 1;
  if (!buf || cap < s.size() + 1) return buf ? SFN_ERR_INVALID_SIZE : SFN_OK;
  memcpy(buf, s.c_str(), s.size() + 1);
  return SFN_OK;
}

// Accepts the array written by sfn_tuning_save (flat objects of integers).
sfn_status sfn_tuning_load(sfn_ctx *cx, const char *json) {
  if (!cx || !json) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = cx->impl;
  const char *p = json;
  std::map<sfn::TuneKey, Schedule> add;
  while ((p = strchr(p, '{')) != nullptr) {
    const char *e = strchr(p, '}');
    if (!e) return c.fail(SFN_ERR_INVALID_ARGUMENT, "unterminated object in tuning table");
    std::string obj(p + 1, e);
    auto get = [&](const char *key, long *v) {
      std::string k = std::string("\"") + key + "\":";
      size_t i = obj.find(k);
      if (i == std::string::npos) return false;
      *v = strtol(obj.c_str() + i + k.size(), nullptr, 10);
      return true;
    };
    long f, ln, b, lay, mlt, lr, lw, mw;
    if (!get("field", &f) || !get("log_n", &ln) || !get("batch", &b) || !get("layout", &lay) ||
        !get("max_log_t", &mlt) || !get("logr", &lr) || !get("log_w", &lw) || !get("mid_log_w", &mw))
      return c.fail(SFN_ERR_INVALID_ARGUMENT, "tuning entry misses a key");
    if (mlt < 1 || mlt > sfn::kMaxLogTile || lr < 1 || lr > 5 || lw < 0 || lw > 8 || mw < 0 || mw > 8)
      return c.fail(SFN_ERR_INVALID_ARGUMENT, "tuning entry out of range");
    Schedule s;
    s.max_log_t = (uint32_t)mlt;
    s.logr = (uint32_t)lr;
    s.log_w = (uint32_t)lw;
    s.mid_log_w = (uint32_t)mw;
    add[sfn::TuneKey{(int)f, (uint32_t)ln, (uint32_t)b, (int)lay}] = s;
    p = e + 1;
  }
  std::lock_guard<std::mutex> lk(c.tune_mu);
  for (auto &kv : add) c.tuned[kv.first] = kv.second;
  return SFN_OK;
}

}  // extern "C"
