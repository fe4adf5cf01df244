// Host runtime behind the C ABI: context, device-resident tables, planner,
// executor, and the host-pointer entry points.
#include "runtime.hpp"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cstdio>

#include "tables.hpp"

namespace sfn {

// ------------------------------------------------------------------ errors
sfn_status Context::fail(sfn_status s, const std::string &msg) {
  std::lock_guard<std::mutex> lk(err_mu);
  last_error = msg;
  return s;
}

void Context::clear_error() {
  std::lock_guard<std::mutex> lk(err_mu);
  last_error.clear();
}

sfn_status Context::hip_fail(hipError_t e, const char *what) {
  std::string msg = std::string(what) + ": " + hipGetErrorString(e);
  {
    std::lock_guard<std::mutex> lk(err_mu);
    last_error.swap(msg);
  }
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SFN_ERR_BUFFER_ALLOC_FAILED;
  return SFN_ERR_GPU_EXECUTION;
}

#define SFN_HIP(call, what)                              \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return hip_fail(e_, what);     \
  } while (0)

Context::~Context() {
  auto rel = [](DevBuf &b) {
    if (b.ptr) (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  };
  for (auto &f : mf_pos)
    for (auto &b : f) rel(b);
  for (auto &b : retired) rel(b);
  rel(bb_small_pair);
  for (auto &b : mf_root) rel(b);
  for (auto &b : mf_pow) rel(b);
  for (auto &kv : tables) {
    rel(kv.second.buf[0]);
    rel(kv.second.buf[1]);
  }
  rel(scratch_in);
  rel(scratch_out);
  rel(scratch_pp);
  rel(scratch_lde);
  if (scratch_ev) (void)hipEventDestroy(scratch_ev);
  for (auto e : tev) (void)hipEventDestroy(e);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  chunks.release();
  if (stream) (void)hipStreamDestroy(stream);
}

namespace {
thread_local const NoAllocScope *t_noalloc = nullptr;
}

NoAllocScope::NoAllocScope(uint32_t *s, size_t b, bool d) : scratch(s), bytes(b), dry(d), prev_(t_noalloc) {
  t_noalloc = this;
}
NoAllocScope::~NoAllocScope() { t_noalloc = prev_; }
const NoAllocScope *NoAllocScope::current() { return t_noalloc; }

// a dry run plans the call and looks every table up, but launches nothing
static bool skip_launch() {
  const NoAllocScope *na = NoAllocScope::current();
  return na && na->dry;
}

static sfn_status not_resident(Context *c, const char *what) {
  return c->fail(SFN_ERR_NOT_READY, std::string(what) +
                                        " is not resident and this entry point allocates nothing: run the allocating "
                                        "entry point once for this shape first");
}

sfn_status Context::acquire_lde_scratch(size_t bytes, hipStream_t s, uint32_t **out) {
  if (const NoAllocScope *na = NoAllocScope::current()) {
    // the library's static scratch: calls that overlap on one device race on it
    if (bytes > na->bytes) return fail(SFN_ERR_INVALID_SIZE, "LDE intermediate exceeds the static scratch");
    *out = na->scratch;
    return SFN_OK;
  }
  if (!scratch_ev && hipEventCreateWithFlags(&scratch_ev, hipEventDisableTiming) != hipSuccess) {
    scratch_ev = nullptr;
    return fail(SFN_ERR_GPU_EXECUTION, "scratch event");
  }
  if (scratch_lde.bytes < bytes && scratch_ev_used) (void)hipEventSynchronize(scratch_ev);
  sfn_status st = ensure(scratch_lde, bytes);
  if (st) return st;
  if (scratch_ev_used) (void)hipStreamWaitEvent(s, scratch_ev, 0);
  *out = (uint32_t *)scratch_lde.ptr;
  return SFN_OK;
}

void Context::release_lde_scratch(hipStream_t s) {
  if (NoAllocScope::current()) return;
  if (hipEventRecord(scratch_ev, s) == hipSuccess) scratch_ev_used = true;
}

sfn_status Context::ensure(DevBuf &b, size_t bytes) {
  if (b.bytes >= bytes && b.ptr) return SFN_OK;
  if (NoAllocScope::current()) return not_resident(this, "a scratch buffer of this size");
  if (b.ptr) {
    // kernels on any stream may still read the old buffer (tables are shared
    // by every stream of the context); growth is rare, so wait for the device
    (void)hipDeviceSynchronize();
    (void)hipFree(b.ptr);
    b.ptr = nullptr;
    b.bytes = 0;
  }
  if (bytes == 0) return SFN_OK;
  hipError_t e = hipMalloc(&b.ptr, bytes);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    (void)hipGetLastError();
    return fail(SFN_ERR_BUFFER_ALLOC_FAILED, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  b.bytes = bytes;
  return SFN_OK;
}

sfn_status Context::upload_table(DevBuf &b, const void *host, size_t bytes) {
  sfn_status s = ensure(b, bytes);
  if (s) return s;
  hipError_t e = hipMemcpy(b.ptr, host, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(b.ptr);  // nothing has read it yet
    b = DevBuf{};
    return hip_fail(e, "table upload");
  }
  table_bytes += bytes;
  return SFN_OK;
}

void Context::release_table(DevBuf &b) {
  if (!b.ptr) return;
  (void)hipFree(b.ptr);
  table_bytes -= b.bytes;
  b = DevBuf{};
}

// ------------------------------------------------------------------ table cache
namespace {
thread_local TablePins *t_pins = nullptr;
}

TablePins::TablePins(Context *ctx) : ctx_(ctx), prev_(t_pins) { t_pins = this; }

TablePins::~TablePins() {
  t_pins = prev_;
  if (!keys_.empty()) ctx_->unpin(keys_);
}

std::vector<TableKey> TablePins::take() {
  std::vector<TableKey> k;
  k.swap(keys_);
  return k;
}

void Context::unpin(const std::vector<TableKey> &keys) {
  std::lock_guard<std::recursive_mutex> g(mu);
  for (const TableKey &k : keys) {
    auto it = tables.find(k);
    if (it != tables.end() && it->second.pins) --it->second.pins;
  }
}

void Context::evict_for(size_t incoming) {
  std::lock_guard<std::recursive_mutex> g(mu);
  std::vector<std::map<TableKey, TableEntry>::iterator> victims;
  size_t keep = cached_bytes;
  while (keep + incoming > table_limit) {
    auto lru = tables.end();
    for (auto it = tables.begin(); it != tables.end(); ++it) {
      if (it->second.pins || std::find(victims.begin(), victims.end(), it) != victims.end()) continue;
      if (lru == tables.end() || it->second.last_use < lru->second.last_use) lru = it;
    }
    if (lru == tables.end()) break;  // everything left is pinned
    victims.push_back(lru);
    keep -= lru->second.bytes;
  }
  if (victims.empty()) return;
  // kernels already issued on any stream may still read these tables
  (void)hipDeviceSynchronize();
  for (auto it : victims) {
    release_table(it->second.buf[0]);
    release_table(it->second.buf[1]);
    cached_bytes -= it->second.bytes;
    tables.erase(it);
    ++evictions;
  }
}

sfn_status Context::cached_table(const TableKey &k, TableBuilder build, const void **p0, const void **p1) {
  std::lock_guard<std::recursive_mutex> g(mu);
  auto it = tables.find(k);
  if (it == tables.end() && NoAllocScope::current()) return not_resident(this, "a twiddle table of this shape");
  if (it == tables.end()) {
    std::vector<uint32_t> h0, h1;
    build(k, h0, h1);
    const size_t need = (h0.size() + h1.size()) * 4;
    evict_for(need);
    TableEntry e;
    sfn_status s = upload_table(e.buf[0], h0.data(), h0.size() * 4);
    if (!s && !h1.empty()) s = upload_table(e.buf[1], h1.data(), h1.size() * 4);
    if (s) {
      release_table(e.buf[0]);
      release_table(e.buf[1]);
      return s;
    }
    e.bytes = need;
    cached_bytes += need;
    it = tables.emplace(k, e).first;
  }
  it->second.last_use = ++table_clock;
  if (t_pins && t_pins->ctx() == this) {
    ++it->second.pins;
    t_pins->add(k);
  }
  *p0 = it->second.buf[0].ptr;
  if (p1) *p1 = it->second.buf[1].ptr;
  return SFN_OK;
}

sfn_status Context::set_table_limit(size_t bytes) {
  std::lock_guard<std::recursive_mutex> g(mu);
  table_limit = bytes ? bytes : kDefaultTableLimit;
  evict_for(0);
  return SFN_OK;
}

namespace {
// Shoup pairs as consecutive (w, q) words
void pair_words(const std::vector<host::Pair> &v, std::vector<uint32_t> &out) {
  static_assert(sizeof(host::Pair) == 8, "two words per pair");
  out.resize(v.size() * 2);
  if (!v.empty()) memcpy(out.data(), v.data(), v.size() * 8);
}
}  // namespace

hipError_t Context::timing_mark(hipStream_t s, bool before, const std::string &label) {
  if (!pass_timing || s != stream || NoAllocScope::current()) return hipSuccess;
  const size_t need = 2 * (size_t)(tcount + 1);
  while (tev.size() < need) {
    hipEvent_t e;
    hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    tev.push_back(e);
  }
  if (before) {
    if (tlabel.size() <= tcount) tlabel.resize(tcount + 1);
    tlabel[tcount] = label;
    return hipEventRecord(tev[2 * tcount], s);
  }
  hipError_t r = hipEventRecord(tev[2 * tcount + 1], s);
  ++tcount;
  return r;
}

// ------------------------------------------------------------------ tables
sfn_status Context::position_table(int field, uint32_t log_max, const uint2 **fwd, const uint2 **inv) {
  std::lock_guard<std::recursive_mutex> g(mu);
  const int fi = mf_index(field);
  if (log_max < kMaxLogTile) log_max = kMaxLogTile;
  if (mf_pos_log[fi] < log_max && NoAllocScope::current()) return not_resident(this, "a position table");
  if (mf_pos_log[fi] < log_max) {
    for (int d = 0; d < 2; ++d) {
      auto t = host::position_table(host::monty_field(field), log_max, d == 1);
      // another thread may hold the smaller table's address for a launch it has
      // not issued yet: keep that buffer until the context is destroyed
      if (mf_pos[fi][d].ptr) {
        retired.push_back(mf_pos[fi][d]);
        mf_pos[fi][d] = DevBuf{};
      }
      sfn_status s = upload_table(mf_pos[fi][d], t.data(), t.size() * sizeof(host::Pair));
      if (s) return s;
    }
    mf_pos_log[fi] = log_max;
  }
  if (fwd) *fwd = (const uint2 *)mf_pos[fi][0].ptr;
  if (inv) *inv = (const uint2 *)mf_pos[fi][1].ptr;
  return SFN_OK;
}

sfn_status Context::bb_small_pair_table(const uint2 **out) {
  std::lock_guard<std::recursive_mutex> g(mu);
  if (!bb_small_pair.ptr && NoAllocScope::current()) return not_resident(this, "the small twiddle table");
  if (!bb_small_pair.ptr) {
    auto f = host::position_table(host::kBabyBear, kMaxLogTile, false);
    auto i = host::position_table(host::kBabyBear, kMaxLogTile, true);
    std::vector<host::Pair> both(i);
    both.insert(both.end(), f.begin(), f.end());
    sfn_status s = upload_table(bb_small_pair, both.data(), both.size() * sizeof(host::Pair));
    if (s) return s;
  }
  *out = (const uint2 *)bb_small_pair.ptr;
  return SFN_OK;
}

sfn_status Context::root_table(int field, const uint2 **out) {
  std::lock_guard<std::recursive_mutex> g(mu);
  DevBuf &r = mf_root[mf_index(field)];
  if (!r.ptr && NoAllocScope::current()) return not_resident(this, "the root table");
  if (!r.ptr) {
    auto t = host::root_table(host::monty_field(field), kMaxLogTile);
    sfn_status s = upload_table(r, t.data(), t.size() * sizeof(host::Pair));
    if (s) return s;
  }
  *out = (const uint2 *)r.ptr;
  return SFN_OK;
}

sfn_status Context::pow_table(int field, const uint32_t **out) {
  std::lock_guard<std::recursive_mutex> g(mu);
  DevBuf &r = mf_pow[mf_index(field)];
  if (!r.ptr && NoAllocScope::current()) return not_resident(this, "the root power table");
  if (!r.ptr) {
    auto t = host::root_powers_m(host::monty_field(field), kPowLoBits, kPowWords);
    if (t.empty()) return fail(SFN_ERR_INVALID_ARGUMENT, "root power table does not fit");
    sfn_status s = upload_table(r, t.data(), t.size() * sizeof(uint32_t));
    if (s) return s;
  }
  *out = (const uint32_t *)r.ptr;
  return SFN_OK;
}

sfn_status Context::twist_m_table(int field, uint32_t lo, uint32_t hi, bool inverse, bool low_major,
                                  const uint32_t **out) {
  const TableKey k{kTwistMont, (uint32_t)field, lo, hi, (uint32_t)inverse, (uint32_t)low_major};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &) {
        const host::Monty &f = host::monty_field((int)k.a);
        auto t = host::to_monty_words(f, host::twist_table(f, k.b, k.c, k.d != 0));
        if (k.e) t = host::low_major(t, k.b, k.c);
        h0 = std::move(t);
      },
      (const void **)out);
}

sfn_status Context::bb_lde_fast(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo, bool low_major,
                                const uint32_t **post, const uint32_t **pointf) {
  const TableKey k{kLdeFast, log_n, added_bits, shift_std, lo, (uint32_t)low_major};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &h1) {
        std::vector<uint32_t> po, pf;
        host::bb_lde_fast_tables(k.a, k.b, k.c, k.d, po, pf);
        if (k.e) {  // every block's post table in (low, p) order
          const size_t n = (size_t)1 << k.a;
          for (size_t b = 0; b < po.size() / n; ++b) {
            std::vector<uint32_t> blk(po.begin() + b * n, po.begin() + (b + 1) * n);
            blk = host::low_major(blk, k.d, k.a);
            std::copy(blk.begin(), blk.end(), po.begin() + b * n);
          }
        }
        h0 = std::move(po);
        h1 = std::move(pf);
      },
      (const void **)post, (const void **)pointf);
}

sfn_status Context::bb_lde_lane(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo,
                                const uint32_t **out) {
  const TableKey k{kLdeLane, log_n, added_bits, shift_std, lo, 0};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &) {
        h0 = host::bb_lde_lane_factors(k.a, k.b, k.c, k.d);
      },
      (const void **)out);
}

sfn_status Context::twist_table(int field, uint32_t lo, uint32_t hi, bool inverse, const uint2 **out) {
  const TableKey k{kTwistPair, (uint32_t)field, lo, hi, (uint32_t)inverse, 0};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &) {
        pair_words(host::twist_table(host::monty_field((int)k.a), k.b, k.c, k.d != 0), h0);
      },
      (const void **)out);
}

sfn_status Context::bb_mid_table(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, const uint2 **out) {
  const TableKey k{kLdeMid, log_n, added_bits, shift_std, 0, 0};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &) {
        pair_words(host::bb_lde_mid_table(k.a, k.b, k.c), h0);
      },
      (const void **)out);
}

sfn_status Context::m31_table(uint32_t log_n, bool inverse, const uint32_t **out) {
  const TableKey k{kM31Block, log_n, (uint32_t)inverse, 0, 0, 0};
  return cached_table(
      k,
      [](const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &) {
        h0 = host::m31_block_table(k.a, k.b != 0);
      },
      (const void **)out);
}

// ------------------------------------------------------------------ validation
// Mirrors the reference's checks: power-of-two sizes via log_n, len == n*batch
// (implicit in the ABI), log_n <= 30 (src/ntt/metal_r2.rs:78) and BabyBear's
// two-adicity 27 (src/field/babybear.rs:70); KoalaBear's two-adicity is 24.
sfn_status validate_ntt(Context *ctx, int field, int variant, int dir, uint32_t log_n, uint32_t batch,
                        int layout) {
  if (field != SFN_FIELD_M31 && field != SFN_FIELD_BABYBEAR && field != SFN_FIELD_KOALABEAR)
    return ctx->fail(SFN_ERR_INVALID_ARGUMENT, "unknown field");
  if (variant < SFN_CT_DIT_R2 || variant > SFN_VARIANT_AUTO)
    return ctx->fail(SFN_ERR_INVALID_ARGUMENT, "unknown variant");
  if (dir != SFN_FORWARD && dir != SFN_INVERSE) return ctx->fail(SFN_ERR_INVALID_ARGUMENT, "unknown direction");
  if (layout != SFN_LAYOUT_SOA && layout != SFN_LAYOUT_ROW_MAJOR)
    return ctx->fail(SFN_ERR_INVALID_ARGUMENT, "unknown layout");
  if (batch == 0) return ctx->fail(SFN_ERR_INVALID_SIZE, "batch must be >= 1");
  const uint32_t max_log = field == SFN_FIELD_M31 ? 30 : host::monty_field(field).max_log;
  if (log_n > max_log) return ctx->fail(SFN_ERR_INVALID_SIZE, "log_n " + std::to_string(log_n) + " exceeds the field limit");
  if (((uint64_t)batch << log_n) > (1ull << 34)) return ctx->fail(SFN_ERR_INVALID_SIZE, "transform too large");
  return SFN_OK;
}

bool ranges_overlap(const void *a, size_t a_bytes, const void *b, size_t b_bytes) {
  const uintptr_t a0 = (uintptr_t)a, b0 = (uintptr_t)b;
  return a_bytes && b_bytes && a0 < b0 + b_bytes && b0 < a0 + a_bytes;
}

sfn_status validate_lde(Context *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits, int layout) {
  if (layout != SFN_LAYOUT_SOA && layout != SFN_LAYOUT_ROW_MAJOR)
    return ctx->fail(SFN_ERR_INVALID_ARGUMENT, "unknown layout");
  if (batch == 0) return ctx->fail(SFN_ERR_INVALID_SIZE, "batch must be >= 1");
  // shift = two_adic_generator(log_n + added_bits + 1) needs <= 27 (src/lde/mod.rs:59)
  if (log_n + added_bits + 1 > 27) return ctx->fail(SFN_ERR_INVALID_SIZE, "log_n + added_bits + 1 exceeds 27");
  if (((uint64_t)batch << (log_n + added_bits)) > (1ull << 34)) return ctx->fail(SFN_ERR_INVALID_SIZE, "LDE too large");
  return SFN_OK;
}

// ------------------------------------------------------------------ planning
// AUTO schedule of one single-column call with no tuning entry: the
// autotuner's choices on MI355X, the same for every field (back-to-back calls
// 1.3-2.2x faster than the batch rule at 2^5..2^20): one pass up to 2^11
// points (radix 8 below 2^7), three radix-8 passes of <= 2^6 at 2^13..2^14,
// two radix-8 passes at 2^15..2^16, passes of up to 2^11 with 8 lanes at
// 2^17..2^20.  2^12 and other sizes keep the batch rule.
Schedule Context::schedule_for_column(int field, uint32_t log_n, int layout) const {
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tuned.find(TuneKey{field, log_n, 1, layout});
    if (it != tuned.end()) return it->second;
  }
  Schedule s;
  if (log_n <= 11) {
    s.max_log_t = 11;
    s.logr = log_n <= 6 ? 3 : 5;
  } else if (log_n >= 13 && log_n <= 14) {
    s.max_log_t = 6;
    s.logr = 3;
    s.log_w = 3;
  } else if (log_n >= 15 && log_n <= 16) {
    s.max_log_t = 10;
    s.logr = 3;
    s.log_w = 2;
  } else if (log_n >= 17 && log_n <= 20) {
    s.max_log_t = 11;
    s.log_w = 3;
  } else {
    s = schedule_for(field, log_n, 1, layout);
  }
  return s;
}

// AUTO schedule of one batched call (batch > 1, at most 2^22 elements) with
// no tuning entry: the autotuner's MI355X choices for BabyBear and M31 at
// 2^8..2^20 x 4..256 columns (back-to-back calls 1.3-3.3x faster than the
// batch rule, which keeps >= ~1024 workgroups per pass and so runs more, smaller
// passes than these launch-bound sizes want).  2^12 and SoA 2^13..2^14 keep the
// batch rule, as do larger calls.
Schedule Context::schedule_for_batch(int field, uint32_t log_n, uint32_t batch, int layout) const {
  const uint64_t elems = (uint64_t)batch << log_n;
  const bool rm = layout == SFN_LAYOUT_ROW_MAJOR;
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tuned.find(TuneKey{field, log_n, batch, layout});
    if (it != tuned.end()) return it->second;
  }
  if (elems > (1ull << 22) || log_n == 12 || (!rm && (log_n == 13 || log_n == 14)))
    return schedule_for(field, log_n, batch, layout);
  Schedule s;
  if (log_n <= 11) {  // one pass
    s.max_log_t = 11;
    s.logr = rm && log_n <= 9 ? 3 : 5;
    s.log_w = rm ? 2 : 3;
  } else if (log_n <= 16) {  // two passes
    s.max_log_t = 10;
    const bool few = elems <= (1ull << 18);
    s.logr = few ? 3 : 5;
    s.log_w = few ? 2 : 4;
  } else {
    s.max_log_t = 11;
    s.log_w = 4;
  }
  return s;
}

// Coset-LDE schedule (2x and more) of a call with no tuning entry: the
// autotuner's MI355X choices at 2^8..2^20 x 1..4096 columns.  Calls of at most
// 2^23 input elements use sub-transforms of up to 2^11 points, so the middle
// pass takes half the layers and each outer side one pass (back-to-back calls
// 1.1-4.8x faster than the batch rule, whose small tiles split these sizes
// into five or more launches); SoA 2^9..2^10 prefers radix-8 outer passes, and
// 2^8 (but row-major batches of >= 8 columns) and smaller keep the batch rule.
// The middle pass takes 32 lanes for tiles of <= 2^8 points and 16 for 2^9
// and 2^10 (larger calls 1.07-2.4x faster than 8); at 2^10 points 16 lanes
// are two workgroups of 8 waves per CU at <= 128 VGPRs, four waves per SIMD
// (2^20 x 256 row-major: middle pass 1797 -> 1492 us against 8 lanes).
Schedule Context::schedule_for_lde(uint32_t log_n, uint32_t batch, int layout) const {
  const uint64_t elems = (uint64_t)batch << log_n;
  const bool rm = layout == SFN_LAYOUT_ROW_MAJOR;
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tuned.find(TuneKey{SFN_TUNE_LDE, log_n, batch, layout});
    if (it != tuned.end()) return it->second;
  }
  Schedule s;
  if (elems > (1ull << 23) || log_n < 8 || (log_n == 8 && (!rm || batch < 8))) {
    s = schedule_for(SFN_TUNE_LDE, log_n, batch, layout);
  } else {
    s.max_log_t = 11;
    s.log_w = 4;
    if (!rm && log_n <= 10) {
      s.max_log_t = 10;
      s.logr = 3;
      s.log_w = 2;
    }
  }
  const uint32_t mlt = std::min<uint32_t>(s.max_log_t, kMaxLogTile);
  const uint32_t mid_t = std::min<uint32_t>(log_n <= mlt ? log_n : (log_n + 1) / 2, mlt);
  s.mid_log_w = mid_t >= 10 ? 4 : mid_t == 9 ? 4 : 5;
  return s;
}

bool Context::has_tuning(const TuneKey &k) const {
  std::lock_guard<std::mutex> lk(tune_mu);
  return tuned.count(k) != 0;
}

Schedule Context::schedule_for(int field, uint32_t log_n, uint32_t batch, int layout) const {
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto it = tuned.find(TuneKey{field, log_n, batch, layout});
    if (it != tuned.end()) return it->second;
  }
  Schedule s;
  // keep at least ~1024 workgroups per pass: small single columns get smaller tiles
  const uint64_t elems = (uint64_t)batch << log_n;
  if (elems < (1ull << 21)) s.max_log_t = 7;
  else if (elems < (1ull << 24)) s.max_log_t = 8;
  return s;
}

namespace {

struct Layout {
  uint64_t cs, rs;
};

inline Layout layout_of(int layout, uint32_t log_rows, uint32_t batch) {
  if (layout == SFN_LAYOUT_SOA) return Layout{(uint64_t)1 << log_rows, 1};
  return Layout{1, batch};
}

// Splits [0, log_n) into passes of at most max_log_t layers, as even as possible.
std::vector<std::pair<uint32_t, uint32_t>> split_layers(uint32_t log_n, uint32_t max_log_t) {
  std::vector<std::pair<uint32_t, uint32_t>> r;  // (lo, log_t), ascending lo
  if (log_n == 0) return r;
  const uint32_t np = (log_n + max_log_t - 1) / max_log_t;
  uint32_t lo = 0;
  for (uint32_t i = 0; i < np; ++i) {
    const uint32_t lt = (log_n - lo) / (np - i) + (((log_n - lo) % (np - i)) ? 1 : 0);
    r.emplace_back(lo, lt);
    lo += lt;
  }
  return r;
}

// Lane geometry of one pass over layers [lo, lo + log_t).
void choose_lanes(PassDesc &p, int layout, uint32_t batch, uint32_t want_log_w) {
  const uint32_t max_log_w_lds = p.log_t >= 14 ? 0 : 14 - p.log_t;  // W * T <= 16384
  uint32_t lw = want_log_w;
  if (layout == SFN_LAYOUT_ROW_MAJOR) {
    p.order = kColFastest;
    p.lane_contig = 1;
    if (p.log_t <= 6) lw = 6;  // small tiles: more lanes
  } else if (p.lo > 0) {
    p.order = kLowFastest;
    p.lane_contig = 1;
    if (p.log_t <= 6) lw = 6;
    if (lw > p.lo) lw = p.lo;  // lanes must be adjacent rows of one column
  } else {
    p.lane_contig = 0;
    p.order = batch >= (1u << 2) ? kColFastest : kHighFastest;
    lw = p.log_t >= 10 ? 2 : (p.log_t >= 8 ? 3 : 5);
  }
  if (lw > max_log_w_lds) lw = max_log_w_lds;
  if (lw > 8) lw = 8;
  p.log_w = lw;
  const uint64_t instances = (uint64_t)batch << (p.log_n - p.log_t);
  p.tiles = (instances + (1ull << lw) - 1) >> lw;
}

// Lane geometry of a tuned pass: blockDim = W << (log_t - 5) threads, one item
// per thread in the full rounds; LDS within the per-workgroup limit.
bool choose_lanes_fast(PassDesc &p, int layout, uint32_t batch, uint32_t want_log_w) {
  if (p.log_t < 6 || p.log_t > (int)kMaxLogTile) return false;
  uint32_t lw;
  // lane orders whose addresses grow with the lane index, so lane 0 bounds the tile
  if (layout == SFN_LAYOUT_ROW_MAJOR) {
    p.order = kColFastest;
    p.lane_contig = 1;
    lw = want_log_w;
    while (lw > 0 && (batch & ((1u << lw) - 1))) --lw;  // W divides the row
  } else if (p.lo > 0) {
    p.order = kLowFastest;
    p.lane_contig = 1;
    lw = std::min<uint32_t>(want_log_w, p.lo);
  } else {
    p.lane_contig = 0;
    p.order = kHighFastest;
    lw = p.log_t >= 13 ? 0 : 13 - p.log_t;  // 256 threads
  }
  while (lw > 0 && (((1u << lw) << (p.log_t - 5)) > 512 || fast_lds_bytes(1, p.log_t, lw) > kMaxLdsBytes)) --lw;
  if (fast_lds_bytes(1, p.log_t, lw) > kMaxLdsBytes) return false;
  p.log_w = lw;
  p.fast = true;
  const uint64_t instances = (uint64_t)batch << (p.log_n - p.log_t);
  p.tiles = (instances + (1ull << lw) - 1) >> lw;
  return true;
}

// Tuned passes whose lanes are adjacent elements: a tile row spans 4 * W
// bytes, so 128 / (4 * W) consecutive tiles share every 128-byte line; eight
// consecutive tiles go to one XCD (with 16 lanes, 512 contiguous bytes of
// every row).  Measured on the 2^20 x 256 LDE against the 128 / (4 * W)
// tiles that share a line: outer passes 1071 -> 968 us (forward) and
// 525 -> 506 us (inverse); four tiles were slower than either.
uint32_t xcd_share_log(const Context *ctx, const PassDesc &p) {
  if (!ctx->xcd_remap || !p.lane_contig || p.log_w >= 5) return 0;
  return 3;
}

}  // namespace

std::vector<PassDesc> plan_ntt_passes(int field, int variant, int dir, uint32_t log_n, uint32_t batch, int layout,
                                      const Schedule &sc, bool allow_fast) {
  (void)field;
  std::vector<PassDesc> passes;
  const bool fwd = dir == SFN_FORWARD;
  if (variant == SFN_CT_DIT_R2) {
    for (uint32_t k = 0; k < log_n; ++k) {
      PassDesc p;
      p.layer_kernel = true;
      p.layer = fwd ? log_n - 1 - k : k;
      p.log_n = log_n;
      passes.push_back(p);
    }
  } else {
    uint32_t logr = sc.logr;
    if (variant == SFN_CT_GS_R2 || variant == SFN_STOCKHAM_R2) logr = 1;
    if (variant == SFN_CT_GS_R4) logr = 2;
    uint32_t mlt = sc.max_log_t;
    if (mlt > kMaxLogTile) mlt = kMaxLogTile;
    auto sp = split_layers(log_n, mlt);
    if (fwd) std::reverse(sp.begin(), sp.end());  // forward runs the high layers first
    for (auto &pr : sp) {
      PassDesc p;
      p.lo = pr.first;
      p.log_t = (int)pr.second;
      p.logr = (int)std::min<uint32_t>(logr, pr.second);
      p.log_n = log_n;
      if (!(allow_fast && logr == 5 && choose_lanes_fast(p, layout, batch, sc.log_w)))
        choose_lanes(p, layout, batch, sc.log_w);
      passes.push_back(p);
    }
  }
  if (!fwd && !passes.empty()) passes.back().scale = true;  // 1/n folded into the last inverse pass
  return passes;
}

// 1/n of the inverse transform: M31 canonical; Montgomery fields as a Shoup
// pair (w, floor(w * 2^32 / p)) of the standard value, which the kernels'
// Shoup multiply applies to Montgomery-form data.
static void scale_factor(int field, uint64_t n, uint32_t *w, uint32_t *q) {
  if (field == SFN_FIELD_M31) {
    *w = host::invmod((uint32_t)(n % host::kM31P), host::kM31P);
    *q = 0;
    return;
  }
  const uint32_t p = host::monty_field(field).p;
  const host::Pair pr = host::shoup(host::invmod((uint32_t)(n % p), p), p);
  *w = pr.w;
  *q = pr.q;
}

static const char *field_tag(int field) {
  return field == SFN_FIELD_M31 ? "m31," : (field == SFN_FIELD_KOALABEAR ? "kb," : "bb,");
}

// ceil(2^64 / batch) for the kernels' lane decoding (PassArgs::batch_recip), or 0
// where it does not apply (one column, or lane indices past 32 bits)
uint64_t batch_reciprocal(uint32_t batch, uint64_t instances) {
  if (batch < 2 || instances > (1ull << 32)) return 0;
  return ~0ull / batch + 1;
}

sfn_status make_pass_args(Context *ctx, int field, int dir, const PassDesc &p, const uint32_t *src, uint32_t *dst,
                          uint64_t src_cs, uint64_t src_rs, uint64_t dst_cs, uint64_t dst_rs, uint32_t batch,
                          uint32_t log_n_real, PassArgs *out) {
  PassArgs &a = *out;
  a = PassArgs{};
  a.src = src;
  a.dst = dst;
  a.src_cs = src_cs;
  a.src_rs = src_rs;
  a.dst_cs = dst_cs;
  a.dst_rs = dst_rs;
  a.log_n = p.log_n;
  a.lo = p.lo;
  a.batch = batch;
  a.log_w = p.log_w;
  a.order = p.order;
  a.lane_contig = p.lane_contig;
  a.instances = (uint64_t)batch << (p.log_n - p.log_t);
  a.batch_recip = batch_reciprocal(batch, a.instances);
  a.src_rev_bits = p.src_rev;
  a.dst_rev_bits = p.dst_rev;
  if (field == SFN_FIELD_M31) {
    const uint32_t *t;
    sfn_status st = ctx->m31_table(p.log_n, dir == 1, &t);
    if (st) return st;
    a.tw = t;
  }
  if (p.scale) {
    a.has_scale = 1;
    scale_factor(field, 1ull << log_n_real, &a.scale_w, &a.scale_q);
  }
  return SFN_OK;
}

// Generic-kernel tables of a Montgomery-field pass (after make_pass_args): the
// position table and, for an outer pass, the four-step twist as Shoup pairs.
// Only the generic kernels read them, so tuned passes never build them.
static sfn_status fill_generic_args(Context *ctx, int field, int dir, const PassDesc &p, PassArgs *out) {
  if (field == SFN_FIELD_M31) return SFN_OK;
  PassArgs &a = *out;
  const uint2 *f, *i;
  sfn_status st = ctx->position_table(field, kMaxLogTile, &f, &i);
  if (st) return st;
  a.tw = dir == 0 ? f : i;
  if (p.lo > 0) {
    const uint2 *tw;
    if ((st = ctx->twist_table(field, p.lo, p.lo + p.log_t, dir == 1, &tw))) return st;
    a.twist = tw;
  }
  return SFN_OK;
}

// Tuned kernels address a workgroup's tile with 32-bit buffer offsets from its
// lane 0: the point extent plus the lane spread must stay below 4 GiB.
static bool fast_offsets_fit(const PassDesc &p, Layout ls, Layout ld) {
  const uint64_t T = 1ull << p.log_t, W = 1ull << p.log_w;
  auto span = [&](Layout L, uint32_t rev) {
    // bit-reversed rows: a lane's points spread over every row of the array
    const uint64_t pts = rev ? (1ull << rev) * L.rs : (T << p.lo) * L.rs;
    uint64_t lane;
    if (p.order == kLowFastest) lane = W * L.rs;
    else if (p.order == kColFastest) lane = W * std::max(L.cs, L.rs);
    else lane = W * ((T << p.lo) * L.rs + L.cs);
    return (pts + lane + 1) * 4;
  };
  return span(ls, p.src_rev) < 0xFFFFFFFFull && span(ld, p.dst_rev) < 0xFFFFFFFFull;
}

// SFNTT_TWIST_GEN=0 reads every four-step twist from its table (A/B measurement).
static bool twist_gen_on() {
  static const bool on = [] {
    const char *e = getenv("SFNTT_TWIST_GEN");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Tuned-kernel fields of a pass's argument block (after make_pass_args):
// buffer extents, and for the Montgomery fields the root table and the
// Montgomery-word twist.
static sfn_status fill_fast_args(Context *ctx, int field, int dir, const PassDesc &p, uint64_t src_elems,
                                 uint64_t dst_elems, PassArgs *out) {
  PassArgs &a = *out;
  a.src_elems = src_elems;
  a.dst_elems = dst_elems;
  if (field == SFN_FIELD_M31) return SFN_OK;
  const uint2 *root;
  sfn_status st = ctx->root_table(field, &root);
  if (st) return st;
  a.tw = root;
  a.twist = nullptr;
  if (p.lo > 0) {
    // lanes of one workgroup share their low row bits when columns run fastest:
    // their twist reads are broadcasts of one table word.  Otherwise every
    // point reads its own word, 4 bytes beside the 8 it moves, and the kernels
    // generate the twist instead.
    a.tab_low_major = p.order == kColFastest;
    if (!a.tab_low_major && twist_gen_on()) {
      const uint32_t *pw;
      if ((st = ctx->pow_table(field, &pw))) return st;
      a.pow_m = pw;
      return SFN_OK;
    }
    const uint32_t *tm;
    if ((st = ctx->twist_m_table(field, p.lo, p.lo + p.log_t, dir == 1, a.tab_low_major != 0, &tm))) return st;
    a.twist_m = tm;
  }
  return SFN_OK;
}

static sfn_status launch_desc(Context *ctx, int field, int dir, const PassDesc &p, const uint32_t *src,
                              uint32_t *dst, Layout ls, Layout ld, uint32_t batch, uint32_t log_n_real,
                              hipStream_t s, uint64_t src_elems = 0, uint64_t dst_elems = 0) {
  if (p.layer_kernel) {
    LayerArgs a{};
    a.src = src;
    a.dst = dst;
    a.src_cs = ls.cs;
    a.src_rs = ls.rs;
    a.dst_cs = ld.cs;
    a.dst_rs = ld.rs;
    a.log_n = p.log_n;
    a.layer = p.layer;
    a.batch = batch;
    a.butterflies = (uint64_t)batch << (p.log_n - 1);
    if (field != SFN_FIELD_M31) {
      const uint2 *f, *i;
      sfn_status st = ctx->position_table(field, p.log_n, &f, &i);
      if (st) return st;
      a.tw = dir == 0 ? f : i;
    } else {
      const uint32_t *t;
      sfn_status st = ctx->m31_table(p.log_n, dir == 1, &t);
      if (st) return st;
      a.tw = t;
    }
    a.has_scale = 0;
    if (p.scale) {
      a.has_scale = 1;
      scale_factor(field, 1ull << log_n_real, &a.scale_w, &a.scale_q);
    }
    (void)ctx->timing_mark(s, true, "k_layer");
    hipError_t e = skip_launch() ? hipSuccess : launch_layer(field, dir, a, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "layer kernel launch");
    (void)ctx->timing_mark(s, false, "");
    return SFN_OK;
  }
  PassArgs a{};
  sfn_status st0 = make_pass_args(ctx, field, dir, p, src, dst, ls.cs, ls.rs, ld.cs, ld.rs, batch, log_n_real, &a);
  if (st0) return st0;
  // the point-adjacent tuned kernel moves 16-byte vectors: needs aligned buffers
  const bool aligned = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % 16 == 0);
  if (p.fast && ctx->allow_fast && (p.lane_contig || aligned) && src_elems && dst_elems &&
      (!(p.src_rev || p.dst_rev) || (p.lane_contig && p.lo == 0)) &&
      fast_offsets_fit(p, ls, ld)) {
    a.xcd_share_log = xcd_share_log(ctx, p);
    if ((st0 = fill_fast_args(ctx, field, dir, p, src_elems, dst_elems, &a))) return st0;
    (void)ctx->timing_mark(s, true, std::string("k_fast<") + field_tag(field) + (dir ? "inv," : "fwd,") +
                                       std::to_string(p.log_t) + (p.lane_contig ? ",lanes" : ",points") +
                                       ">lo=" + std::to_string(p.lo));
    hipError_t e = skip_launch() ? hipSuccess : launch_fast(field, dir, p.log_t, p.lane_contig != 0, a, p.tiles, s);
    if (e != hipSuccess) return ctx->hip_fail(e, "tuned pass launch");
    (void)ctx->timing_mark(s, false, "");
    return SFN_OK;
  }
  if (p.src_rev || p.dst_rev)  // bit-reversed rows are addressed by the tuned lane-contiguous kernel only
    return ctx->fail(SFN_ERR_UNSUPPORTED, "bit-reversed row addressing needs the tuned kernel");
  if (p.fast) {  // fall back to the generic kernel with its own geometry
    PassDesc q = p;
    q.fast = false;
    choose_lanes(q, ls.cs == 1 && ls.rs != 1 ? SFN_LAYOUT_ROW_MAJOR : SFN_LAYOUT_SOA, batch, 4);
    return launch_desc(ctx, field, dir, q, src, dst, ls, ld, batch, log_n_real, s);
  }
  if ((st0 = fill_generic_args(ctx, field, dir, p, &a))) return st0;
  (void)ctx->timing_mark(s, true, std::string("k_pass<") + field_tag(field) + (dir ? "inv," : "fwd,") +
                                     std::to_string(p.log_t) + "," + std::to_string(p.logr) + ">lo=" +
                                     std::to_string(p.lo));
  hipError_t e = skip_launch() ? hipSuccess : launch_pass(field, dir, p.log_t, p.logr, a, p.tiles, s);
  if (e != hipSuccess) return ctx->hip_fail(e, "pass kernel launch");
  (void)ctx->timing_mark(s, false, "");
  return SFN_OK;
}

sfn_status Context::run_ntt(int field, int variant, int dir, uint32_t log_n, uint32_t batch, int layout,
                            const uint32_t *d_in, uint32_t *d_out, hipStream_t s, const Schedule *force) {
  TablePins pins(this);  // the tables this call's kernels read stay resident while they are issued
  const uint64_t total = (uint64_t)batch << log_n;
  if (log_n == 0) {  // identity (src/ntt/metal_r2.rs:81-83)
    if (d_in != d_out) {
      hipError_t e = skip_launch() ? hipSuccess : launch_copy(d_in, d_out, total, s);
      if (e != hipSuccess) return hip_fail(e, "copy launch");
    }
    return SFN_OK;
  }
  timing_begin_sequence(s);
  // the autotuner's schedule is the AUTO variant's; V1-V4 keep their own tiling
  Schedule sc = force ? *force
                      : (variant != SFN_VARIANT_AUTO ? schedule_for(-1, log_n, batch, layout)
                         : batch == 1                  ? schedule_for_column(field, log_n, layout)
                                                       : schedule_for_batch(field, log_n, batch, layout));
  const bool fwd = dir == SFN_FORWARD;
  std::vector<PassDesc> passes = plan_ntt_passes(field, variant, dir, log_n, batch, layout, sc, true);
  (void)fwd;
  // buffers: pass 0 reads the input; V3 alternates between out and scratch and
  // ends in out; the others work in place on out.
  const size_t bytes = total * sizeof(uint32_t);
  uint32_t *pp = nullptr;
  if (variant == SFN_STOCKHAM_R2 && passes.size() > 1) {
    sfn_status st = ensure(scratch_pp, bytes);
    if (st) return st;
    pp = (uint32_t *)scratch_pp.ptr;
  }
  const Layout L = layout_of(layout, log_n, batch);
  const uint32_t *src = d_in;
  for (size_t i = 0; i < passes.size(); ++i) {
    uint32_t *dst = d_out;
    if (pp) dst = ((passes.size() - 1 - i) % 2 == 0) ? d_out : pp;
    sfn_status st = launch_desc(this, field, dir, passes[i], src, dst, L, L, batch, log_n, s, total, total);
    if (st) return st;
    src = dst;
  }
  return SFN_OK;
}

sfn_status Context::run_lde(uint32_t log_n, uint32_t batch, uint32_t added_bits, uint32_t shift_raw,
                            int layout, const uint32_t *d_in, uint32_t *d_out, hipStream_t s,
                            const Schedule *force, bool natural) {
  TablePins pins(this);
  const uint32_t LN = log_n + added_bits;
  const uint32_t shift_std =
      shift_raw ? host::bb_from_monty(shift_raw) : host::bb_root(LN + 1);  // src/lde/mod.rs:59
  timing_begin_sequence(s);
  if (log_n == 0) {
    // one row per column: the interpolant is that value, so every output row
    // repeats it (the reference's iDFT of one point is a no-op, its shift power
    // s^0 = 1, and the forward DFT of (c, 0, .., 0) is c everywhere).  The row
    // copy with a zero source row stride broadcasts the input row.
    const Layout Lo = layout_of(layout, LN, batch);
    (void)timing_mark(s, true, "k_rows_bitrev<broadcast>");
    hipError_t e = skip_launch() ? hipSuccess : launch_rows_bitrev(d_in, d_out, LN, batch, 1, 0, Lo.cs, Lo.rs, s);
    if (e != hipSuccess) return hip_fail(e, "LDE broadcast launch");
    (void)timing_mark(s, false, "");
    return SFN_OK;
  }
  Schedule sc = force ? *force : schedule_for_lde(log_n, batch, layout);
  // middle pass covers [lo, log_n); K1 (inverse) and K3 (forward) cover [0, lo),
  // each in one or more passes of at most max_log_t layers
  uint32_t mlt = std::min<uint32_t>(sc.max_log_t, kMaxLogTile);
  uint32_t mid_t = log_n <= mlt ? log_n : (log_n + 1) / 2;
  if (mid_t > mlt) mid_t = mlt;
  const uint32_t lo = log_n - mid_t;
  const Layout Lin = layout_of(layout, log_n, batch);
  const Layout Lout = layout_of(layout, LN, batch);
  const uint64_t n_in = (uint64_t)batch << log_n, n_out = (uint64_t)batch << LN;
  sfn_status st;
  auto outer = split_layers(lo, mlt);  // ascending (lo', log_t')
  // Natural row order.  Row-major columns with tuned outer passes: the first
  // inverse pass reads the input rows bit-reversed and the last forward pass
  // writes the output rows bit-reversed, so the permutation costs no traffic;
  // the transform then runs in a scratch array, since a pass that moves rows
  // cannot work in place.  Otherwise explicit row permutations of in and out.
  uint32_t *work = d_out;
  bool fused = false;
  PassDesc f1, f3;
  if (natural && layout == SFN_LAYOUT_ROW_MAJOR && allow_fast && sc.logr == 5 && !outer.empty() &&
      outer[0].first == 0 && outer[0].second >= 6) {
    f1.lo = f3.lo = 0;
    f1.log_t = f3.log_t = (int)outer[0].second;
    f1.logr = f3.logr = 5;
    f1.log_n = log_n;
    f3.log_n = LN;
    if (choose_lanes_fast(f1, layout, batch, sc.log_w) && choose_lanes_fast(f3, layout, batch, sc.log_w)) {
      f1.src_rev = log_n;
      f3.dst_rev = LN;
      if (f1.lane_contig && f3.lane_contig && fast_offsets_fit(f1, Lin, Lout) && fast_offsets_fit(f3, Lout, Lout)) {
        if ((st = acquire_lde_scratch(n_out * 4, s, &work))) return st;
        fused = true;
      }
    }
  }
  // K1: inverse layers [0, lo) (no normalisation) into the first n rows of work
  const uint32_t *mid_src = d_in;
  Layout mid_ls = Lin;
  if (natural && !fused) {
    // natural-order input: its rows, bit-reversed, become the first n rows of
    // out, where the transform then runs in place
    (void)timing_mark(s, true, "k_rows_bitrev<in>");
    hipError_t e = skip_launch() ? hipSuccess : launch_rows_bitrev(d_in, d_out, log_n, batch, Lin.cs, Lin.rs, Lout.cs, Lout.rs, s);
    if (e != hipSuccess) return hip_fail(e, "row permutation launch");
    (void)timing_mark(s, false, "");
    mid_src = d_out;
    mid_ls = Lout;
  }
  for (size_t i = 0; i < outer.size(); ++i) {
    PassDesc p;
    p.lo = outer[i].first;
    p.log_t = (int)outer[i].second;
    p.logr = (int)std::min<uint32_t>(sc.logr, outer[i].second);
    p.log_n = log_n;
    if (!(p.logr == 5 && choose_lanes_fast(p, layout, batch, sc.log_w))) choose_lanes(p, layout, batch, sc.log_w);
    if (fused && i == 0) p = f1;
    st = launch_desc(this, SFN_FIELD_BABYBEAR, 1, p, mid_src, work, mid_ls, Lout, batch, log_n, s,
                     mid_src == d_in ? n_in : n_out, n_out);
    if (st) return st;
    mid_src = work;
    mid_ls = Lout;
  }
  // K2: inverse [lo, log_n) + n^-1 (shift*zeta_b)^j + forward [log_n-1, lo] per block
  {
    PassArgs a{};
    a.src = mid_src;
    a.dst = work;
    a.src_cs = mid_ls.cs;
    a.src_rs = mid_ls.rs;
    a.dst_cs = Lout.cs;
    a.dst_rs = Lout.rs;
    a.log_n = log_n;
    a.lo = lo;
    a.batch = batch;
    PassDesc p;
    p.lo = lo;
    p.log_t = (int)mid_t;
    p.log_n = log_n;
    choose_lanes(p, layout, batch, sc.mid_log_w);
    // the middle kernel runs W << (mid_t - k) threads (one register item each):
    // at most 1024; lanes over low row bits must stay within one column (W <= 2^lo)
    const uint32_t k = std::min<uint32_t>(5, mid_t);
    uint32_t lw = p.log_w;
    while (lw > 0 && ((1u << lw) << (mid_t - k)) > 1024) --lw;
    while (((1u << lw) << (mid_t - k)) < 256 && (p.order != kLowFastest || lw < lo) && lw < 6 &&
           ((2u << lw) << (mid_t - k)) <= 1024)
      ++lw;
    a.log_w = lw;
    a.order = p.order;
    a.lane_contig = p.lane_contig;
    a.instances = (uint64_t)batch << lo;
    a.batch_recip = batch_reciprocal(batch, a.instances);
    a.blocks = 1u << added_bits;
    a.dst_block_off = ((uint64_t)1 << log_n) * Lout.rs;
    // tuned middle kernel: lanes adjacent in memory, one item per thread
    PassDesc q;
    q.lo = lo;
    q.log_t = (int)mid_t;
    q.log_n = log_n;
    Layout mid_ld = Lout;  // per block the kernel rebases at the block start
    // lanes adjacent in memory; or, when the middle pass is the whole transform
    // (lo == 0, SoA), one column per lane with 16-byte vector accesses.  With
    // outer passes the tuned middle pass goes with their radix-32 schedule.
    const bool mid_aligned = ((uintptr_t)mid_src % 16 == 0) && ((uintptr_t)work % 16 == 0);
    if (allow_fast && (sc.logr == 5 || lo == 0) && choose_lanes_fast(q, layout, batch, sc.mid_log_w) &&
        (q.lane_contig || (lo == 0 && mid_aligned)) && fast_offsets_fit(q, mid_ls, mid_ld)) {
      const uint2 *root;
      const uint32_t *tim = nullptr, *post, *pointf;
      if ((st = root_table(SFN_FIELD_BABYBEAR, &root))) return st;
      const bool low_major = q.order == kColFastest;
      // lanes over low row bits (SoA) read a twist word per point: the kernel
      // generates the inverse twist instead (as the tuned outer passes do)
      const bool gen = q.lane_contig && !low_major && twist_gen_on();
      if (gen) {
        // and its post factors: the forward twist walked as the inverse one,
        // times the lane's n^-1 sigma_b^low
        const uint32_t *pw, *lanef;
        if ((st = pow_table(SFN_FIELD_BABYBEAR, &pw))) return st;
        if ((st = bb_lde_lane(log_n, added_bits, shift_std, lo, &lanef))) return st;
        a.pow_m = pw;
        a.lanef_m = lanef;
      } else if ((st = twist_m_table(SFN_FIELD_BABYBEAR, lo, log_n, true, low_major, &tim))) {
        return st;
      }
      if ((st = bb_lde_fast(log_n, added_bits, shift_std, lo, low_major, &post, &pointf))) return st;
      a.tab_low_major = low_major;
      a.tw = root;
      // one column per lane covers the whole transform: no four-step twist
      a.twist_m = q.lane_contig ? tim : nullptr;
      a.post_m = post;
      a.pointf_m = pointf;
      a.src_elems = mid_src == d_in ? n_in : n_out;
      a.dst_elems = n_out;
      a.log_w = q.log_w;
      a.order = q.order;
      a.lane_contig = q.lane_contig;
      a.xcd_share_log = xcd_share_log(this, q);
      (void)timing_mark(s, true, "k_mid_fast<" + std::to_string(mid_t) + (q.lane_contig ? ",lanes" : ",points") +
                                     ">lo=" + std::to_string(lo));
      hipError_t e =
          skip_launch() ? hipSuccess : launch_lde_mid_fast((int)mid_t, q.lane_contig != 0, a, q.tiles, s);
      if (e != hipSuccess) return hip_fail(e, "LDE middle pass launch");
      (void)timing_mark(s, false, "");
    } else {
      // the generic middle kernel's tables: Shoup-pair twists and block factors
      const uint2 *pair, *twi, *twf, *mid;
      if ((st = twist_table(SFN_FIELD_BABYBEAR, lo, log_n, true, &twi))) return st;
      if ((st = twist_table(SFN_FIELD_BABYBEAR, lo, log_n, false, &twf))) return st;
      if ((st = bb_mid_table(log_n, added_bits, shift_std, &mid))) return st;
      a.twist = twi;
      a.twist_fwd = twf;
      a.mid = mid;
      if ((st = bb_small_pair_table(&pair))) return st;
      a.tw = pair;
      const uint64_t tiles = (a.instances + (1ull << lw) - 1) >> lw;
      (void)timing_mark(s, true, "k_lde_mid<" + std::to_string(mid_t) + ">lo=" + std::to_string(lo));
      hipError_t e = skip_launch() ? hipSuccess : launch_lde_mid((int)mid_t, a, tiles, s);
      if (e != hipSuccess) return hip_fail(e, "LDE middle pass launch");
      (void)timing_mark(s, false, "");
    }
  }
  // K3: forward layers [lo-1 .. 0] on every output block (the N-row array)
  for (size_t i = outer.size(); i-- > 0;) {
    PassDesc p;
    p.lo = outer[i].first;
    p.log_t = (int)outer[i].second;
    p.logr = (int)std::min<uint32_t>(sc.logr, outer[i].second);
    p.log_n = LN;
    if (!(p.logr == 5 && choose_lanes_fast(p, layout, batch, sc.log_w))) choose_lanes(p, layout, batch, sc.log_w);
    if (fused && i == 0) p = f3;
    st = launch_desc(this, SFN_FIELD_BABYBEAR, 0, p, work, fused && i == 0 ? d_out : work, Lout, Lout, batch, LN, s,
                     n_out, n_out);
    if (st) return st;
  }
  if (fused) release_lde_scratch(s);
  if (natural && !fused) {  // bit-reversed output rows -> natural order, in place
    (void)timing_mark(s, true, "k_rows_bitrev<out>");
    hipError_t e = skip_launch() ? hipSuccess : launch_rows_bitrev(d_out, d_out, LN, batch, Lout.cs, Lout.rs, Lout.cs, Lout.rs, s);
    if (e != hipSuccess) return hip_fail(e, "row permutation launch");
    (void)timing_mark(s, false, "");
  }
  return SFN_OK;
}

// ------------------------------------------------------------------ multi-job launches
namespace {

// Lanes of a tuned multi-job tile (SoA columns): W * T = 8192 points, one
// 32-point item per thread of a 256-thread workgroup.
bool choose_lanes_multi(PassDesc &p, uint32_t batch) {
  if (p.log_t < 6 || p.log_t > (int)kMaxLogTile || p.logr != 5) return false;
  uint32_t lw = std::min<uint32_t>(7, 14 - (uint32_t)p.log_t);  // lane records <= 4 KiB
  const uint64_t instances = (uint64_t)batch << (p.log_n - p.log_t);
  while (lw > 0 && (1ull << (lw - 1)) >= instances) --lw;  // no wholly idle lane groups
  if (p.lo > 0) {
    p.order = kLowFastest;  // lanes: adjacent low rows of one column
    p.lane_contig = 1;
    lw = std::min<uint32_t>(lw, p.lo);
  } else {
    p.order = kHighFastest;  // lanes: contiguous sub-transforms (points adjacent)
    p.lane_contig = 0;
  }
  const size_t T = (size_t)1 << p.log_t;
  if (kMultiRootBytes + (kLaneInfoBytes << lw) + ((T + T / 32) << lw) * 4 > kMaxLdsBytes / 2) return false;
  p.log_w = lw;
  p.fast = true;
  p.tiles = (instances + (1ull << lw) - 1) >> lw;
  return true;
}

}  // namespace

sfn_status build_multi(Context *ctx, const std::vector<MultiJob> &jobs_in, std::vector<MultiLaunch> &out) {
  out.clear();
  // Consecutive jobs of one shape whose columns follow each other in memory
  // (input and output) are one batched job: their lanes then share tiles.
  std::vector<MultiJob> jobs;
  for (const MultiJob &m : jobs_in) {
    if (!jobs.empty()) {
      MultiJob &b = jobs.back();
      const uint64_t span = (uint64_t)b.batch << b.log_n;
      if (b.field == m.field && b.dir == m.dir && b.log_n == m.log_n && m.log_n > 0 &&
          (uint64_t)b.batch + m.batch <= 0xFFFFFFFFull && m.src == b.src + span && m.dst == b.dst + span &&
          ((b.src == b.dst) == (m.src == m.dst))) {
        b.batch += m.batch;
        continue;
      }
    }
    jobs.push_back(m);
  }
  std::vector<std::vector<PassDesc>> jp(jobs.size());
  size_t nl = 0;
  for (size_t j = 0; j < jobs.size(); ++j) {
    const MultiJob &m = jobs[j];
    if (m.log_n == 0) continue;
    Schedule sc = ctx->schedule_for(m.field, m.log_n, m.batch, SFN_LAYOUT_SOA);
    // Many jobs fill the device together, so passes may be large: up to 2^11
    // points, also for 2^22 columns (two 11-layer passes against three
    // shorter ones, measured on 512 uneven columns: 2.361-2.362 vs
    // 2.373-2.384 ms; with the earlier 256-thread kernel the three passes won,
    // 2.44 vs 2.50 ms).  A tuned entry keeps its pass size; the radix is
    // always the tuned kernel's.
    if (!ctx->has_tuning(TuneKey{m.field, m.log_n, m.batch, SFN_LAYOUT_SOA})) sc.max_log_t = kMaxLogTile;
    sc.logr = 5;
    jp[j] = plan_ntt_passes(m.field, SFN_VARIANT_AUTO, m.dir, m.log_n, m.batch, SFN_LAYOUT_SOA, sc, false);
    nl = std::max(nl, jp[j].size());
  }
  out.resize(nl);
  for (size_t k = 0; k < nl; ++k) {
    MultiLaunch &L = out[k];
    // tuned when every pass of this launch fits the tuned kernel and its
    // Montgomery-field passes share one field (the kernel stages one root table)
    bool tuned = ctx->allow_fast;
    int mfield = -1;
    std::vector<PassDesc> q(jobs.size());
    for (size_t j = 0; j < jobs.size() && tuned; ++j) {
      if (k >= jp[j].size()) continue;
      const MultiJob &m = jobs[j];
      q[j] = jp[j][k];
      if (m.field != SFN_FIELD_M31) {
        if (mfield >= 0 && mfield != m.field) tuned = false;
        mfield = m.field;
      }
      const Layout lay{(uint64_t)1 << m.log_n, 1};
      const bool aligned = ((uintptr_t)m.src % 16 == 0) && ((uintptr_t)m.dst % 16 == 0);
      tuned = tuned && choose_lanes_multi(q[j], m.batch) && (q[j].lane_contig || aligned) &&
              fast_offsets_fit(q[j], lay, lay);
    }
    L.tuned = tuned;
    size_t tile_bytes = 0;
    for (size_t j = 0; j < jobs.size(); ++j) {
      if (k >= jp[j].size()) continue;
      const MultiJob &m = jobs[j];
      const PassDesc &p = tuned ? q[j] : jp[j][k];
      const uint64_t n = 1ull << m.log_n;
      const uint32_t *src = k == 0 ? m.src : m.dst;  // later passes work in place on dst
      SegDesc d{};
      sfn_status st = make_pass_args(ctx, m.field, m.dir, p, src, m.dst, n, 1, n, 1, m.batch, m.log_n, &d.a);
      if (st) return st;
      d.kind = (uint32_t)(m.field * 2 + m.dir);
      d.log_t = (uint32_t)p.log_t;
      if (tuned) {
        if ((st = fill_fast_args(ctx, m.field, m.dir, p, (uint64_t)m.batch << m.log_n, (uint64_t)m.batch << m.log_n,
                                 &d.a)))
          return st;
        if (p.lane_contig) {
          d.kind += 6;
          if (p.log_w < 5) L.xs = std::max<uint32_t>(L.xs, std::min<uint32_t>(3, 5 - p.log_w));
        }
        if (m.field != SFN_FIELD_M31) L.mf_root = d.a.tw;  // the context's root table of that field
        L.lanes_bytes = std::max<uint32_t>(L.lanes_bytes, (uint32_t)kLaneInfoBytes << p.log_w);
        const size_t T = (size_t)1 << p.log_t;
        tile_bytes = std::max(tile_bytes, ((T + T / 32) << p.log_w) * 4);
      } else {
        if ((st = fill_generic_args(ctx, m.field, m.dir, p, &d.a))) return st;
        L.lds = std::max(L.lds, pass_lds_bytes(m.field, p.log_t, p.log_w));
      }
      d.tile_begin = L.tiles;
      L.tiles += p.tiles;
      L.segs.push_back(d);
    }
    if (tuned) L.lds = (L.mf_root ? kMultiRootBytes : 0) + L.lanes_bytes + tile_bytes;
  }
  return SFN_OK;
}

uint32_t multi_grid(const Context *ctx, const MultiLaunch &L) {
  const uint64_t cus = (uint64_t)std::max(1, ctx->props.multiProcessorCount);
  if (L.tuned) {
    // resident workgroups per CU by LDS (at most 4 of 4 waves), a multiple of 8
    const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(2, kMaxLdsBytes / std::max<size_t>(L.lds, 1)));
    const uint64_t cap = std::max<uint64_t>(8, (cus * per_cu) & ~7ull);
    const uint64_t want = (L.tiles + 7) & ~7ull;
    return (uint32_t)std::min(cap, want);
  }
  const uint32_t per_cu = L.lds > 80 * 1024 ? 1 : (L.lds > 40 * 1024 ? 2 : 4);
  return (uint32_t)std::min<uint64_t>(L.tiles, cus * per_cu * 2);
}

hipError_t launch_multi_any(const MultiLaunch &L, const SegDesc *d_segs, uint32_t grid, hipStream_t s) {
  if (L.tuned)
    return launch_multi_fast(d_segs, (uint32_t)L.segs.size(), L.tiles, L.lds, grid, L.xs, L.mf_root, L.lanes_bytes, s);
  return launch_multi(d_segs, (uint32_t)L.segs.size(), L.tiles, L.lds, grid, s);
}

}  // namespace sfn

// ======================================================================= C ABI
using sfn::Context;

namespace {
// Intermediate array of sfn_coset_lde_natural_device_noalloc: a static device
// array, so that entry point allocates nothing.  Every GPU the library runs on
// holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns for that device (kept per device).
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchBytes / 4];
std::mutex g_scratch_mu;
uint32_t *g_scratch_ptr[kMaxDevices] = {};

// With `device` current: the address of this device's copy of `scratch`.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return e != hipSuccess ? e : hipErrorInvalidValue;
  g_scratch_ptr[device] = (uint32_t *)p;
  return hipSuccess;
}

uint32_t *scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

static uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static hipStream_t pick_stream(Context &c, void *stream) { return stream ? (hipStream_t)stream : c.stream; }

// Host-pointer call: upload, run between an event pair, download.
template <class Run>
static sfn_status host_call(Context &c, const uint32_t *in, size_t in_elems, uint32_t *out, size_t out_elems,
                            bool in_place_ok, uint64_t *device_ns, Run run) {
  (void)hipSetDevice(c.device);
  sfn_status st = c.ensure(c.scratch_in, in_elems * 4);
  if (st) return st;
  uint32_t *d_in = (uint32_t *)c.scratch_in.ptr;
  uint32_t *d_out = d_in;
  if (!in_place_ok || out_elems != in_elems) {
    if ((st = c.ensure(c.scratch_out, out_elems * 4))) return st;
    d_out = (uint32_t *)c.scratch_out.ptr;
  }
  hipError_t e = hipMemcpyAsync(d_in, in, in_elems * 4, hipMemcpyHostToDevice, c.stream);
  if (e != hipSuccess) return c.hip_fail(e, "upload");
  if ((e = hipEventRecord(c.ev0, c.stream)) != hipSuccess) return c.hip_fail(e, "event");
  if ((st = run(d_in, d_out))) return st;
  if ((e = hipEventRecord(c.ev1, c.stream)) != hipSuccess) return c.hip_fail(e, "event");
  e = hipMemcpyAsync(out, d_out, out_elems * 4, hipMemcpyDeviceToHost, c.stream);
  if (e != hipSuccess) return c.hip_fail(e, "download");
  if ((e = hipStreamSynchronize(c.stream)) != hipSuccess) return c.hip_fail(e, "GPU execution");
  float ms = 0.f;
  if ((e = hipEventElapsedTime(&ms, c.ev0, c.ev1)) != hipSuccess) return c.hip_fail(e, "event timing");
  if (device_ns) *device_ns = (uint64_t)((double)ms * 1e6);
  c.timing_pending = true;
  return SFN_OK;
}

namespace sfn {

// Column chunks of a host <-> device batch call: at least kChunkBytes of input
// each.  SoA chunks are contiguous (up to 8); row-major chunks are column
// windows moved by 2D copies, kept at >= 64 columns (256-byte row pieces, which
// the copy engines move at full rate) and at most 4.
constexpr size_t kChunkBytes = (size_t)32 << 20;

uint32_t chunk_columns(uint64_t in_bytes, uint32_t batch, int layout) {
  const bool rm = layout == SFN_LAYOUT_ROW_MAJOR;
  uint64_t chunks = std::min<uint64_t>(rm ? 4 : 8, in_bytes / kChunkBytes);
  // row-major chunks move as 2D copies of (columns x 4)-byte row pieces: at
  // 512 bytes and up they run at the contiguous rate (measured 55.8 vs 57.6
  // GB/s; 256 bytes: 43.8, 128 bytes: 30.6)
  if (rm) chunks = std::min<uint64_t>(chunks, batch / 128);
  if (chunks <= 1) return 0;
  uint32_t q = (uint32_t)((batch + chunks - 1) / chunks);
  if (rm) q = (q + 15) & ~15u;  // whole tiles of the tuned kernels' lanes
  return q < batch ? q : 0;
}

sfn_status run_chunked(Context &c, ChunkStreams &cs, hipStream_t after, const uint32_t *in, uint64_t in_rows,
                       uint32_t *out, uint64_t out_rows, uint32_t batch, int layout, uint32_t q, uint32_t *d_in,
                       uint32_t *d_out, const ChunkRun &run, double *kernel_ns) {
  const bool rm = layout == SFN_LAYOUT_ROW_MAJOR;
  const uint32_t chunks = (batch + q - 1) / q;
  hipError_t e;
  for (auto &x : cs.s)
    if (!x && (e = hipStreamCreateWithFlags(&x, hipStreamNonBlocking)) != hipSuccess) {
      x = nullptr;
      return c.hip_fail(e, "transfer stream");
    }
  while (cs.ev.size() < 2 * (size_t)chunks + 1) {
    hipEvent_t ev;
    if ((e = hipEventCreate(&ev)) != hipSuccess) return c.hip_fail(e, "event");
    cs.ev.push_back(ev);
  }
  // work already queued on `after` (earlier calls) comes first
  if (after) {
    if ((e = hipEventRecord(cs.ev[2 * chunks], after)) != hipSuccess) return c.hip_fail(e, "event");
    for (auto x : cs.s)
      if (x != after && (e = hipStreamWaitEvent(x, cs.ev[2 * chunks], 0)) != hipSuccess)
        return c.hip_fail(e, "stream order");
  }
  for (uint32_t k = 0; k < chunks; ++k) {
    hipStream_t s = cs.s[k % ChunkStreams::kStreams];
    const uint32_t c0 = k * q, cols = std::min(q, batch - c0);
    uint32_t *ci = d_in + (size_t)c0 * in_rows, *co = d_out + (size_t)c0 * out_rows;
    if (rm)
      e = hipMemcpy2DAsync(ci, (size_t)cols * 4, in + c0, (size_t)batch * 4, (size_t)cols * 4, in_rows,
                           hipMemcpyHostToDevice, s);
    else
      e = hipMemcpyAsync(ci, in + (size_t)c0 * in_rows, (size_t)cols * in_rows * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return c.hip_fail(e, "upload");
    if ((e = hipEventRecord(cs.ev[2 * k], s)) != hipSuccess) return c.hip_fail(e, "event");
    sfn_status st = run(cols, ci, co, s);
    if (st) return st;
    if ((e = hipEventRecord(cs.ev[2 * k + 1], s)) != hipSuccess) return c.hip_fail(e, "event");
    if (rm)
      e = hipMemcpy2DAsync(out + c0, (size_t)batch * 4, co, (size_t)cols * 4, (size_t)cols * 4, out_rows,
                           hipMemcpyDeviceToHost, s);
    else
      e = hipMemcpyAsync(out + (size_t)c0 * out_rows, co, (size_t)cols * out_rows * 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return c.hip_fail(e, "download");
  }
  for (auto x : cs.s)
    if ((e = hipStreamSynchronize(x)) != hipSuccess) return c.hip_fail(e, "GPU execution");
  double ns = 0;
  for (uint32_t k = 0; k < chunks; ++k) {
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, cs.ev[2 * k], cs.ev[2 * k + 1])) != hipSuccess)
      return c.hip_fail(e, "event timing");
    ns += (double)ms * 1e6;
  }
  if (kernel_ns) *kernel_ns = ns;
  return SFN_OK;
}

}  // namespace sfn

// The same call split into column chunks that run on the context's transfer
// streams in turn, so the upload of one chunk, the kernels of another and the
// download of a third overlap (the two directions of the PCIe link run at
// once).  run(cols, d_in, d_out, stream) transforms `cols` columns.  device_ns
// is the sum of the chunks' kernel spans.  Calls that do not split take host_call.
template <class Run>
static sfn_status host_call_chunked(Context &c, const uint32_t *in, uint64_t in_rows, uint32_t *out, uint64_t out_rows,
                                    uint32_t batch, int layout, bool in_place_ok, uint64_t *device_ns, Run run) {
  const size_t in_elems = (size_t)batch * in_rows, out_elems = (size_t)batch * out_rows;
  const uint32_t q = sfn::chunk_columns(in_elems * 4, batch, layout);
  if (!q)
    return host_call(c, in, in_elems, out, out_elems, in_place_ok, device_ns,
                     [&](const uint32_t *di, uint32_t *dout) { return run(batch, di, dout, c.stream); });
  (void)hipSetDevice(c.device);
  sfn_status st = c.ensure(c.scratch_in, in_elems * 4);
  if (st) return st;
  uint32_t *d_in = (uint32_t *)c.scratch_in.ptr;
  uint32_t *d_out = d_in;
  if (!in_place_ok || out_elems != in_elems) {
    if ((st = c.ensure(c.scratch_out, out_elems * 4))) return st;
    d_out = (uint32_t *)c.scratch_out.ptr;
  }
  double ns = 0;
  st = sfn::run_chunked(c, c.chunks, c.stream, in, in_rows, out, out_rows, batch, layout, q, d_in, d_out, run, &ns);
  if (st) return st;
  if (device_ns) *device_ns = (uint64_t)ns;
  c.timing_pending = false;  // the context stream's event pair does not bracket this call
  return SFN_OK;
}

extern "C" {

const char *sfn_status_string(sfn_status s) {
  switch (s) {
    case SFN_OK: return "ok";
    case SFN_ERR_INVALID_SIZE: return "invalid size";
    case SFN_ERR_DEVICE_NOT_FOUND: return "device not found";
    case SFN_ERR_BUFFER_ALLOC_FAILED: return "buffer allocation failed";
    case SFN_ERR_KERNEL_LOAD: return "kernel load error";
    case SFN_ERR_GPU_EXECUTION: return "GPU execution error";
    case SFN_ERR_CORRECTNESS: return "correctness mismatch";
    case SFN_ERR_UNSUPPORTED: return "unsupported";
    case SFN_ERR_INVALID_ARGUMENT: return "invalid argument";
    case SFN_ERR_BUSY: return "busy";
    case SFN_ERR_NOT_READY: return "not ready";
  }
  return "unknown status";
}

sfn_status sfn_ctx_create(int device, sfn_ctx **out) {
  if (!out) return SFN_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count) {
    (void)hipGetLastError();
    return SFN_ERR_DEVICE_NOT_FOUND;
  }
  if (hipSetDevice(device) != hipSuccess) return SFN_ERR_DEVICE_NOT_FOUND;
  auto *c = new (std::nothrow) sfn_ctx();
  if (!c) return SFN_ERR_BUFFER_ALLOC_FAILED;
  Context &x = c->impl;
  x.device = device;
  if (hipGetDeviceProperties(&x.props, device) != hipSuccess ||
      hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&x.ev0) != hipSuccess || hipEventCreate(&x.ev1) != hipSuccess) {
    delete c;
    return SFN_ERR_GPU_EXECUTION;
  }
  const char *generic_only = getenv("SFNTT_GENERIC_ONLY");
  x.allow_fast = !(generic_only && generic_only[0] == '1');
  const char *xcd = getenv("SFNTT_XCD_REMAP");
  x.xcd_remap = !(xcd && xcd[0] == '0');
  if (strncmp(x.props.gcnArchName, "gfx950", 6) != 0) {
    delete c;
    return SFN_ERR_DEVICE_NOT_FOUND;  // the kernels are built for gfx950 only
  }
  if (scratch_setup(device) != hipSuccess) {  // this device's static scratch (allocation-free entry point)
    (void)hipGetLastError();
    delete c;
    return SFN_ERR_GPU_EXECUTION;
  }
  *out = c;
  return SFN_OK;
}

void sfn_ctx_destroy(sfn_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->impl.device);
  (void)hipStreamSynchronize(ctx->impl.stream);
  delete ctx;
}

const char *sfn_last_error(const sfn_ctx *ctx) {
  if (!ctx) return "";
  // a per-thread copy: another thread may replace the context's message meanwhile
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(ctx->impl.err_mu);
  copy = ctx->impl.last_error;
  return copy.c_str();
}

sfn_status sfn_device_info_get(sfn_ctx *ctx, sfn_device_info *info) {
  if (!ctx || !info) return SFN_ERR_INVALID_ARGUMENT;
  const hipDeviceProp_t &p = ctx->impl.props;
  memset(info, 0, sizeof(*info));
  strncpy(info->name, p.name, sizeof(info->name) - 1);
  strncpy(info->gcn_arch, p.gcnArchName, sizeof(info->gcn_arch) - 1);
  info->compute_units = p.multiProcessorCount;
  info->lds_bytes_per_cu = (int)p.maxSharedMemoryPerMultiProcessor;
  info->hbm_bytes = p.totalGlobalMem;
  info->clock_khz = p.clockRate;
  info->mem_clock_khz = p.memoryClockRate;
  info->mem_bus_width = p.memoryBusWidth;
  return SFN_OK;
}

sfn_status sfn_ntt(sfn_ctx *ctx, sfn_field field, sfn_variant variant, sfn_direction dir, uint32_t log_n,
                   uint32_t batch, sfn_layout layout, const uint32_t *in, uint32_t *out, uint64_t *device_ns) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_ntt(&c, field, variant, dir, log_n, batch, layout);
  if (st) return st;
  if (!in || !out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  const size_t elems = (size_t)batch << log_n;
  if (log_n == 0) {  // identity, no GPU work (src/ntt/metal_r2.rs:81-83)
    if (in != out) memmove(out, in, elems * 4);
    if (device_ns) *device_ns = 0;
    return SFN_OK;
  }
  if (variant == SFN_STOCKHAM_R2)  // its ping-pong array is one per context: one stream
    return host_call(c, in, elems, out, elems, true, device_ns, [&](const uint32_t *di, uint32_t *dout) {
      return c.run_ntt(field, variant, dir, log_n, batch, layout, di, dout, c.stream);
    });
  return host_call_chunked(c, in, 1ull << log_n, out, 1ull << log_n, batch, layout, true, device_ns,
                           [&](uint32_t cols, const uint32_t *di, uint32_t *dout, hipStream_t s) {
                             return c.run_ntt(field, variant, dir, log_n, cols, layout, di, dout, s);
                           });
}

sfn_status sfn_coset_lde_batch(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                               uint32_t shift_raw, sfn_layout layout, const uint32_t *in, uint32_t *out,
                               uint64_t *device_ns) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_lde(&c, log_n, batch, added_bits, layout);
  if (st) return st;
  if (!in || !out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  return host_call_chunked(c, in, 1ull << log_n, out, 1ull << (log_n + added_bits), batch, layout, false, device_ns,
                           [&](uint32_t cols, const uint32_t *di, uint32_t *dout, hipStream_t s) {
                             return c.run_lde(log_n, cols, added_bits, shift_raw, layout, di, dout, s);
                           });
}

sfn_status sfn_coset_lde_natural(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                               uint32_t shift_raw, sfn_layout layout, const uint32_t *in, uint32_t *out,
                               uint64_t *device_ns) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_lde(&c, log_n, batch, added_bits, layout);
  if (st) return st;
  if (!in || !out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  return host_call_chunked(c, in, 1ull << log_n, out, 1ull << (log_n + added_bits), batch, layout, false, device_ns,
                           [&](uint32_t cols, const uint32_t *di, uint32_t *dout, hipStream_t s) {
                             return c.run_lde(log_n, cols, added_bits, shift_raw, layout, di, dout, s, nullptr, true);
                           });
}

sfn_status sfn_pointwise_mul(sfn_ctx *ctx, sfn_field field, const uint32_t *a, const uint32_t *b, uint32_t *out,
                             size_t len) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  if (field != SFN_FIELD_M31 && field != SFN_FIELD_BABYBEAR && field != SFN_FIELD_KOALABEAR)
    return c.fail(SFN_ERR_INVALID_ARGUMENT, "unknown field");
  if (len == 0) return SFN_OK;
  if (!a || !b || !out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  (void)hipSetDevice(c.device);
  sfn_status st;
  if ((st = c.ensure(c.scratch_in, len * 8))) return st;
  if ((st = c.ensure(c.scratch_out, len * 4))) return st;
  uint32_t *da = (uint32_t *)c.scratch_in.ptr, *db = da + len, *dout = (uint32_t *)c.scratch_out.ptr;
  hipError_t e;
  if ((e = hipMemcpyAsync(da, a, len * 4, hipMemcpyHostToDevice, c.stream))) return c.hip_fail(e, "upload");
  if ((e = hipMemcpyAsync(db, b, len * 4, hipMemcpyHostToDevice, c.stream))) return c.hip_fail(e, "upload");
  if ((e = sfn::launch_pointwise(field, da, db, dout, len, c.stream))) return c.hip_fail(e, "launch");
  if ((e = hipMemcpyAsync(out, dout, len * 4, hipMemcpyDeviceToHost, c.stream))) return c.hip_fail(e, "download");
  if ((e = hipStreamSynchronize(c.stream))) return c.hip_fail(e, "GPU execution");
  return SFN_OK;
}

// ---- device-resident tier ----
sfn_status sfn_buf_alloc(sfn_ctx *ctx, size_t bytes, void **dptr) {
  if (!ctx || !dptr) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  *dptr = nullptr;
  if (bytes == 0) return c.fail(SFN_ERR_BUFFER_ALLOC_FAILED, "zero-length allocation");  // src/gpu.rs:94-96
  (void)hipSetDevice(c.device);
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    return c.fail(SFN_ERR_BUFFER_ALLOC_FAILED, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return SFN_OK;
}

sfn_status sfn_buf_free(sfn_ctx *ctx, void *dptr) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  if (!dptr) return SFN_OK;
  (void)hipSetDevice(ctx->impl.device);
  hipError_t e = hipFree(dptr);
  if (e != hipSuccess) return ctx->impl.hip_fail(e, "hipFree");
  return SFN_OK;
}

sfn_status sfn_upload(sfn_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, pick_stream(c, stream));
  return e == hipSuccess ? SFN_OK : c.hip_fail(e, "upload");
}

sfn_status sfn_download(sfn_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, pick_stream(c, stream));
  return e == hipSuccess ? SFN_OK : c.hip_fail(e, "download");
}

sfn_status sfn_synchronize(sfn_ctx *ctx, void *stream) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  hipError_t e = hipStreamSynchronize(pick_stream(c, stream));
  return e == hipSuccess ? SFN_OK : c.hip_fail(e, "GPU execution");
}

sfn_status sfn_twiddles_prepare(sfn_ctx *ctx, sfn_field field, uint32_t log_n) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  (void)hipSetDevice(c.device);
  if (field == SFN_FIELD_BABYBEAR || field == SFN_FIELD_KOALABEAR) {
    const uint32_t max_log = sfn::host::monty_field(field).max_log;
    if (log_n > max_log) return c.fail(SFN_ERR_INVALID_SIZE, "log_n exceeds " + std::to_string(max_log));
    return c.position_table(field, log_n, nullptr, nullptr);
  }
  if (field != SFN_FIELD_M31) return c.fail(SFN_ERR_INVALID_ARGUMENT, "unknown field");
  if (log_n > 30) return c.fail(SFN_ERR_INVALID_SIZE, "log_n exceeds 30");
  const uint32_t *t;
  sfn_status st = c.m31_table(log_n, false, &t);
  if (st) return st;
  return c.m31_table(log_n, true, &t);
}

sfn_status sfn_twiddles_bytes(sfn_ctx *ctx, size_t *bytes) {
  if (!ctx || !bytes) return SFN_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::recursive_mutex> g(ctx->impl.mu);
  *bytes = ctx->impl.table_bytes;
  return SFN_OK;
}

sfn_status sfn_twiddles_limit(sfn_ctx *ctx, size_t max_bytes) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(ctx->impl.device);
  return ctx->impl.set_table_limit(max_bytes);
}

sfn_status sfn_ntt_device(sfn_ctx *ctx, sfn_field field, sfn_variant variant, sfn_direction dir, uint32_t log_n,
                          uint32_t batch, sfn_layout layout, const uint32_t *d_in, uint32_t *d_out, void *stream) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_ntt(&c, field, variant, dir, log_n, batch, layout);
  if (st) return st;
  if (!d_in || !d_out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  (void)hipSetDevice(c.device);
  hipStream_t s = pick_stream(c, stream);
  const bool own = s == c.stream, timed = own && c.call_timing;
  if (timed) (void)hipEventRecord(c.ev0, s);
  st = c.run_ntt(field, variant, dir, log_n, batch, layout, d_in, d_out, s);
  if (timed) (void)hipEventRecord(c.ev1, s);
  if (own) c.timing_pending = timed && st == SFN_OK;
  return st;
}

sfn_status sfn_coset_lde_batch_device(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                                      uint32_t shift_raw, sfn_layout layout, const uint32_t *d_in, uint32_t *d_out,
                                      void *stream) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_lde(&c, log_n, batch, added_bits, layout);
  if (st) return st;
  if (!d_in || !d_out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  // the first pass reads column b at b*n and writes it at b*N: an overlapping
  // output would overwrite input columns not read yet
  if (sfn::ranges_overlap(d_in, ((size_t)batch << log_n) * 4, d_out, ((size_t)batch << (log_n + added_bits)) * 4))
    return c.fail(SFN_ERR_INVALID_ARGUMENT, "LDE input and output must not overlap");
  (void)hipSetDevice(c.device);
  hipStream_t s = pick_stream(c, stream);
  const bool own = s == c.stream, timed = own && c.call_timing;
  if (timed) (void)hipEventRecord(c.ev0, s);
  st = c.run_lde(log_n, batch, added_bits, shift_raw, layout, d_in, d_out, s);
  if (timed) (void)hipEventRecord(c.ev1, s);
  if (own) c.timing_pending = timed && st == SFN_OK;
  return st;
}

sfn_status sfn_coset_lde_natural_device(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                                        uint32_t shift_raw, sfn_layout layout, const uint32_t *d_in,
                                        uint32_t *d_out, void *stream) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_lde(&c, log_n, batch, added_bits, layout);
  if (st) return st;
  if (!d_in || !d_out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  // the first pass reads column b at b*n and writes it at b*N: an overlapping
  // output would overwrite input columns not read yet
  if (sfn::ranges_overlap(d_in, ((size_t)batch << log_n) * 4, d_out, ((size_t)batch << (log_n + added_bits)) * 4))
    return c.fail(SFN_ERR_INVALID_ARGUMENT, "LDE input and output must not overlap");
  (void)hipSetDevice(c.device);
  hipStream_t s = pick_stream(c, stream);
  const bool own = s == c.stream, timed = own && c.call_timing;
  if (timed) (void)hipEventRecord(c.ev0, s);
  st = c.run_lde(log_n, batch, added_bits, shift_raw, layout, d_in, d_out, s, nullptr, true);
  if (timed) (void)hipEventRecord(c.ev1, s);
  if (own) c.timing_pending = timed && st == SFN_OK;
  return st;
}

sfn_status sfn_coset_lde_natural_device_noalloc(sfn_ctx *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits,
                                                uint32_t shift_raw, sfn_layout layout, const uint32_t *d_in,
                                                uint32_t *d_out, void *stream) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  sfn_status st = sfn::validate_lde(&c, log_n, batch, added_bits, layout);
  if (st) return st;
  if (!d_in || !d_out) return c.fail(SFN_ERR_INVALID_ARGUMENT, "null buffer");
  const size_t in_bytes = ((size_t)batch << log_n) * 4, out_bytes = ((size_t)batch << (log_n + added_bits)) * 4;
  if (sfn::ranges_overlap(d_in, in_bytes, d_out, out_bytes))
    return c.fail(SFN_ERR_INVALID_ARGUMENT, "LDE input and output must not overlap");
  if (out_bytes > kNoAllocScratchBytes)
    return c.fail(SFN_ERR_INVALID_SIZE, "output of " + std::to_string(out_bytes) +
                                            " bytes exceeds the allocation-free LDE's static scratch of " +
                                            std::to_string(kNoAllocScratchBytes) + " bytes");
  uint32_t *scr = scratch_of(c.device);
  if (!scr) return c.fail(SFN_ERR_NOT_READY, "static scratch address not set up for this device");
  (void)hipSetDevice(c.device);
  {  // every table resident?  A dry run looks them up without launching anything
    sfn::NoAllocScope dry(scr, kNoAllocScratchBytes, true);
    if ((st = c.run_lde(log_n, batch, added_bits, shift_raw, layout, d_in, d_out, c.stream, nullptr, true))) return st;
  }
  sfn::NoAllocScope na(scr, kNoAllocScratchBytes, false);
  hipStream_t s = pick_stream(c, stream);
  const bool own = s == c.stream, timed = own && c.call_timing;
  if (timed) (void)hipEventRecord(c.ev0, s);
  st = c.run_lde(log_n, batch, added_bits, shift_raw, layout, d_in, d_out, s, nullptr, true);
  if (timed) (void)hipEventRecord(c.ev1, s);
  if (own) c.timing_pending = timed && st == SFN_OK;
  return st;
}

sfn_status sfn_call_timing(sfn_ctx *ctx, int enable) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  ctx->impl.call_timing = enable != 0;
  return SFN_OK;
}

sfn_status sfn_last_device_ns(sfn_ctx *ctx, uint64_t *ns) {
  if (!ctx || !ns) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  if (!c.timing_pending)
    return c.fail(SFN_ERR_INVALID_ARGUMENT,
                  "no timed launch sequence: the last call on the context stream failed or ran with call timing off");
  hipError_t e = hipEventSynchronize(c.ev1);
  if (e != hipSuccess) return c.hip_fail(e, "event");
  float ms = 0.f;
  e = hipEventElapsedTime(&ms, c.ev0, c.ev1);
  if (e != hipSuccess) return c.hip_fail(e, "event timing");
  *ns = (uint64_t)((double)ms * 1e6);
  return SFN_OK;
}

sfn_status sfn_pass_timing(sfn_ctx *ctx, int enable) {
  if (!ctx) return SFN_ERR_INVALID_ARGUMENT;
  ctx->impl.pass_timing = enable != 0;
  ctx->impl.tcount = 0;
  return SFN_OK;
}

sfn_status sfn_pass_times(sfn_ctx *ctx, uint64_t *ns, uint32_t cap, uint32_t *count) {
  if (!ctx || !count) return SFN_ERR_INVALID_ARGUMENT;
  Context &c = ctx->impl;
  *count = c.tcount;
  for (uint32_t i = 0; i < c.tcount && i < cap && ns; ++i) {
    hipError_t e = hipEventSynchronize(c.tev[2 * i + 1]);
    if (e != hipSuccess) return c.hip_fail(e, "event");
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, c.tev[2 * i], c.tev[2 * i + 1]);
    if (e != hipSuccess) return c.hip_fail(e, "event timing");
    ns[i] = (uint64_t)((double)ms * 1e6);
  }
  return SFN_OK;
}

const char *sfn_pass_label(sfn_ctx *ctx, uint32_t index) {
  if (!ctx || index >= ctx->impl.tcount || index >= ctx->impl.tlabel.size()) return "";
  return ctx->impl.tlabel[index].c_str();
}

}  // extern "C"

uint64_t sfn_now_ns() { return now_ns(); }
