// Host runtime: context, device-resident table cache, planner and executor.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sfntt.h"
#include "kernels.hpp"

namespace sfn {

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// One pass (or one V1 layer) of a plan.
struct PassDesc {
  bool layer_kernel = false;  // V1: one butterfly per thread for `layer`
  uint32_t layer = 0;
  bool lde_mid = false;
  bool fast = false;          // tuned kernel (fast.hip): radix-32, one item per thread
  int log_t = 0, logr = 0;
  uint32_t lo = 0;
  uint32_t log_n = 0;         // rows of the (virtual) transform the pass indexes
  uint32_t log_w = 0;
  uint32_t order = 0, lane_contig = 0;
  uint64_t tiles = 0;
  int src_sel = 0, dst_sel = 0;  // 0 = caller input, 1 = caller output, 2 = scratch
  bool scale = false;
  uint32_t src_rev = 0, dst_rev = 0;  // rows bit-reversed over this many bits (tuned passes, lo == 0)
};

struct Schedule {
  uint32_t max_log_t = 10;  // largest LDS sub-transform
  uint32_t logr = 5;        // register radix
  uint32_t log_w = 4;       // lanes per workgroup (strided passes)
  uint32_t mid_log_w = 3;   // lanes per workgroup of the LDE middle pass
};

// Key of a per-shape device table in the context's bounded cache.
enum TableKind : uint32_t {
  kTwistPair = 1,   // four-step twist, Shoup pairs: (field, lo, hi, inverse)
  kTwistMont = 2,   // four-step twist, Montgomery words: (field, lo, hi, inverse, low_major)
  kLdeMid = 3,      // generic LDE middle-pass blocks: (log_n, added_bits, shift)
  kLdeFast = 4,     // tuned LDE post / point-factor tables: (log_n, added_bits, shift, lo, low_major)
  kM31Block = 5,    // M31 block twiddles: (log_n, inverse)
  kLdeLane = 6      // tuned LDE per-lane post factors n^-1 sigma_b^low: (log_n, added_bits, shift, lo)
};
struct TableKey {
  uint32_t kind, a, b, c, d, e;
  bool operator<(const TableKey &o) const {
    return std::tie(kind, a, b, c, d, e) < std::tie(o.kind, o.a, o.b, o.c, o.d, o.e);
  }
};
struct TableEntry {
  DevBuf buf[2];
  size_t bytes = 0;
  uint64_t last_use = 0;  // LRU clock
  uint32_t pins = 0;      // open TablePins scopes / plans that hold it
};

// Streams and events of a host <-> device call split into column chunks
// (created on first use; the context and every server worker own one set).
struct ChunkStreams {
  static constexpr int kStreams = 3;
  hipStream_t s[kStreams] = {nullptr, nullptr, nullptr};
  std::vector<hipEvent_t> ev;  // a start / end pair per chunk, and one ordering event
  void release() {
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.clear();
    for (auto &x : s)
      if (x) {
        (void)hipStreamDestroy(x);
        x = nullptr;
      }
  }
};

// While a NoAllocScope is open on a thread (the allocation-free entry point),
// the runtime allocates no device memory: a table or scratch buffer that is
// not resident fails the call with SFN_ERR_NOT_READY instead of being built,
// and the natural-order LDE's intermediate array is `scratch` (the library's
// static device array, `bytes` long) instead of the context's.  A dry scope
// also launches nothing (the call only finds out whether it could run).
class NoAllocScope {
 public:
  NoAllocScope(uint32_t *scratch, size_t bytes, bool dry);
  ~NoAllocScope();
  NoAllocScope(const NoAllocScope &) = delete;
  NoAllocScope &operator=(const NoAllocScope &) = delete;
  static const NoAllocScope *current();
  uint32_t *scratch;
  size_t bytes;
  bool dry;

 private:
  const NoAllocScope *prev_;
};

class Context;
// While a TablePins scope is open on a thread, every cached table the thread
// obtains from `ctx` stays resident until the scope closes (its launches are
// enqueued by then) or, after take(), until the new owner calls unpin.
class TablePins {
 public:
  explicit TablePins(Context *ctx);
  ~TablePins();
  TablePins(const TablePins &) = delete;
  TablePins &operator=(const TablePins &) = delete;
  std::vector<TableKey> take();  // the pins move to the caller (plans keep theirs)
  void add(const TableKey &k) { keys_.push_back(k); }
  Context *ctx() const { return ctx_; }

 private:
  Context *ctx_;
  TablePins *prev_;
  std::vector<TableKey> keys_;
};

struct TuneKey {
  int field;
  uint32_t log_n, batch;
  int layout;
  bool operator<(const TuneKey &o) const {
    return std::tie(field, log_n, batch, layout) < std::tie(o.field, o.log_n, o.batch, o.layout);
  }
};

class Context {
 public:
  int device = 0;
  hipStream_t stream = nullptr;
  ChunkStreams chunks;  // host-pointer calls split into column chunks
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // (ev0, ev1) bracket the last launch sequence on `stream` (read by sfn_last_device_ns)
  bool timing_pending = false;
  // the *_device entry points record that pair only when enabled: the two
  // event records cost ~8-10 us of device time per call
  bool call_timing = false;
  std::string last_error;  // written under err_mu (server workers share the context)
  mutable std::mutex err_mu;
  hipDeviceProp_t props{};
  std::recursive_mutex mu;  // serialises table-cache updates (server threads share it)

  // device-resident tables; Montgomery-field tables are indexed by mf_index(field)
  // (0 = BabyBear, 1 = KoalaBear)
  static int mf_index(int field) { return field == 2 ? 1 : 0; }
  DevBuf mf_pos[2][2];     // position tables (fwd, inv), size 2^mf_pos_log pairs
  uint32_t mf_pos_log[2] = {0, 0};
  DevBuf bb_small_pair;    // [inv small | fwd small], 2^kMaxLogTile each (LDE middle pass)
  DevBuf mf_root[2];       // g_{2^11}^e, e < 2^11, Shoup pairs (tuned kernels)
  DevBuf mf_pow[2];        // the field root's powers, Montgomery words (generated twists)
  // Per-shape tables (twists, LDE factors, M31 block twiddles): least recently
  // used entries are freed when the cache would exceed table_limit bytes (an
  // entry larger than the limit is still admitted alone).  Pinned entries are
  // never evicted; an eviction waits for the device first, since kernels of
  // any stream may still read an unpinned table.
  std::map<TableKey, TableEntry> tables;
  size_t table_limit = kDefaultTableLimit;
  size_t cached_bytes = 0;  // bytes of `tables`
  uint64_t table_clock = 0;
  uint64_t evictions = 0;
  static constexpr size_t kDefaultTableLimit = (size_t)4 << 30;
  size_t table_bytes = 0;       // every device table the context holds
  std::vector<DevBuf> retired;  // outgrown position tables, freed with the context

  // scratch for the host-pointer API and V3 ping-pong
  DevBuf scratch_in, scratch_out, scratch_pp;
  // intermediate array of the natural-order LDE; calls on any stream take
  // turns on it through scratch_ev (recorded after each use)
  DevBuf scratch_lde;
  hipEvent_t scratch_ev = nullptr;
  bool scratch_ev_used = false;

  std::map<TuneKey, Schedule> tuned;
  mutable std::mutex tune_mu;  // guards `tuned` (server workers read it while a caller tunes)
  bool allow_fast = true;  // SFNTT_GENERIC_ONLY=1 at context creation turns the tuned kernels off
  bool xcd_remap = true;   // SFNTT_XCD_REMAP=0 at context creation keeps workgroup id = tile

  // per-kernel event timing of the last launch sequence on `stream`
  bool pass_timing = false;
  std::vector<hipEvent_t> tev;     // 2 per kernel, grown on demand
  std::vector<std::string> tlabel; // kernel labels of the current sequence
  uint32_t tcount = 0;
  void timing_begin_sequence(hipStream_t s) {
    if (s == stream) tcount = 0;
  }
  hipError_t timing_mark(hipStream_t s, bool before, const std::string &label);

  ~Context();
  sfn_status fail(sfn_status s, const std::string &msg);
  sfn_status hip_fail(hipError_t e, const char *what);
  void clear_error();

  sfn_status ensure(DevBuf &b, size_t bytes);
  // scratch_lde of at least `bytes`, ordered after its previous user on any stream
  sfn_status acquire_lde_scratch(size_t bytes, hipStream_t s, uint32_t **out);
  void release_lde_scratch(hipStream_t s);
  sfn_status upload_table(DevBuf &b, const void *host, size_t bytes);
  void release_table(DevBuf &b);
  // Cached table `k`: built by build(k, h0, h1) (host words of up to two buffers)
  // on a miss; pinned in the innermost open TablePins scope of this thread.
  using TableBuilder = void (*)(const TableKey &k, std::vector<uint32_t> &h0, std::vector<uint32_t> &h1);
  sfn_status cached_table(const TableKey &k, TableBuilder build, const void **p0, const void **p1 = nullptr);
  void unpin(const std::vector<TableKey> &keys);
  // frees unpinned least-recently-used entries until `incoming` more bytes fit
  void evict_for(size_t incoming);
  sfn_status set_table_limit(size_t bytes);
  // field: SFN_FIELD_BABYBEAR or SFN_FIELD_KOALABEAR for the Montgomery-field tables
  sfn_status position_table(int field, uint32_t log_max, const uint2 **fwd, const uint2 **inv);
  sfn_status bb_small_pair_table(const uint2 **out);
  sfn_status root_table(int field, const uint2 **out);
  sfn_status pow_table(int field, const uint32_t **out);
  // low_major: entry (low << (hi - lo)) | p instead of (p << lo) | low
  sfn_status twist_m_table(int field, uint32_t lo, uint32_t hi, bool inverse, bool low_major,
                           const uint32_t **out);
  sfn_status bb_lde_fast(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo, bool low_major,
                         const uint32_t **post, const uint32_t **pointf);
  sfn_status bb_lde_lane(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo, const uint32_t **out);
  sfn_status twist_table(int field, uint32_t lo, uint32_t hi, bool inverse, const uint2 **out);
  sfn_status bb_mid_table(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, const uint2 **out);
  sfn_status m31_table(uint32_t log_n, bool inverse, const uint32_t **out);

  // Autotuned schedule of (field, log_n, batch, layout) if one is stored (field 2 =
  // the LDE), else the default for the size; field -1 always gives the default.
  Schedule schedule_for(int field, uint32_t log_n, uint32_t batch, int layout) const;
  Schedule schedule_for_column(int field, uint32_t log_n, int layout) const;  // one AUTO column
  Schedule schedule_for_batch(int field, uint32_t log_n, uint32_t batch, int layout) const;  // AUTO batch
  Schedule schedule_for_lde(uint32_t log_n, uint32_t batch, int layout) const;                // coset LDE
  bool has_tuning(const TuneKey &k) const;

  // Plans and runs a transform on device buffers.
  sfn_status run_ntt(int field, int variant, int dir, uint32_t log_n, uint32_t batch, int layout,
                     const uint32_t *d_in, uint32_t *d_out, hipStream_t s, const Schedule *force = nullptr);
  // natural: rows in natural order on both sides (input row i = P(g_n^i), output
  // row k = P(shift * g_N^k)) instead of bit-reversed.
  sfn_status run_lde(uint32_t log_n, uint32_t batch, uint32_t added_bits, uint32_t shift_raw, int layout,
                     const uint32_t *d_in, uint32_t *d_out, hipStream_t s, const Schedule *force = nullptr,
                     bool natural = false);
};

std::vector<PassDesc> plan_ntt_passes(int field, int variant, int dir, uint32_t log_n, uint32_t batch, int layout,
                                      const Schedule &sc, bool allow_fast = false);
sfn_status make_pass_args(Context *ctx, int field, int dir, const PassDesc &p, const uint32_t *src, uint32_t *dst,
                          uint64_t src_cs, uint64_t src_rs, uint64_t dst_cs, uint64_t dst_rs, uint32_t batch,
                          uint32_t log_n_real, PassArgs *out);

sfn_status validate_ntt(Context *ctx, int field, int variant, int dir, uint32_t log_n, uint32_t batch,
                        int layout);

// ---- many transforms in one launch per pass (plans, server groups) ----
struct MultiJob {
  int field, dir;
  uint32_t log_n, batch;  // SoA columns
  const uint32_t *src;
  uint32_t *dst;
};
struct MultiLaunch {
  std::vector<SegDesc> segs;  // tile_begin ascending
  uint64_t tiles = 0;
  size_t lds = 0;
  bool tuned = false;  // k_multi_fast (radix-32 tiles) instead of the generic k_multi
  uint32_t xs = 0;     // XCD grouping of tiles (tuned launches)
  const void *mf_root = nullptr;  // tuned launches with Montgomery-field segments: their root table
  uint32_t lanes_bytes = 0;       // tuned launches: lane-record region
};
// Pass k of every job goes into launch k.  A launch uses the tuned kernel when
// every one of its passes fits it (tile 2^6..2^11, offsets, alignment).
sfn_status build_multi(Context *ctx, const std::vector<MultiJob> &jobs, std::vector<MultiLaunch> &out);
// Persistent grid for a launch on this device.
uint32_t multi_grid(const Context *ctx, const MultiLaunch &L);
hipError_t launch_multi_any(const MultiLaunch &L, const SegDesc *d_segs, uint32_t grid, hipStream_t s);
sfn_status validate_lde(Context *ctx, uint32_t log_n, uint32_t batch, uint32_t added_bits, int layout);

// Host <-> device batch calls in column chunks: chunk_columns gives the
// columns per chunk for `in_bytes` of input (0: one piece is better), and
// run_chunked uploads chunk k, runs run(cols, d_in_k, d_out_k, stream) and
// downloads it on stream k mod 3 of `cs`, after the work already queued on
// `after` (may be null).  d_in / d_out hold the whole batch (chunk k at column
// k * q; row-major chunks compact).  kernel_ns: the sum of the chunks' spans.
using ChunkRun = std::function<sfn_status(uint32_t cols, const uint32_t *d_in, uint32_t *d_out, hipStream_t s)>;
uint32_t chunk_columns(uint64_t in_bytes, uint32_t batch, int layout);
sfn_status run_chunked(Context &c, ChunkStreams &cs, hipStream_t after, const uint32_t *in, uint64_t in_rows,
                       uint32_t *out, uint64_t out_rows, uint32_t batch, int layout, uint32_t q, uint32_t *d_in,
                       uint32_t *d_out, const ChunkRun &run, double *kernel_ns);
// [a, a + a_bytes) and [b, b + b_bytes) share a byte
bool ranges_overlap(const void *a, size_t a_bytes, const void *b, size_t b_bytes);

}  // namespace sfn

struct sfn_ctx {
  sfn::Context impl;
};
