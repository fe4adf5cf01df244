// Kernel argument blocks and launchers shared by kernels.hip and the runtime.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfn {

constexpr int kPassThreads = 256;  // workgroup size of the pass kernels
constexpr int kMaxLogTile = 11;    // largest LDS sub-transform (2^11 points)
constexpr size_t kMaxLdsBytes = 160 * 1024;  // LDS per workgroup on gfx950
// Powers of a Montgomery field's primitive 2^LOG_ROOT-th root g (LOG_ROOT <= 27),
// Montgomery form: words [0, 2^14) hold g^j, words 2^14 + j hold g^(j * 2^14)
// (j < 2^(LOG_ROOT - 14), the rest of the 2^13 words unused), so g^E for any
// E < 2^LOG_ROOT is the product of two entries.
constexpr uint32_t kPowLoBits = 14;
constexpr size_t kPowWords = ((size_t)1 << kPowLoBits) + ((size_t)1 << 13);

// Lane enumeration orders of a pass (a lane = one sub-transform instance).
enum LaneOrder : uint32_t {
  kLowFastest = 0,  // lanes walk the low row bits, then columns, then high bits
  kColFastest = 1,  // lanes walk columns, then low bits, then high bits
  kHighFastest = 2  // lanes walk the high row bits, then columns (needs lo == 0)
};

// One pass = layers [lo, lo + LOG_T) of a transform of 2^log_n rows, on every
// column.  Row r of column c sits at c*cs + r*rs (SoA: cs = n, rs = 1;
// row-major: cs = 1, rs = batch).
struct PassArgs {
  const uint32_t *src;
  uint32_t *dst;
  uint64_t src_cs, src_rs;
  uint64_t dst_cs, dst_rs;
  uint64_t instances;   // batch << (log_n - LOG_T)
  // ceil(2^64 / batch) when batch >= 2 and instances <= 2^32 (lane decoding
  // divides by batch with two multiplies), else 0 (a 64-bit division)
  uint64_t batch_recip;
  uint32_t log_n;
  uint32_t lo;
  uint32_t batch;
  uint32_t log_w;       // lanes per workgroup = 1 << log_w
  uint32_t order;       // LaneOrder
  uint32_t lane_contig; // 1: consecutive lanes are adjacent in memory; 0: points are
  const void *tw;       // BabyBear: position table (Shoup pairs, layer t at 2^t - 1);
                        // M31: the transform's block table (layer l at n - 2^(log_n-l))
  const uint2 *twist;   // BabyBear four-step twist [(p << lo) | low], or null
  uint32_t has_scale;   // multiply every stored value by `scale` (1/n)
  uint32_t scale_w, scale_q;
  // coset-LDE middle pass only
  const uint2 *mid;       // blocks tables of 2^log_n Shoup pairs
  const uint2 *twist_fwd; // forward twist of the blocks' high pass
  uint32_t blocks;
  uint64_t dst_block_off; // element offset between consecutive output blocks
  // tuned kernels (fast.hip): buffer extents and single-word Montgomery tables
  uint64_t src_elems, dst_elems;  // elements addressable from src / dst
  const uint32_t *twist_m;        // four-step twist of this pass, Montgomery form
  const uint32_t *post_m;         // LDE: per-block post-twist x n^-1 * sigma_b^low (blocks x 2^log_n)
  const uint32_t *pointf_m;       // LDE: per-block point factors sigma_b^(p << lo) (blocks x 2^LOG_T)
  uint32_t xcd_share_log;         // 2^s consecutive tiles share 128-byte lines: keep them on one XCD
  uint32_t tab_low_major;         // twist_m / post_m entries ordered (low, p) instead of (p, low)
  // non-null (with twist_m null): the pass's four-step twist is generated
  // from these root powers (kPowWords words) instead of read per point
  const uint32_t *pow_m;
  // LDE middle pass over lanes of low row bits, with pow_m: the post factors
  // are generated, times word b * 2^lo + low of this table (n^-1 sigma_b^low)
  const uint32_t *lanef_m;
  // rows stored in bit-reversed order over this many bits (0: storage order);
  // tuned lane-contiguous passes with lo == 0 only
  uint32_t src_rev_bits, dst_rev_bits;
};

// V1 (one launch per layer) arguments.
struct LayerArgs {
  const uint32_t *src;
  uint32_t *dst;
  uint64_t src_cs, src_rs, dst_cs, dst_rs;
  uint64_t butterflies;  // batch << (log_n - 1)
  uint32_t log_n;
  uint32_t layer;
  uint32_t batch;
  const void *tw;  // BabyBear: full position table; M31: block table
  uint32_t has_scale, scale_w, scale_q;
};

// One segment of a multi-job launch: a pass of one job over its own tiles.
struct SegDesc {
  PassArgs a;
  uint64_t tile_begin;  // first global tile of this segment (segments sorted)
  uint32_t kind;        // field * 2 + direction (+ 6 for the tuned kernel when lanes are adjacent)
  uint32_t log_t;       // sub-transform size (register radix min(5, log_t))
};

constexpr size_t kLaneInfoBytes = 32;  // per-lane decode record kept in LDS

// Dynamic LDS of a pass: [lane records][Montgomery-field small table][tile, W x (T + T/32)]
inline size_t pass_lds_bytes(int field, int log_t, uint32_t log_w) {
  const size_t T = (size_t)1 << log_t;
  const size_t TP = T + (T >> 5);
  return (kLaneInfoBytes << log_w) + (field != 0 ? T * 8 : 0) + (TP << log_w) * 4;
}
// LDE middle pass: [lane records][inverse + forward small tables][tile]
inline size_t lde_mid_lds_bytes(int log_t, uint32_t log_w) {
  const size_t T = (size_t)1 << log_t;
  const size_t TP = T + (T >> 5);
  return (kLaneInfoBytes << log_w) + 2 * T * 8 + (TP << log_w) * 4;
}
hipError_t launch_pass(int field, int dir, int log_t, int logr, const PassArgs &a,
                       uint64_t tiles, hipStream_t s);
hipError_t launch_lde_mid(int log_t, const PassArgs &a, uint64_t tiles, hipStream_t s);
hipError_t launch_layer(int field, int dir, const LayerArgs &a, hipStream_t s);
hipError_t launch_pointwise(int field, const uint32_t *a, const uint32_t *b, uint32_t *out,
                            uint64_t len, hipStream_t s);
hipError_t launch_multi(const SegDesc *d_segs, uint32_t nseg, uint64_t total_tiles, size_t lds, uint32_t grid,
                        hipStream_t s);
size_t fast_lds_bytes(int field, int log_t, uint32_t log_w);
hipError_t launch_fast(int field, int dir, int log_t, bool contig, const PassArgs &a, uint64_t tiles,
                       hipStream_t s);
// contig: lanes adjacent in memory; otherwise each lane is one column of
// adjacent points (lo == 0, 16-byte aligned columns)
hipError_t launch_lde_mid_fast(int log_t, bool contig, const PassArgs &a, uint64_t tiles, hipStream_t s);
// Tuned multi-job pass: segments' kind = field * 2 + dir (+ 4 when lanes are
// adjacent in memory), log_t in [6, 11]; kMultiFastThreads threads per
// workgroup (at most 128 VGPRs: two workgroups, four waves per SIMD).
// bb_root: the root table of the launch's Montgomery field (2^11 Shoup pairs,
// of which the kernel stages the half e <= 2^10) when any segment has one, else
// null; lanes_bytes: the lane-record region.
// LDS = (bb_root ? kMultiRootBytes : 0) + lanes_bytes + the largest tile.
constexpr int kMultiFastThreads = 512;
// half root table of the many-job kernel: g^e for e <= 2^10, 16-byte padded
constexpr size_t kMultiRootBytes = ((((size_t)1 << 10) + 1) * 8 + 15) & ~(size_t)15;
hipError_t launch_multi_fast(const SegDesc *d_segs, uint32_t nseg, uint64_t total_tiles, size_t lds, uint32_t grid,
                             uint32_t xs, const void *bb_root, uint32_t lanes_bytes, hipStream_t s);
hipError_t launch_copy(const uint32_t *src, uint32_t *dst, uint64_t len, hipStream_t s);
// One launch copying many ranges: word i of [segs[k].begin, segs[k+1].begin)
// (the last ends at total) goes from segs[k].src to segs[k].dst, both indexed
// from the segment's start.  segs is a device array of nseg entries, begin ascending.
struct CopySeg {
  const uint32_t *src;
  uint32_t *dst;
  uint64_t begin;
};
hipError_t launch_copy_segments(const CopySeg *d_segs, uint32_t nseg, uint64_t total, hipStream_t s);
// dst row r = src row bitrev_L(r) for a 2^L-row, `batch`-column array with the
// given (column, row) strides; src == dst permutes in place (same strides).
hipError_t launch_rows_bitrev(const uint32_t *src, uint32_t *dst, uint32_t L, uint32_t batch, uint64_t scs,
                              uint64_t srs, uint64_t dcs, uint64_t drs, hipStream_t s);

}  // namespace sfn
