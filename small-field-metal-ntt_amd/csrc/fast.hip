// Tuned pass kernels (radix-32 register blocks, LOG_T in [6, 11]).
//
// Compared with the generic LDS-staged pass (kernels.hip) these:
//  - load the first register round straight from HBM into registers and store
//    the last round straight from registers, so a 2^10-point pass touches LDS
//    once per element (one exchange) and each thread has 32 independent loads
//    in flight;
//  - address HBM through buffer resources rebased per workgroup: one 32-bit
//    VGPR offset per item, the per-point stride in a scalar offset, so the 32
//    loads and 32 stores of an item cost no vector ALU work for addresses;
//  - run exactly one item (one lane x 32 points) per thread in the full
//    rounds (blockDim = W * 2^(LOG_T-5));
//  - keep one root table g_T^e (e < T) in LDS that serves every layer and both
//    directions (layer t, position j uses e = j * 2^(LOG_T-t-1), negated for
//    the inverse);
//  - stream the four-step twists as single-word Montgomery constants, or,
//    where the lanes of a workgroup do not share their twist, generate them
//    (two root powers and a geometric sequence per item);
//  - lay LDS out so the exchanges are bank-conflict free: lane-fastest
//    (p + p/32) * W + lane when lanes are adjacent in memory, lane-major
//    lane * (T + T/32) + p + p/32 when points are.
// The host only selects these kernels when every offset inside a workgroup
// fits a 32-bit buffer offset.
#include <stdlib.h>

#include <type_traits>

#include "pass_common.hpp"

namespace sfn {

#ifndef SFN_PART
#define SFN_PART -1
#endif
#define SFN_IN_PART(k) (SFN_PART < 0 || SFN_PART == (k))

namespace fast {

constexpr int KR = 5;
constexpr int kRsrcWord3 = 0x00020000;  // raw buffer, gfx9 family

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc make_rsrc(const void *base, uint64_t bytes) {
  const uint32_t n = bytes > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)n, kRsrcWord3);
}
__device__ __forceinline__ uint32_t bload(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ void bstore(uint32_t v, Rsrc r, uint32_t voff, uint32_t soff) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)voff, (int)soff, 0);
}

template <int K>
__host__ __device__ constexpr uint32_t brev_c(uint32_t i) {
  uint32_t r = 0;
  for (int b = 0; b < K; ++b) r |= ((i >> b) & 1u) << (K - 1 - b);
  return r;
}

// v[i] *= tab[off + i * step] for the R values of an item (Montgomery table words).
template <class F, int R>
__device__ __forceinline__ void mul_table(uint32_t (&v)[R], Rsrc tab, uint32_t off_bytes, uint32_t step_bytes) {
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = F::mul_mont(v[i], bload(tab, off_bytes, i * step_bytes));
}

// Same with the table staged in LDS, indexed by the point alone.
template <class F, int R, int B0>
__device__ __forceinline__ void mul_lds_table(uint32_t (&v)[R], const uint32_t *tab, uint32_t pbase) {
#pragma unroll
  for (int i = 0; i < R; ++i) v[i] = F::mul_mont(v[i], tab[pbase | (i << B0)]);
}

// Copies T words from global to LDS with 16-byte accesses (T >= 4 * blockDim
// is not required: threads loop).
__device__ __forceinline__ void stage_words(uint32_t *dst, const uint32_t *src, uint32_t T) {
  for (uint32_t e = threadIdx.x * 4; e < T; e += blockDim.x * 4)
    *reinterpret_cast<uint4 *>(dst + e) = *reinterpret_cast<const uint4 *>(src + e);
}

// Montgomery-field twiddles from the LDS root table (Shoup pairs).
template <int LOG_T, int DIR>
struct RootTw {
  static constexpr bool kNeg = false;
  static constexpr bool kFoldRoots = true;
  static constexpr bool kSplitNeg = false;
  const uint2 *tab;
  __device__ __forceinline__ uint2 get(int t, uint32_t pos) const {
    uint32_t e = pos << (LOG_T - t - 1);
    if (DIR == 1) e = (0u - e) & ((1u << LOG_T) - 1);
    return tab[e];
  }
  // the same entry of the compile-time table (t, pos constants)
  template <class F>
  __device__ static __forceinline__ uint2 cget(int t, uint32_t pos) {
    uint32_t e = pos << (LOG_T - t - 1);
    if (DIR == 1) e = (0u - e) & ((1u << LOG_T) - 1);
    return make_uint2(kConstRoots<F, LOG_T>.w[e], kConstRoots<F, LOG_T>.q[e]);
  }
};

// Half root table g_T^e, e <= T/2 (T/2 + 1 Shoup pairs).  Layer t, position j
// uses e = j * 2^(LOG_T-t-1) < T/2.  The inverse needs g^-e = -g^(T/2-e)
// (g^(T/2) = -1), so it returns g^(T/2-e) and its CT butterflies take the
// twiddle as negated (kNeg); e = 0 gives -g^(T/2) = 1 as it should.
template <int LOG_T, int DIR>
struct HalfRootTw {
  static constexpr bool kNeg = DIR == 1;
  static constexpr bool kFoldRoots = true;
  static constexpr bool kSplitNeg = DIR == 1;
  const uint2 *tab;
  __device__ __forceinline__ uint2 get(int t, uint32_t pos) const {
    const uint32_t e = pos << (LOG_T - t - 1);
    return tab[DIR == 1 ? (1u << (LOG_T - 1)) - e : e];
  }
  // get(t, low | m << B0) of the inverse (entry T/2 - e) for m <= mmax,
  // addressed from the layer's largest position: the per-thread base is
  // shared by every m of the layer, and each m is a constant non-negative
  // offset from it.  With low < 2^B0 and mmax < 2^(t - B0), the base index
  // T/2 - (low + (mmax << B0)) << sh is at least T/2 - (2^t - 1) << sh = 2^sh.
  template <int B0>
  __device__ __forceinline__ uint2 get_split(int t, uint32_t low, uint32_t m, uint32_t mmax) const {
    const uint32_t sh = LOG_T - t - 1;
    const uint2 *base = tab + ((1u << (LOG_T - 1)) - (low << sh) - ((mmax << B0) << sh));
    return base[((mmax - m) << B0) << sh];
  }
  template <class F>
  __device__ static __forceinline__ uint2 cget(int t, uint32_t pos) {
    const uint32_t e = pos << (LOG_T - t - 1);
    const uint32_t i = DIR == 1 ? (1u << (LOG_T - 1)) - e : e;
    return make_uint2(kConstRoots<F, LOG_T>.w[i], kConstRoots<F, LOG_T>.q[i]);
  }
};

template <class F, int DIR, int LOG_T>
struct TwSel {
  using type = RootTw<LOG_T, DIR>;
};
template <int DIR, int LOG_T>
struct TwSel<M31, DIR, LOG_T> {
  using type = TwCtx<M31>;
};

template <int LOG_T, bool CONTIG>
__device__ __forceinline__ uint32_t lidx(uint32_t lane, uint32_t p, uint32_t log_w) {
  if constexpr (CONTIG) return ((p + (p >> 5)) << log_w) + lane;
  else return lane * (uint32_t)((1 << LOG_T) + (1 << LOG_T) / 32) + p + (p >> 5);
}

// Per-workgroup state shared by the rounds.  All offsets are relative to the
// workgroup's first lane (lane 0 has the lowest address of the tile).
struct RoundIO {
  Rsrc src, dst, twist, pow, lanef;
  const uint32_t *srcp;  // plain pointers (point-adjacent variant), already offset to lane 0
  uint32_t *dstp;
  uint64_t sbase0, dbase0;  // element index of lane 0's row 0 in src / dst
  uint32_t sps, dps;        // elements between consecutive points
  uint32_t lo;
  uint32_t tps, tls;  // twist / post entry of (p, low) = (p << tps) + (low << tls)
  // bit-reversed rows (lo == 0): point p of a lane sits bitrev_T(p) * srev
  // elements after the lane's base (0: rows in storage order)
  uint32_t srev, drev;
  uint32_t low0;      // low row bits of lane 0
  // LDS copies of this workgroup's twist / post tables, indexed by the point
  // (LDE middle pass with column-fastest lanes, which all share low0)
  const uint32_t *ltw, *lpost;
  bool has_twist;
  bool twist_gen;  // has_twist: generate the twist from `pow` (else read `twist`)
  bool post_gen;   // has_post: generate the post factors from `pow` and `lanef` (else read `post`)
  uint32_t has_scale, scale_w, scale_q;
  // LDE blocks: per-point factor table (one block) and post-twist table (one block)
  Rsrc pointf, post;
  bool has_pointf, has_post;
};

// First touch of a full round: the item's R values from HBM into registers
// (zeros for a lane past the end), no arithmetic.
template <class F, int DIR, int LOG_T, bool CONTIG, int K, int B0>
__device__ __forceinline__ void load_item(uint32_t (&v)[1 << K], const LaneInfo &li, const RoundIO &io,
                                          uint32_t pbase) {
  constexpr int R = 1 << K;
  if (!li.valid) {
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = 0;
    return;
  }
  if constexpr (!CONTIG) {
    // points are adjacent rows: plain global loads, 16-byte vectors when
    // a thread owns R consecutive points
    const uint32_t *sp = io.srcp + (li.sbase - io.sbase0) + (uint64_t)pbase * io.sps;
    if constexpr (B0 == 0 && K >= 2) {
      const uint4 *s4 = reinterpret_cast<const uint4 *>(sp);
#pragma unroll
      for (int i = 0; i < R / 4; ++i) {
        const uint4 q = s4[i];
        v[4 * i] = q.x; v[4 * i + 1] = q.y; v[4 * i + 2] = q.z; v[4 * i + 3] = q.w;
      }
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) v[i] = sp[(uint64_t)(i << B0) * io.sps];
    }
  } else if (io.srev) {
    // bitrev_T(pbase | i << B0) = bitrev_T(pbase) | bitrev_K(i) << (T - B0 - K)
    const uint32_t pr = __brev(pbase) >> (32 - LOG_T);
    const uint32_t roff = (uint32_t)(li.sbase - io.sbase0 + (uint64_t)pr * io.srev) * 4u;
    const uint32_t rstep = io.srev * 4u;
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = bload(io.src, roff, (brev_c<K>(i) << (LOG_T - B0 - K)) * rstep);
  } else {
    const uint32_t voff = (uint32_t)(li.sbase - io.sbase0 + (uint64_t)pbase * io.sps) * 4u;
    const uint32_t step = (io.sps << B0) * 4u;
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = bload(io.src, voff, i * step);
  }
}

// g_hi^(+-e) in Montgomery form, e < 2^hi, for the pass's four-step root
// g_hi = g^(2^(LOG_ROOT - hi)), hi = lo + LOG_T: g^E with E = e * 2^(LOG_ROOT - hi)
// (negated mod 2^LOG_ROOT for the inverse) is the product of two table words.
template <class F, int DIR, int LOG_T>
__device__ __forceinline__ uint32_t twist_root_pow(const RoundIO &io, uint32_t e) {
  constexpr uint32_t kMask = (1u << F::LOG_ROOT) - 1;
  uint32_t E = e << (F::LOG_ROOT - LOG_T - io.lo);
  if (DIR == 1) E = (0u - E) & kMask;
  return F::mul(bload(io.pow, (E & ((1u << kPowLoBits) - 1)) * 4u, 0),
                bload(io.pow, (E >> kPowLoBits) * 4u, 4u << kPowLoBits));
}

// Multiplies the item's values by their four-step twist, generated: the twist
// of (point p, low row bits l) is g_hi^(+-l * bitrev(p, LOG_T)), and for the
// item's points p = pbase | i << B0 (pbase zero on bits [B0, B0 + K)),
// bitrev(p) = bitrev(pbase) + bitrev_K(i) << (LOG_T - B0 - K): so the twist of
// point i is s * r^bitrev_K(i) with s = g_hi^(l * bitrev(pbase)) and
// r = g_hi^(l << (LOG_T - B0 - K)), a geometric sequence walked with one
// Montgomery product per point.  The walk stays in [0, 2p) (no final
// reduction: s < 2p, r < p keep s * r < p * 2^32), and the product with the
// value (< p) is reduced to [0, p).
// SCALED: every factor also times the lane's word of `lanef` (the LDE's
// post factors: forward twist x n^-1 sigma_b^low).
template <class F, int DIR, int LOG_T, int K, int B0, bool SCALED = false>
__device__ __forceinline__ void twist_gen(uint32_t (&v)[1 << K], const LaneInfo &li, const RoundIO &io,
                                          uint32_t pbase) {
  constexpr int R = 1 << K;
  const uint32_t brp = __brev(pbase) >> (32 - LOG_T);
  uint32_t s = twist_root_pow<F, DIR, LOG_T>(io, li.low * brp);
  if constexpr (SCALED) s = F::mul(s, bload(io.lanef, li.low * 4u, 0));
  const uint32_t r = twist_root_pow<F, DIR, LOG_T>(io, li.low << (LOG_T - B0 - K));
#pragma unroll
  for (int k = 0; k < R; ++k) {
    v[brev_c<K>(k)] = F::mul(v[brev_c<K>(k)], s);
    if (k + 1 < R) s = F::mul_lazy(s, r);
  }
}

// Four-step twist of an inverse Montgomery-field pass, applied to loaded values.
template <class F, int DIR, int LOG_T, int K, int B0, bool LT>
__device__ __forceinline__ void twist_item(uint32_t (&v)[1 << K], const LaneInfo &li, const RoundIO &io,
                                           uint32_t pbase) {
  constexpr int R = 1 << K;
  if constexpr (F::MONTY && DIR == 1) {
    if (io.has_twist && li.valid) {
      if constexpr (LT) mul_lds_table<F, R, B0>(v, io.ltw, pbase);
      else if (io.twist_gen) twist_gen<F, DIR, LOG_T, K, B0>(v, li, io, pbase);
      else mul_table<F, R>(v, io.twist, (((uint32_t)pbase << io.tps) + (li.low << io.tls)) * 4u, 4u << (B0 + io.tps));
    }
  }
}

// Last touch of a full round: forward twist / post factors / scale, then the
// item's R values from registers to HBM (nothing for a lane past the end).
template <class F, int DIR, int LOG_T, bool CONTIG, int K, int B0, bool LT>
__device__ __forceinline__ void store_item(uint32_t (&v)[1 << K], const LaneInfo &li, const RoundIO &io,
                                           uint32_t pbase) {
  constexpr int R = 1 << K;
  if (!li.valid) return;
  if constexpr (F::MONTY && DIR == 0) {
    const uint32_t toff = (((uint32_t)pbase << io.tps) + (li.low << io.tls)) * 4u;
    if (io.has_twist) {
      if (io.twist_gen) twist_gen<F, DIR, LOG_T, K, B0>(v, li, io, pbase);
      else mul_table<F, R>(v, io.twist, toff, 4u << (B0 + io.tps));
    }
    if constexpr (LT) {
      // the LDE middle pass always has post factors (and leaves the values
      // unreduced for them); keep the 32 table reads after the butterflies:
      // hoisting them above the round costs 32 live registers at the peak
      __builtin_amdgcn_sched_barrier(0);
      mul_lds_table<F, R, B0>(v, io.lpost, pbase);
    } else if (io.has_post) {
      if (io.post_gen) twist_gen<F, DIR, LOG_T, K, B0, true>(v, li, io, pbase);
      else mul_table<F, R>(v, io.post, toff, 4u << (B0 + io.tps));
    }
  }
  if (io.has_scale) {
#pragma unroll
    for (int i = 0; i < R; ++i) {
      if constexpr (F::MONTY) v[i] = F::mul_tw(v[i], make_uint2(io.scale_w, io.scale_q));
      else v[i] = M31::mul(v[i], io.scale_w);
    }
  }
  if constexpr (!CONTIG) {
    uint32_t *dp = io.dstp + (li.dbase - io.dbase0) + (uint64_t)pbase * io.dps;
    if constexpr (B0 == 0 && K >= 2) {
      uint4 *d4 = reinterpret_cast<uint4 *>(dp);
#pragma unroll
      for (int i = 0; i < R / 4; ++i) d4[i] = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) dp[(uint64_t)(i << B0) * io.dps] = v[i];
    }
  } else if (io.drev) {
    const uint32_t pr = __brev(pbase) >> (32 - LOG_T);
    const uint32_t roff = (uint32_t)(li.dbase - io.dbase0 + (uint64_t)pr * io.drev) * 4u;
    const uint32_t rstep = io.drev * 4u;
#pragma unroll
    for (int i = 0; i < R; ++i) bstore(v[i], io.dst, roff, (brev_c<K>(i) << (LOG_T - B0 - K)) * rstep);
  } else {
    const uint32_t voff = (uint32_t)(li.dbase - io.dbase0 + (uint64_t)pbase * io.dps) * 4u;
    const uint32_t step = (io.dps << B0) * 4u;
#pragma unroll
    for (int i = 0; i < R; ++i) bstore(v[i], io.dst, voff, i * step);
  }
}

// Round r of a pass: loads from HBM when FIRST, else from LDS; stores to HBM
// when LAST, else to LDS.  Items: W << (LOG_T - K).
template <class F, int DIR, int LOG_T, bool CONTIG, int K, int B0, bool FIRST, bool LAST, bool LT, class TW>
__device__ __forceinline__ void run_round(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes, TW tw,
                                          const RoundIO &io) {
  constexpr int R = 1 << K;
  constexpr int LOG_NB = LOG_T - K;
  const uint32_t W = 1u << log_w;
  const uint32_t items = W << LOG_NB;
  for (uint32_t it = threadIdx.x; it < items; it += blockDim.x) {
    uint32_t lane, rb;
    if constexpr (CONTIG) { lane = it & (W - 1); rb = it >> log_w; }
    else { rb = it & ((1u << LOG_NB) - 1); lane = it >> LOG_NB; }
    const uint32_t pbase = round_pbase<K, B0>(rb);
    const LaneInfo &li = lanes[lane];
    uint32_t v[R];
    if constexpr (FIRST) {
      load_item<F, DIR, LOG_T, CONTIG, K, B0>(v, li, io, pbase);
      twist_item<F, DIR, LOG_T, K, B0, LT>(v, li, io, pbase);
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) v[i] = tile[lidx<LOG_T, CONTIG>(lane, pbase | (i << B0), log_w)];
    }
    if constexpr (!F::MONTY) tw.hb = li.hb;
    // the LDE middle pass's forward last round multiplies every value by its
    // post factor (store_item), which takes unreduced values
    round_regs<F, DIR, K, B0, LT && LAST && DIR == 0>(v, pbase, tw);
    if constexpr (LAST) {
      store_item<F, DIR, LOG_T, CONTIG, K, B0, LT>(v, li, io, pbase);
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) tile[lidx<LOG_T, CONTIG>(lane, pbase | (i << B0), log_w)] = v[i];
    }
  }
}

// Rounds [r, rend) of a pass whose rounds are KRT layers each (the partial
// round as the schedule places it).
template <class F, int DIR, int LOG_T, bool CONTIG, int r, int rend, bool LT = false, int KRT = KR, class TW>
__device__ __forceinline__ void rounds(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes, const TW &tw,
                                       const RoundIO &io) {
  using S = Sched<LOG_T, KRT>;
  if constexpr (r < rend) {
    constexpr int K = S::template k<DIR, r>();
    constexpr int B0 = S::template b0<DIR, r>();
    run_round<F, DIR, LOG_T, CONTIG, K, B0, r == 0, r == S::NR - 1, LT>(tile, log_w, lanes, tw, io);
    if constexpr (r + 1 < rend) __syncthreads();
    rounds<F, DIR, LOG_T, CONTIG, r + 1, rend, LT, KRT>(tile, log_w, lanes, tw, io);
  }
}

// LDS bytes of a tuned pass: lane records (W x LaneInfo), root table (T x uint2,
// Montgomery fields) and tile (W x (T + T/32) words)
__host__ __device__ constexpr size_t lds_bytes(int field, int log_t, uint32_t log_w) {
  return (kLaneInfoBytes << log_w) + (field != 0 ? ((size_t)8 << log_t) : 0) +
         ((((size_t)1 << log_t) + ((size_t)1 << log_t) / 32) << log_w) * 4;
}

template <class F, int DIR, int LOG_T>
__device__ __forceinline__ typename TwSel<F, DIR, LOG_T>::type make_tw(const PassArgs &a, unsigned char *tabp) {
  typename TwSel<F, DIR, LOG_T>::type tw;
  if constexpr (F::MONTY) {
    tw.tab = reinterpret_cast<const uint2 *>(tabp);
  } else {
    tw.tbl = reinterpret_cast<const uint32_t *>(a.tw);
    tw.n_log = a.log_n;
    tw.lo = a.lo;
    tw.hi = a.lo + LOG_T;
    tw.hb = 0;
  }
  return tw;
}

// a.tw (Montgomery fields) = root table g_{2^11}^e, e < 2^11 (Shoup pairs)
template <int LOG_T>
__device__ __forceinline__ void load_root(const PassArgs &a, uint2 *tab) {
  const uint2 *root = reinterpret_cast<const uint2 *>(a.tw);
  for (uint32_t e = threadIdx.x; e < (1u << LOG_T); e += blockDim.x) tab[e] = root[e << (kMaxLogTile - LOG_T)];
}

// The half table g_T^e, e <= T/2, from the same root table.
template <int LOG_T>
__device__ __forceinline__ void load_root_half(const PassArgs &a, uint2 *tab) {
  const uint2 *root = reinterpret_cast<const uint2 *>(a.tw);
  for (uint32_t e = threadIdx.x; e <= (1u << (LOG_T - 1)); e += blockDim.x) tab[e] = root[e << (kMaxLogTile - LOG_T)];
}

// Common workgroup set-up: lane records in LDS, buffer resources rebased at
// lane 0 (decoded by every thread, so the bases stay scalar).
template <int LOG_T>
__device__ __forceinline__ void setup(const PassArgs &a, LaneInfo *lanes, RoundIO &io, uint32_t tile) {
  const uint32_t log_w = a.log_w;
  const uint64_t g0 = (uint64_t)tile << log_w;
  for (uint32_t l = threadIdx.x; l < (1u << log_w); l += blockDim.x) decode_lane(a, LOG_T, g0 + l, lanes[l]);
  LaneInfo l0;
  decode_lane(a, LOG_T, g0, l0);
  io.sbase0 = l0.sbase;
  io.dbase0 = l0.dbase;
  io.src = make_rsrc(a.src + l0.sbase, (a.src_elems - l0.sbase) * 4);
  io.dst = make_rsrc(a.dst + l0.dbase, (a.dst_elems - l0.dbase) * 4);
  io.srcp = a.src + l0.sbase;
  io.dstp = a.dst + l0.dbase;
  io.sps = (uint32_t)(((uint64_t)1 << a.lo) * a.src_rs);
  io.dps = (uint32_t)(((uint64_t)1 << a.lo) * a.dst_rs);
  io.lo = a.lo;
  io.low0 = l0.low;
  io.srev = a.src_rev_bits ? (uint32_t)(a.src_rs << (a.src_rev_bits - LOG_T)) : 0u;
  io.drev = a.dst_rev_bits ? (uint32_t)(a.dst_rs << (a.dst_rev_bits - LOG_T)) : 0u;
  io.ltw = io.lpost = nullptr;
  io.tps = a.tab_low_major ? 0u : a.lo;
  io.tls = a.tab_low_major ? (uint32_t)LOG_T : 0u;
  io.twist_gen = a.pow_m != nullptr;
  io.has_twist = a.twist_m != nullptr || io.twist_gen;
  if (a.twist_m) io.twist = make_rsrc(a.twist_m, (uint64_t)4 << (a.lo + LOG_T));
  if (io.twist_gen) io.pow = make_rsrc(a.pow_m, kPowWords * 4);
  io.has_scale = a.has_scale;
  io.scale_w = a.scale_w;
  io.scale_q = a.scale_q;
  io.has_pointf = io.has_post = false;
  io.post_gen = false;
}

// One tile (W lanes) of a tuned pass; any blockDim (threads loop over items).
// Montgomery-field twiddles come from an LDS root table of 2^ROOT_LOG entries
// (g_{2^ROOT_LOG}^e): the twiddle of layer t, position j is entry
// j << (ROOT_LOG - t - 1) whatever the sub-transform size, so one table of
// 2^11 entries serves every tile size.  The caller fills it before the call
// (the barrier after the lane set-up publishes it).
// LW >= 0: the launch's lanes per workgroup are 2^LW (a.log_w == LW), known at
// compile time, so every LDS slot index of the rounds folds to an immediate
// offset from one per-item base.
// KRT: layers per register round (5; 6 gives a 2^11-point pass two rounds,
// 6 + 5, instead of 5 + 5 + 1, for 64 values per item in the 6-layer round).
// HALF: the root table holds only e <= 2^(ROOT_LOG-1) (HalfRootTw).
template <class F, int DIR, int LOG_T, bool CONTIG, int ROOT_LOG, int LW = -1, int KRT = KR, bool HALF = false>
__device__ __forceinline__ void fast_tile(const PassArgs &a, uint32_t tile_id, LaneInfo *lanes, const uint2 *root,
                                          uint32_t *tile) {
  static_assert(ROOT_LOG >= LOG_T, "root table too small");
  const uint32_t log_w = LW >= 0 ? (uint32_t)LW : a.log_w;
  RoundIO io;
  setup<LOG_T>(a, lanes, io, tile_id);
  __syncthreads();
  if constexpr (F::MONTY) {
    std::conditional_t<HALF, HalfRootTw<ROOT_LOG, DIR>, RootTw<ROOT_LOG, DIR>> tw;
    tw.tab = root;
    rounds<F, DIR, LOG_T, CONTIG, 0, Sched<LOG_T, KRT>::NR, false, KRT>(tile, log_w, lanes, tw, io);
  } else {
    const auto tw = make_tw<F, DIR, LOG_T>(a, nullptr);
    rounds<F, DIR, LOG_T, CONTIG, 0, Sched<LOG_T, KRT>::NR, false, KRT>(tile, log_w, lanes, tw, io);
  }
}

// ---- two-round pass with a half-tile exchange ----
// A pass of two radix-32 rounds (LOG_T = 10) exchanges every point once
// between its rounds.  Here the exchange goes through half a tile in two
// phases: first the points whose top bit is 0, then those whose top bit is 1
// (point p sits at slot p mod T/2).  Where the top point bit is one of the
// item's own 32 points, a thread moves 16 values in each phase; where it is
// a bit of the item's base, the thread moves all 32 in one phase.  The
// workgroup holds half the LDS of the one-phase exchange, so three workgroups
// of 16 lanes fit on a CU instead of two (six waves per SIMD instead of four).
template <int LOG_T>
__host__ __device__ constexpr size_t half_tile_words(uint32_t log_w) {
  return (((size_t)1 << (LOG_T - 1)) + ((size_t)1 << (LOG_T - 1)) / 32) << log_w;
}

template <int LOG_T, int K, int B0, int H, class Fn>
__device__ __forceinline__ void half_slots(uint32_t pbase, Fn &&fn) {
  constexpr int R = 1 << K;
  if constexpr (B0 + K == LOG_T) {
    // top point bit = top bit of the item index: values [H * R/2, (H+1) * R/2)
#pragma unroll
    for (int i = H * (R / 2); i < (H + 1) * (R / 2); ++i) fn(i);
  } else {
    if (((pbase >> (LOG_T - 1)) & 1u) == (uint32_t)H) {
#pragma unroll
      for (int i = 0; i < R; ++i) fn(i);
    }
  }
}

template <class F, int DIR, int LOG_T, class TW>
__device__ __forceinline__ void half_exchange_pass(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes, TW tw,
                                                   const RoundIO &io) {
  using S = Sched<LOG_T, KR>;
  static_assert(S::NR == 2, "two rounds");
  constexpr int K0 = S::template k<DIR, 0>(), B00 = S::template b0<DIR, 0>();
  constexpr int K1 = S::template k<DIR, 1>(), B01 = S::template b0<DIR, 1>();
  static_assert(K0 == KR && K1 == KR, "radix-32 rounds");
  constexpr int R = 1 << KR;
  constexpr uint32_t HM = (1u << (LOG_T - 1)) - 1;  // slot of point p: p & HM
  // one item per thread in both rounds (blockDim == W << (LOG_T - 5)), lanes fastest
  const uint32_t W = 1u << log_w;
  const uint32_t lane = threadIdx.x & (W - 1);
  const uint32_t rb = threadIdx.x >> log_w;
  const LaneInfo &li = lanes[lane];
  if constexpr (!F::MONTY) tw.hb = li.hb;
  const uint32_t pb0 = round_pbase<K0, B00>(rb), pb1 = round_pbase<K1, B01>(rb);
  uint32_t v[R], u[R];
  load_item<F, DIR, LOG_T, true, K0, B00>(v, li, io, pb0);
  twist_item<F, DIR, LOG_T, K0, B00, false>(v, li, io, pb0);
  round_regs<F, DIR, K0, B00>(v, pb0, tw);
  auto slot = [&](uint32_t p) { const uint32_t q = p & HM; return ((q + (q >> 5)) << log_w) + lane; };
  half_slots<LOG_T, K0, B00, 0>(pb0, [&](int i) { tile[slot(pb0 | ((uint32_t)i << B00))] = v[i]; });
  __syncthreads();
  half_slots<LOG_T, K1, B01, 0>(pb1, [&](int i) { u[i] = tile[slot(pb1 | ((uint32_t)i << B01))]; });
  __syncthreads();  // every phase-0 value is read before the slots are reused
  half_slots<LOG_T, K0, B00, 1>(pb0, [&](int i) { tile[slot(pb0 | ((uint32_t)i << B00))] = v[i]; });
  __syncthreads();
  half_slots<LOG_T, K1, B01, 1>(pb1, [&](int i) { u[i] = tile[slot(pb1 | ((uint32_t)i << B01))]; });
  round_regs<F, DIR, K1, B01>(u, pb1, tw);
  store_item<F, DIR, LOG_T, true, K1, B01, false>(u, li, io, pb1);
}

template <class F, int DIR, int LOG_T, bool CONTIG, int LW = -1>
__global__ __launch_bounds__(512) void k_fast(const PassArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  // LDS: [tile][root table][lane records]: with the tile first, more of an
  // item's exchange slots are immediate offsets (< 64 KiB) from one base
  const uint32_t log_w = LW >= 0 ? (uint32_t)LW : a.log_w;
  uint32_t *tile = reinterpret_cast<uint32_t *>(smem);
  uint2 *root = reinterpret_cast<uint2 *>(smem + ((((size_t)1 << LOG_T) + ((size_t)1 << LOG_T) / 32) << log_w) * 4);
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(reinterpret_cast<unsigned char *>(root) +
                                                 (F::MONTY ? ((size_t)8 << LOG_T) : 0));
  if constexpr (F::MONTY) load_root<LOG_T>(a, root);
  fast_tile<F, DIR, LOG_T, CONTIG, LOG_T, LW>(a, xcd_tile(blockIdx.x, gridDim.x, a.xcd_share_log), lanes, root, tile);
}

// The same pass with the half-tile exchange (CONTIG, LOG_T = 10).
// LDS: [half tile][root table][lane records].
template <class F, int DIR, int LOG_T, int LW = -1>
__global__ __launch_bounds__(512) void k_fast_half(const PassArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t log_w = LW >= 0 ? (uint32_t)LW : a.log_w;  // a.log_w == LW when LW >= 0
  uint32_t *tile = reinterpret_cast<uint32_t *>(smem);
  uint2 *root = reinterpret_cast<uint2 *>(smem + ((half_tile_words<LOG_T>(log_w) * 4 + 15) & ~(size_t)15));
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(reinterpret_cast<unsigned char *>(root) +
                                                 (F::MONTY ? ((size_t)8 << LOG_T) : 0));
  if constexpr (F::MONTY) load_root<LOG_T>(a, root);
  RoundIO io;
  setup<LOG_T>(a, lanes, io, xcd_tile(blockIdx.x, gridDim.x, a.xcd_share_log));
  __syncthreads();
  if constexpr (F::MONTY) {
    RootTw<LOG_T, DIR> tw;
    tw.tab = root;
    half_exchange_pass<F, DIR, LOG_T>(tile, log_w, lanes, tw, io);
  } else {
    half_exchange_pass<F, DIR, LOG_T>(tile, log_w, lanes, make_tw<F, DIR, LOG_T>(a, nullptr), io);
  }
}

__host__ __device__ constexpr size_t lds_bytes_half(int field, int log_t, uint32_t log_w) {
  return (kLaneInfoBytes << log_w) + (field != 0 ? ((size_t)8 << log_t) : 0) +
         (((((((size_t)1 << (log_t - 1)) + ((size_t)1 << (log_t - 1)) / 32) << log_w) * 4) + 15) & ~(size_t)15);
}

// ---- coset LDE middle pass ----
// inverse layers [lo, log_n) (pre-twist on load), then for every output block
// b: x sigma_b^(p << lo), forward layers [log_n-1, lo], x post_b (forward
// twist x n^-1 x sigma_b^low) on store, where sigma_b = shift * zeta_b.
// The inverse's last round and the forward's first round own the same points
// (one item per thread), so the coefficients stay in registers across blocks.
// LT: the workgroup's lanes share their low row bits (columns run fastest), so
// the twist, point-factor and post tables it needs are T words each; they are
// staged in LDS ([tables: 2T words] between the root table and the tile)
// instead of being read per lane from global memory.
// CONTIG: lanes adjacent in memory (the middle layers of a multi-pass LDE);
// otherwise every lane is one whole column of points adjacent in memory (an
// LDE of at most 2^11 rows, SoA: lo == 0, 16-byte vector loads and stores).
// LDS of the middle pass: the tile (W x (T + T/32) words), half root table
// (T/2 + 1 Shoup pairs, padded to 16 bytes), LT tables (2T words) and lane records.
__host__ __device__ constexpr size_t mid_root_bytes(int log_t) {
  return ((((size_t)1 << (log_t - 1)) + 1) * 8 + 15) & ~(size_t)15;
}
// tile bytes padded to 16, so the tables after it take 16-byte accesses
__host__ __device__ constexpr size_t mid_tile_bytes(int log_t, uint32_t log_w) {
  return ((((((size_t)1 << log_t) + ((size_t)1 << log_t) / 32) << log_w) * 4) + 15) & ~(size_t)15;
}
__host__ __device__ constexpr size_t mid_lds_bytes(int log_t, uint32_t log_w, bool lt) {
  return mid_tile_bytes(log_t, log_w) + mid_root_bytes(log_t) + (lt ? ((size_t)8 << log_t) : 0) +
         (kLaneInfoBytes << log_w);
}

// 16 lanes (512 threads, two workgroups per CU): at most 128 VGPRs, so four
// waves per SIMD; otherwise three.
template <int LOG_T, bool LT, int LW = -1, bool CONTIG = true>
__global__ __launch_bounds__(512, LW == 4 ? 4 : 3) void k_mid_fast(const PassArgs a) {
  using F = BabyBear;
  using S = Sched<LOG_T, KR>;
  static_assert(S::NR >= 2, "middle pass needs two rounds");
  constexpr int K = KR;
  constexpr int B0 = LOG_T - KR;  // top chunk
  static_assert(S::template k<1, S::NR - 1>() == K && S::template b0<1, S::NR - 1>() == B0, "inverse last round");
  static_assert(S::template k<0, 0>() == K && S::template b0<0, 0>() == B0, "forward first round");
  constexpr int R = 1 << K;
  constexpr uint32_t T = 1u << LOG_T;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t log_w = LW >= 0 ? (uint32_t)LW : a.log_w;  // a.log_w == LW when LW >= 0
  // LDS: [tile][half root table][LT tables][lane records]: with the tile
  // first, every exchange slot of an item is an immediate offset (< 64 KiB)
  // from one per-thread base, so no address registers are held per slot
  uint32_t *tile = reinterpret_cast<uint32_t *>(smem);
  unsigned char *tabp = smem + mid_tile_bytes(LOG_T, log_w);
  uint32_t *tabs = reinterpret_cast<uint32_t *>(tabp + mid_root_bytes(LOG_T));
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(tabs + (LT ? 2 * T : 0));
  RoundIO io;
  setup<LOG_T>(a, lanes, io, xcd_tile(blockIdx.x, gridDim.x, a.xcd_share_log));
  load_root_half<LOG_T>(a, reinterpret_cast<uint2 *>(tabp));
  // LT tables: [point factors of the current block: T words][post of the
  // current block: T words]; the post half holds the inverse twist first.
  if constexpr (LT) {
    stage_words(tabs + T, a.twist_m + ((uint64_t)io.low0 << LOG_T), T);
    stage_words(tabs, a.pointf_m, T);
    io.ltw = tabs + T;
  }
  __syncthreads();
  HalfRootTw<LOG_T, 1> twi;
  HalfRootTw<LOG_T, 0> twf;
  twi.tab = twf.tab = reinterpret_cast<const uint2 *>(tabp);
  // inverse rounds 0 .. NR-2 (round 0 loads from HBM with the inverse twist)
  rounds<F, 1, LOG_T, CONTIG, 0, S::NR - 1, LT>(tile, log_w, lanes, twi, io);
  __syncthreads();
  // inverse last round: one item per thread (blockDim == W << (LOG_T - K))
  const uint32_t W = 1u << log_w;
  const uint32_t it = threadIdx.x;
  const uint32_t lane = CONTIG ? it & (W - 1) : it >> (LOG_T - K);
  const uint32_t rb = CONTIG ? it >> log_w : it & ((1u << (LOG_T - K)) - 1);
  const uint32_t pbase = round_pbase<K, B0>(rb);
  uint32_t c[R];
#pragma unroll
  for (int i = 0; i < R; ++i) c[i] = tile[lidx<LOG_T, CONTIG>(lane, pbase | (i << B0), log_w)];
  // unreduced: every coefficient is multiplied by a point factor next
  round_regs<F, 1, K, B0, true>(c, pbase, twi);
  const uint64_t n = 1ull << a.log_n;
  io.has_twist = false;  // the forward twist is part of post_b
  io.has_post = true;
  // LT: the next block's tables are written between the two barriers of the
  // current block's exchange (by then every thread has used the current point
  // factors, and the previous post table); when one 16-byte word per thread
  // covers a table they are fetched into registers a round ahead.
  const uint32_t e4 = threadIdx.x * 4;
  const bool pre = LT && blockDim.x * 4 >= T;
  uint4 po = make_uint4(0, 0, 0, 0), pf = po;
  const uint32_t *post_lane = a.post_m + ((uint64_t)io.low0 << LOG_T);
  if (pre && e4 < T) {
    po = *reinterpret_cast<const uint4 *>(post_lane + e4);
    if (a.blocks > 1) pf = *reinterpret_cast<const uint4 *>(a.pointf_m + T + e4);
  }
  for (uint32_t b = 0; b < a.blocks; ++b) {
    // point factors of block b, stored item-major: entry pbase * R + i is the
    // factor of point pbase | (i << B0), so an item reads R consecutive words
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = c[i];
    if constexpr (LT) {
      mul_lds_table<F, R, 0>(v, tabs + pbase * R, 0);
    } else {
      const Rsrc pf_r = make_rsrc(a.pointf_m + ((uint64_t)b << LOG_T), (uint64_t)4 << LOG_T);
      mul_table<F, R>(v, pf_r, pbase * (4u * R), 4u);
      if (a.lanef_m) {
        io.lanef = make_rsrc(a.lanef_m + ((uint64_t)b << a.lo), (uint64_t)4 << a.lo);
        io.post_gen = true;
      } else {
        io.post = make_rsrc(a.post_m + b * n, n * 4);
      }
    }
    round_regs<F, 0, K, B0>(v, pbase, twf);
    __syncthreads();  // every thread is done reading the tile (and the tables in use)
    {
      // the same slots as lidx, from a lane and pbase re-derived in the loop
      // with their known-zero bits (W - 1 and the item's own point bits), so
      // the 32 slots are immediate offsets from one base
      const uint32_t ln = loop_local(lane) & (W - 1);
      const uint32_t pb = loop_local(pbase) & ((T - 1) & ~((uint32_t)(R - 1) << B0));
#pragma unroll
      for (int i = 0; i < R; ++i) tile[lidx<LOG_T, CONTIG>(ln, pb | (i << B0), log_w)] = v[i];
    }
    if constexpr (LT) {
      if (pre) {
        if (e4 < T) {
          *reinterpret_cast<uint4 *>(tabs + T + e4) = po;
          if (b + 1 < a.blocks) *reinterpret_cast<uint4 *>(tabs + e4) = pf;
          if (b + 1 < a.blocks) po = *reinterpret_cast<const uint4 *>(post_lane + (b + 1) * n + e4);
          if (b + 2 < a.blocks) pf = *reinterpret_cast<const uint4 *>(a.pointf_m + ((uint64_t)(b + 2) << LOG_T) + e4);
        }
      } else {
        stage_words(tabs + T, post_lane + b * n, T);
        if (b + 1 < a.blocks) stage_words(tabs, a.pointf_m + ((uint64_t)(b + 1) << LOG_T), T);
      }
      io.lpost = tabs + T;
    }
    __syncthreads();
    io.dst = make_rsrc(a.dst + b * a.dst_block_off + io.dbase0, (a.dst_elems - b * a.dst_block_off - io.dbase0) * 4);
    io.dstp = a.dst + b * a.dst_block_off + io.dbase0;
    rounds<F, 0, LOG_T, CONTIG, 1, S::NR, LT>(tile, log_w, lanes, twf, io);
  }
}

// SFNTT_HALF_EXCHANGE=0 keeps the one-phase exchange (A/B measurement).
static bool half_exchange_on() {
  static const bool on = [] {
    const char *e = getenv("SFNTT_HALF_EXCHANGE");
    return !(e && e[0] == '0');
  }();
  return on;
}

// 16 lanes fixed at compile time in the half-tile exchange kernels (measured:
// LDE inverse pass 502-515 -> 480 us, forward unchanged); SFNTT_HALF_LW4=0 keeps
// the runtime lane count (A/B measurement)
static bool half_lw4_on() {
  static const bool on = [] {
    const char *e = getenv("SFNTT_HALF_LW4");
    return !(e && e[0] == '0');
  }();
  return on;
}

template <class F, int DIR, int LOG_T, bool CONTIG>
static hipError_t launch_t(const PassArgs &a, uint64_t tiles, hipStream_t s) {
  const size_t lds = lds_bytes(F::ID, LOG_T, a.log_w);
  const uint64_t threads = (uint64_t)1 << (a.log_w + LOG_T - KR);
  if (lds > kMaxLdsBytes || threads > 512 || tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  // forward Montgomery-field 2^10-point passes over 16 adjacent lanes (the
  // coset LDE's last pass, batched forward transforms): lane count fixed at
  // compile time.  Measured: 3.5 % faster there; the inverse and M31 passes
  // ran 6-7 % slower with it, so they keep the runtime lane count.
  constexpr bool kLw4 = CONTIG && LOG_T == 10 && F::MONTY && DIR == 0;
  auto kern = k_fast<F, DIR, LOG_T, CONTIG>;
  auto kern4 = k_fast<F, DIR, LOG_T, CONTIG, kLw4 ? 4 : -1>;
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    (void)hipFuncSetAttribute((const void *)kern4, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  if constexpr (CONTIG && LOG_T == 10) {
    // lanes adjacent in memory, 2^10-point tiles: half-tile exchange
    if (half_exchange_on()) {
      static const bool configured_h = [&] {
        (void)hipFuncSetAttribute((const void *)k_fast_half<F, DIR, LOG_T>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
        (void)hipFuncSetAttribute((const void *)k_fast_half<F, DIR, LOG_T, 4>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
        return true;
      }();
      (void)configured_h;
      const size_t lds_h = lds_bytes_half(F::ID, LOG_T, a.log_w);
      if (a.log_w == 4 && half_lw4_on())
        hipLaunchKernelGGL((k_fast_half<F, DIR, LOG_T, 4>), dim3((uint32_t)tiles), dim3((uint32_t)threads), lds_h, s,
                           a);
      else
        hipLaunchKernelGGL((k_fast_half<F, DIR, LOG_T>), dim3((uint32_t)tiles), dim3((uint32_t)threads), lds_h, s, a);
      return hipGetLastError();
    }
  }
  if (kLw4 && a.log_w == 4) hipLaunchKernelGGL(kern4, dim3((uint32_t)tiles), dim3((uint32_t)threads), lds, s, a);
  else hipLaunchKernelGGL(kern, dim3((uint32_t)tiles), dim3((uint32_t)threads), lds, s, a);
  return hipGetLastError();
}

template <class F, int DIR>
static hipError_t launch_fd(int log_t, bool contig, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  switch (log_t) {
#define SFN_CASE(L)                                                   \
  case L:                                                             \
    return contig ? launch_t<F, DIR, L, true>(a, tiles, s) : launch_t<F, DIR, L, false>(a, tiles, s);
    SFN_CASE(6) SFN_CASE(7) SFN_CASE(8) SFN_CASE(9) SFN_CASE(10) SFN_CASE(11)
#undef SFN_CASE
    default: return hipErrorInvalidValue;
  }
}

}  // namespace fast

#if SFN_IN_PART(0)
hipError_t launch_fast_m31_fwd(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<M31, 0>(lt, c, a, t, s); }
#endif
#if SFN_IN_PART(1)
hipError_t launch_fast_m31_inv(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<M31, 1>(lt, c, a, t, s); }
#endif
#if SFN_IN_PART(2)
hipError_t launch_fast_bb_fwd(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<BabyBear, 0>(lt, c, a, t, s); }
#endif
#if SFN_IN_PART(3)
hipError_t launch_fast_bb_inv(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<BabyBear, 1>(lt, c, a, t, s); }
#endif
#if SFN_IN_PART(6)
hipError_t launch_fast_kb_fwd(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<KoalaBear, 0>(lt, c, a, t, s); }
#endif
#if SFN_IN_PART(7)
hipError_t launch_fast_kb_inv(int lt, bool c, const PassArgs &a, uint64_t t, hipStream_t s) { return fast::launch_fd<KoalaBear, 1>(lt, c, a, t, s); }
#endif

#if SFN_IN_PART(4)
hipError_t launch_fast_m31_fwd(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_m31_inv(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_bb_fwd(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_bb_inv(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_kb_fwd(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_kb_inv(int, bool, const PassArgs &, uint64_t, hipStream_t);

size_t fast_lds_bytes(int field, int log_t, uint32_t log_w) { return fast::lds_bytes(field, log_t, log_w); }

hipError_t launch_fast(int field, int dir, int log_t, bool contig, const PassArgs &a, uint64_t tiles,
                       hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  if (field == 0) return dir == 0 ? launch_fast_m31_fwd(log_t, contig, a, tiles, s)
                                  : launch_fast_m31_inv(log_t, contig, a, tiles, s);
  if (field == 2) return dir == 0 ? launch_fast_kb_fwd(log_t, contig, a, tiles, s)
                                  : launch_fast_kb_inv(log_t, contig, a, tiles, s);
  return dir == 0 ? launch_fast_bb_fwd(log_t, contig, a, tiles, s) : launch_fast_bb_inv(log_t, contig, a, tiles, s);
}

template <int LOG_T, bool LT, bool CONTIG>
static hipError_t launch_mid_fast_lt(const PassArgs &a, uint64_t tiles, size_t lds, hipStream_t s) {
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)fast::k_mid_fast<LOG_T, LT, -1, CONTIG>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    (void)hipFuncSetAttribute((const void *)fast::k_mid_fast<LOG_T, LT, 3, CONTIG>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    (void)hipFuncSetAttribute((const void *)fast::k_mid_fast<LOG_T, LT, 4, CONTIG>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  const uint64_t threads = (uint64_t)1 << (a.log_w + LOG_T - fast::KR);
  // eight or sixteen adjacent lanes (2^10-point tiles): lane count fixed at compile time
  if (CONTIG && LOG_T == 10 && a.log_w == 3)
    hipLaunchKernelGGL((fast::k_mid_fast<LOG_T, LT, 3, CONTIG>), dim3((uint32_t)tiles), dim3((uint32_t)threads), lds,
                       s, a);
  else if (CONTIG && LOG_T == 10 && a.log_w == 4)
    hipLaunchKernelGGL((fast::k_mid_fast<LOG_T, LT, 4, CONTIG>), dim3((uint32_t)tiles), dim3((uint32_t)threads), lds,
                       s, a);
  else
    hipLaunchKernelGGL((fast::k_mid_fast<LOG_T, LT, -1, CONTIG>), dim3((uint32_t)tiles), dim3((uint32_t)threads),
                       lds, s, a);
  return hipGetLastError();
}

// LDS-staged tables when the lanes share their low row bits (tab_low_major)
// and the extra 2T words fit; otherwise per-lane global table reads.
template <int LOG_T>
static hipError_t launch_mid_fast_t(bool contig, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  const size_t lds = fast::mid_lds_bytes(LOG_T, a.log_w, false);
  const uint64_t threads = (uint64_t)1 << (a.log_w + LOG_T - fast::KR);
  if (lds > kMaxLdsBytes || threads > 512 || tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  if (!contig) return launch_mid_fast_lt<LOG_T, false, false>(a, tiles, lds, s);
  const size_t lds_lt = fast::mid_lds_bytes(LOG_T, a.log_w, true);
  if (a.tab_low_major && lds_lt <= kMaxLdsBytes) return launch_mid_fast_lt<LOG_T, true, true>(a, tiles, lds_lt, s);
  return launch_mid_fast_lt<LOG_T, false, true>(a, tiles, lds, s);
}

hipError_t launch_lde_mid_fast(int log_t, bool contig, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  switch (log_t) {
    case 6: return launch_mid_fast_t<6>(contig, a, tiles, s);
    case 7: return launch_mid_fast_t<7>(contig, a, tiles, s);
    case 8: return launch_mid_fast_t<8>(contig, a, tiles, s);
    case 9: return launch_mid_fast_t<9>(contig, a, tiles, s);
    case 10: return launch_mid_fast_t<10>(contig, a, tiles, s);
    case 11: return launch_mid_fast_t<11>(contig, a, tiles, s);
    default: return hipErrorInvalidValue;
  }
}
#endif  // SFN_IN_PART(4)

#if SFN_IN_PART(5)
// ---- many jobs in one launch, tuned tiles ----
// Segments (one per job pass) own consecutive ranges of a global tile index.
// A persistent grid (a multiple of 8 workgroups) walks the tiles with a grid
// stride, so every wave reaches the end of the list and the launch drains.
// Global tile ids are permuted within groups of 8 * 2^xs like xcd_tile, so
// tiles that share 128-byte lines run on one XCD.  Field, direction, tile size
// and lane order are selected per segment (uniform across the workgroup).
namespace fast {
template <class F, int DIR, bool CONTIG>
__device__ __forceinline__ void multi_tile(uint32_t log_t, const PassArgs &a, uint32_t tile, LaneInfo *lanes,
                                           const uint2 *root, uint32_t *tile_lds) {
  switch (log_t) {
#define SFN_CASE(L) \
  case L: fast_tile<F, DIR, L, CONTIG, kMaxLogTile, -1, KR, true>(a, tile, lanes, root, tile_lds); break;
    SFN_CASE(6) SFN_CASE(7) SFN_CASE(8) SFN_CASE(9) SFN_CASE(10) SFN_CASE(11)
#undef SFN_CASE
    default: break;
  }
}

// LDS: [half root table of the launch's Montgomery field (BabyBear or
// KoalaBear; one per launch): g^e for e <= 2^10 of the 2^11-entry table, loaded
// once][lane records: lanes_bytes][tile].  Every tile runs radix-32 rounds (a
// 2^11-point tile three, 5 + 5 + 1 layers), which keeps the kernel at 128
// VGPRs: 512 threads per workgroup and two workgroups per CU, four waves per
// SIMD (measured on 512 uneven columns: 2.64 -> 2.44 ms against 256 threads with
// a 6 + 5 round schedule for 2^11, 144 VGPRs and three waves per SIMD).
__global__ __launch_bounds__(kMultiFastThreads, 4) void k_multi_fast(const SegDesc *segs, uint32_t nseg,
                                                                  uint64_t total_tiles, uint32_t xs,
                                                                  const uint2 *root_g, uint32_t lanes_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint2 *root = reinterpret_cast<uint2 *>(smem);
  const size_t root_bytes = root_g ? kMultiRootBytes : 0;
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(smem + root_bytes);
  uint32_t *tile_lds = reinterpret_cast<uint32_t *>(smem + root_bytes + lanes_bytes);
  if (root_g)
    for (uint32_t e = threadIdx.x; e <= (1u << (kMaxLogTile - 1)); e += blockDim.x) root[e] = root_g[e];
  const uint64_t gs = 8ull << xs;
  const uint64_t full = total_tiles - total_tiles % gs;
  for (uint64_t t0 = blockIdx.x; t0 < total_tiles; t0 += gridDim.x) {
    uint64_t t = t0;
    if (xs && t0 < full) {
      const uint64_t r = t0 & (gs - 1);
      t = (t0 - r) + ((r & 7u) << xs) + (r >> 3);
    }
    uint32_t lo = 0, hi = nseg;  // segs[lo].tile_begin <= t < segs[hi].tile_begin
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (segs[mid].tile_begin <= t) lo = mid;
      else hi = mid;
    }
    const PassArgs a = segs[lo].a;
    const uint32_t local = (uint32_t)(t - segs[lo].tile_begin);
    const uint32_t log_t = segs[lo].log_t;
    switch (segs[lo].kind) {  // field * 2 + dir, + 6 when lanes are adjacent in memory
      case 0: multi_tile<M31, 0, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 1: multi_tile<M31, 1, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 2: multi_tile<BabyBear, 0, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 3: multi_tile<BabyBear, 1, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 4: multi_tile<KoalaBear, 0, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 5: multi_tile<KoalaBear, 1, false>(log_t, a, local, lanes, root, tile_lds); break;
      case 6: multi_tile<M31, 0, true>(log_t, a, local, lanes, root, tile_lds); break;
      case 7: multi_tile<M31, 1, true>(log_t, a, local, lanes, root, tile_lds); break;
      case 8: multi_tile<BabyBear, 0, true>(log_t, a, local, lanes, root, tile_lds); break;
      case 9: multi_tile<BabyBear, 1, true>(log_t, a, local, lanes, root, tile_lds); break;
      case 10: multi_tile<KoalaBear, 0, true>(log_t, a, local, lanes, root, tile_lds); break;
      case 11: multi_tile<KoalaBear, 1, true>(log_t, a, local, lanes, root, tile_lds); break;
      default: break;
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}
}  // namespace fast

hipError_t launch_multi_fast(const SegDesc *d_segs, uint32_t nseg, uint64_t total_tiles, size_t lds, uint32_t grid,
                             uint32_t xs, const void *bb_root, uint32_t lanes_bytes, hipStream_t s) {
  if (total_tiles == 0) return hipSuccess;
  if (lds > kMaxLdsBytes || grid == 0 || xs > 3) return hipErrorInvalidConfiguration;
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)fast::k_multi_fast, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  hipLaunchKernelGGL(fast::k_multi_fast, dim3(grid), dim3(kMultiFastThreads), lds, s, d_segs, nseg, total_tiles, xs,
                     reinterpret_cast<const uint2 *>(bb_root), lanes_bytes);
  return hipGetLastError();
}
#endif  // SFN_IN_PART(5)

}  // namespace sfn
