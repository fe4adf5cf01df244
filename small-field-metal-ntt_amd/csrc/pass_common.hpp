// Pieces shared by the pass kernels: round schedule, lane decoding, register
// rounds and twiddle fetchers.
#pragma once
#include "field.hpp"
#include "kernels.hpp"

namespace sfn {

template <int LOG_T>
struct Tile {
  static constexpr int T = 1 << LOG_T;
  static constexpr int TP = T + (T >> 5);  // lane pitch: one pad word per 32 points
  __device__ static __forceinline__ int idx(int p) { return p + (p >> 5); }
};

// Round schedule: forward covers bits top-down in chunks of LOGR with the
// partial chunk last (lowest bits); inverse covers bits bottom-up with the
// partial chunk first, so the inverse's last round and the forward's first
// round own the same points.
template <int LOG_T, int LOGR>
struct Sched {
  static constexpr int NR = (LOG_T + LOGR - 1) / LOGR;
  static constexpr int REM = LOG_T - (NR - 1) * LOGR;
  template <int DIR, int r>
  static constexpr int k() {
    return DIR == 0 ? (r == NR - 1 ? REM : LOGR) : (r == 0 ? REM : LOGR);
  }
  template <int DIR, int r>
  static constexpr int b0() {
    return DIR == 0 ? (r == NR - 1 ? 0 : LOG_T - (r + 1) * LOGR) : (r == 0 ? 0 : REM + (r - 1) * LOGR);
  }
};

struct LaneInfo {
  uint64_t sbase, dbase;
  uint64_t hb;
  uint32_t low, valid;
};

// floor(n / d) for n < 2^32 from m = ceil(2^64 / d), d >= 2: (m * n) >> 64, exact
// for every such n (Lemire, Kaser and Kurz, "Faster remainder by direct
// computation", 2019, theorem 1 with F = 64, N = 32).  m * n = m_hi * n * 2^32 +
// m_lo * n, and the low 32 bits of m_lo * n cannot carry into bit 64.
__device__ __forceinline__ uint32_t div_by_recip(uint32_t n, uint64_t m) {
  const uint64_t t = (uint64_t)(uint32_t)(m >> 32) * n + __umulhi((uint32_t)m, n);
  return (uint32_t)(t >> 32);
}

// x / batch and x % batch
__device__ __forceinline__ void div_batch(const PassArgs &a, uint64_t x, uint64_t &q, uint32_t &r) {
  if (a.batch == 1) {
    q = x;
    r = 0;
  } else if (a.batch_recip) {  // x < 2^32 (the host checked instances <= 2^32)
    const uint32_t qq = div_by_recip((uint32_t)x, a.batch_recip);
    q = qq;
    r = (uint32_t)x - qq * a.batch;
  } else {
    q = x / a.batch;
    r = (uint32_t)(x % a.batch);
  }
}

__device__ __forceinline__ void decode_lane(const PassArgs &a, int log_t, uint64_t g, LaneInfo &li) {
  const uint32_t lo = a.lo;
  const uint32_t hi = lo + log_t;
  const uint32_t lo_mask = (1u << lo) - 1;
  uint32_t c, low;
  uint64_t hb;
  li.valid = g < a.instances;
  if (!li.valid) g = 0;
  if (a.order == kLowFastest) {
    low = (uint32_t)g & lo_mask;
    div_batch(a, g >> lo, hb, c);
  } else if (a.order == kColFastest) {
    uint64_t r;
    div_batch(a, g, r, c);
    low = (uint32_t)r & lo_mask;
    hb = r >> lo;
  } else {
    const uint32_t nhb_log = a.log_n - hi;
    hb = g & ((1ull << nhb_log) - 1);
    uint64_t r = g >> nhb_log;
    low = 0;
    c = (uint32_t)r;
  }
  const uint64_t row0 = (hb << hi) | low;
  // logical row r of a bit-reversed array lives at physical row bitrev(r)
  const uint64_t srow = a.src_rev_bits ? (uint64_t)(__brev((uint32_t)row0) >> (32 - a.src_rev_bits)) : row0;
  const uint64_t drow = a.dst_rev_bits ? (uint64_t)(__brev((uint32_t)row0) >> (32 - a.dst_rev_bits)) : row0;
  li.sbase = (uint64_t)c * a.src_cs + srow * a.src_rs;
  li.dbase = (uint64_t)c * a.dst_cs + drow * a.dst_rs;
  li.hb = hb;
  li.low = low;
}

// ---- twiddle fetch inside a register round ----
// Montgomery fields: small position table staged in LDS (layer t at offset 2^t - 1).
// M31: block table in global memory, index (hb << (hi-l-1)) | (p >> (t+1)).
template <class F, bool MONTY = F::MONTY>
struct TwCtx;

template <class F>
struct TwCtx<F, true> {
  static constexpr bool kNeg = false;
  static constexpr bool kFoldRoots = false;
  static constexpr bool kSplitNeg = false;
  const uint2 *stw;  // LDS
  __device__ __forceinline__ uint2 get(int t, uint32_t pos) const { return stw[(1u << t) - 1 + pos]; }
};

template <>
struct TwCtx<M31, false> {
  static constexpr bool kNeg = false;
  static constexpr bool kFoldRoots = false;
  static constexpr bool kSplitNeg = false;
  const uint32_t *tbl;  // global, whole transform
  uint32_t n_log, lo, hi;
  uint64_t hb;
  // layer t of the sub-transform (global layer lo + t), block index inside the sub-transform
  __device__ __forceinline__ uint32_t get(int t, uint32_t blk) const {
    const uint32_t l = lo + t;
    const uint64_t off = (1ull << n_log) - (1ull << (n_log - l));
    return __ldg(tbl + off + ((hb << (hi - l - 1)) | blk));
  }
  // get(t, base | m) for a base whose bits below m's range are clear: the
  // words of one layer are then one per-thread address plus constant offsets m
  // (folded into the loads), instead of an address computed per word.
  __device__ __forceinline__ uint32_t get_split(int t, uint32_t base, uint32_t m) const {
    const uint32_t l = lo + t;
    const uint64_t off = (1ull << n_log) - (1ull << (n_log - l));
    const uint32_t *p = tbl + off + (hb << (hi - l - 1)) + base;
    return __ldg(p + m);
  }
};

// One register round: K layers on v[0..2^K) holding points pbase | (i << B0).
// A fetcher with kNeg returns the negated twiddle -w for the inverse (CT)
// butterflies, which then compute (x - (-w)y, x + (-w)y) = (x + wy, x - wy).
// LAZY_OUT (Montgomery fields): the round's last layer leaves its outputs
// unreduced, in [0, 2p), for a caller that multiplies every output by a
// Montgomery table factor next (mul_mont takes any 32-bit value).
template <class F, int DIR, int K, int B0, bool LAZY_OUT = false, class TW>
__device__ __forceinline__ void round_regs(uint32_t (&v)[1 << K], uint32_t pbase, const TW &tw) {
  static_assert(!LAZY_OUT || F::MONTY, "unreduced outputs: Montgomery fields only");
  constexpr int R = 1 << K;
  constexpr bool GS = F::MONTY == (DIR == 0);  // Montgomery fwd and M31 inv are GS
  if constexpr (DIR == 0) {
#pragma unroll
    for (int j = K - 1; j >= 0; --j) {
#pragma unroll
      for (int m = 0; m < (1 << (F::MONTY ? j : (K - 1 - j))); ++m) {
        typename F::Tw w;
        if constexpr (F::MONTY) {
          // a round over the lowest point bits: every position is a constant
          if constexpr (B0 == 0 && TW::kFoldRoots) w = TW::template cget<F>(j, (uint32_t)m);
          else w = tw.get(B0 + j, (pbase & ((1u << B0) - 1)) | ((uint32_t)m << B0));
        } else {
          // pbase >> (B0 + j + 1) is zero below bit K - 1 - j, and m < 2^(K - 1 - j)
          w = tw.get_split(B0 + j, pbase >> (B0 + j + 1), (uint32_t)m);
        }
#pragma unroll
        for (int q = 0; q < (1 << (F::MONTY ? (K - 1 - j) : j)); ++q) {
          // BB: twiddle depends on i mod 2^j -> i = m + q*2^(j+1)
          // M31: twiddle depends on i >> (j+1) -> i = (m << (j+1)) + q
          const int i = F::MONTY ? (m + (q << (j + 1))) : ((m << (j + 1)) + q);
          const bool lz = LAZY_OUT && j == 0;
          if constexpr (F::MONTY && B0 == 0) {
            if (m == 0) {  // position 0 of every layer: the twiddle is 1
              if constexpr (LAZY_OUT) bfly_unit_l<F>(v[i], v[i + (1 << j)], lz, lz);
              else bfly_unit<F>(v[i], v[i + (1 << j)]);
              continue;
            }
          }
          if constexpr (GS && LAZY_OUT) {
            if (lz) bfly_gs_lazy<F>(v[i], v[i + (1 << j)], w);
            else bfly_gs<F>(v[i], v[i + (1 << j)], w);
          }
          else if constexpr (GS) bfly_gs<F>(v[i], v[i + (1 << j)], w);
          else bfly_ct<F>(v[i], v[i + (1 << j)], w);
        }
      }
    }
  } else {
    // Montgomery CT layers: both outputs of butterfly (i, i + 2^j) are the
    // multiplied operands of layer j + 1 when bit j + 1 of i is set, and then
    // stay unreduced (bfly_ct_l) unless that product is a twiddle-1 butterfly
    // (position m = 0 in a round over the lowest point bits, only possible for
    // the x output).  The round's last layer leaves canonical values.
    constexpr bool kLazy = F::MONTY && !GS;
#pragma unroll
    for (int j = 0; j < K; ++j) {
#pragma unroll
      for (int m = 0; m < (1 << (F::MONTY ? j : (K - 1 - j))); ++m) {
        typename F::Tw w;
        if constexpr (F::MONTY) {
          if constexpr (B0 == 0 && TW::kFoldRoots) w = TW::template cget<F>(j, (uint32_t)m);
          else if constexpr (TW::kSplitNeg)
            w = tw.template get_split<B0>(B0 + j, pbase & ((1u << B0) - 1), (uint32_t)m, (1u << j) - 1);
          else w = tw.get(B0 + j, (pbase & ((1u << B0) - 1)) | ((uint32_t)m << B0));
        } else {
          // pbase >> (B0 + j + 1) is zero below bit K - 1 - j, and m < 2^(K - 1 - j)
          w = tw.get_split(B0 + j, pbase >> (B0 + j + 1), (uint32_t)m);
        }
#pragma unroll
        for (int q = 0; q < (1 << (F::MONTY ? (K - 1 - j) : j)); ++q) {
          const int i = F::MONTY ? (m + (q << (j + 1))) : ((m << (j + 1)) + q);
          const bool last = LAZY_OUT && j + 1 == K;
          const bool ly = last || (kLazy && j + 1 < K && ((i >> (j + 1)) & 1) != 0);
          const bool lx = last || (ly && !(B0 == 0 && m == 0));
          if constexpr (F::MONTY && B0 == 0) {
            if (m == 0) {
              if constexpr (kLazy) bfly_unit_l<F>(v[i], v[i + (1 << j)], lx, ly);
              else bfly_unit<F>(v[i], v[i + (1 << j)]);
              continue;
            }
          }
          if constexpr (GS) bfly_gs<F>(v[i], v[i + (1 << j)], w);
          else if constexpr (kLazy && TW::kNeg) bfly_ct_neg_l<F>(v[i], v[i + (1 << j)], w, lx, ly);
          else if constexpr (kLazy) bfly_ct_l<F>(v[i], v[i + (1 << j)], w, lx, ly);
          else if constexpr (TW::kNeg) bfly_ct_neg<F>(v[i], v[i + (1 << j)], w);
          else bfly_ct<F>(v[i], v[i + (1 << j)], w);
        }
      }
    }
  }
  (void)R;
}

// Tile of workgroup g.  Workgroup ids go to the 8 XCDs round-robin (id mod 8)
// and every XCD has its own L2, so 2^s consecutive tiles that share 128-byte
// lines are given ids that land on one XCD: in each group of 8 * 2^s ids,
// id r -> tile (r mod 8) * 2^s + r / 8.  Ids past the last full group keep
// their tile.  s = 0 is the identity.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t g, uint32_t grid, uint32_t s) {
  const uint32_t gs = 8u << s;
  if (s == 0 || g >= grid - grid % gs) return g;
  const uint32_t r = g & (gs - 1);
  return (g - r) + ((r & 7u) << s) + (r >> 3);
}

// Returns x unchanged as a value the compiler cannot move out of the
// enclosing loop (an empty asm statement).  Address arithmetic on it then stays
// next to its loads and stores and folds into their immediate offsets, where
// loop-invariant code motion would otherwise hold one register per address.
__device__ __forceinline__ uint32_t loop_local(uint32_t x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <int K, int B0>
__device__ __forceinline__ uint32_t round_pbase(uint32_t rb) {
  return (rb & ((1u << B0) - 1)) | ((rb >> B0) << (B0 + K));
}


}  // namespace sfn
