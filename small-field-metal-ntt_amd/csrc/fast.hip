// Tuned pass kernels (radix-32 register blocks, LOG_T in [6, 11]).
//
// Compared with the generic LDS-staged pass (kernels.hip) these:
//  - load the first register round straight from HBM into registers and store
//    the last round straight from registers, so a 2^10-point pass touches LDS
//    once per element (one exchange) and each thread has 32 independent loads
//    in flight;
//  - run exactly one item (one lane x 32 points) per thread in the full
//    rounds (blockDim = W * 2^(LOG_T-5));
//  - keep one root table g_T^e (e < T) in LDS that serves every layer and both
//    directions (layer t, position j uses e = j * 2^(LOG_T-t-1), negated for
//    the inverse);
//  - lay LDS out so the exchanges are bank-conflict free: lane-fastest
//    (p + p/32) * W + lane when lanes are adjacent in memory, lane-major
//    lane * (T + T/32) + p + p/32 when points are.
#include "pass_common.hpp"

namespace sfn {

#ifndef SFN_PART
#define SFN_PART -1
#endif
#define SFN_IN_PART(k) (SFN_PART < 0 || SFN_PART == (k))

namespace fast {

constexpr int KR = 5;

// BabyBear twiddles from the LDS root table.
template <int LOG_T, int DIR>
struct RootTw {
  const uint2 *tab;
  __device__ __forceinline__ uint2 get(int t, uint32_t pos) const {
    uint32_t e = pos << (LOG_T - t - 1);
    if (DIR == 1) e = (0u - e) & ((1u << LOG_T) - 1);
    return tab[e];
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

// Per-launch state shared by the rounds.
struct RoundIO {
  const uint32_t *src;
  uint32_t *dst;
  uint64_t sps, dps;    // element stride between consecutive points
  uint64_t dst_off;     // extra offset added to every store (LDE output block)
  const uint2 *twist_in, *twist_out;  // BabyBear four-step twists (or null)
  uint32_t lo;
  uint32_t has_scale, scale_w, scale_q;
};

// Round r of a pass: loads from HBM when FIRST, else from LDS; stores to HBM
// when LAST, else to LDS.  Items: W << (LOG_T - K).
template <class F, int DIR, int LOG_T, bool CONTIG, int K, int B0, bool FIRST, bool LAST, class TW>
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
      if (li.valid) {
        if constexpr (!CONTIG && B0 == 0 && K >= 2) {
          // R consecutive words per thread (points are adjacent rows)
          const uint4 *s4 = reinterpret_cast<const uint4 *>(io.src + li.sbase + pbase);
#pragma unroll
          for (int i = 0; i < R / 4; ++i) {
            const uint4 q = s4[i];
            v[4 * i] = q.x; v[4 * i + 1] = q.y; v[4 * i + 2] = q.z; v[4 * i + 3] = q.w;
          }
        } else {
#pragma unroll
          for (int i = 0; i < R; ++i) v[i] = io.src[li.sbase + (uint64_t)(pbase | (i << B0)) * io.sps];
        }
        if constexpr (F::ID == 1 && DIR == 1) {
          if (io.twist_in) {
#pragma unroll
            for (int i = 0; i < R; ++i)
              v[i] = BabyBear::mul_tw(v[i], io.twist_in[((uint64_t)(pbase | (i << B0)) << io.lo) | li.low]);
          }
        }
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = 0;
      }
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) v[i] = tile[lidx<LOG_T, CONTIG>(lane, pbase | (i << B0), log_w)];
    }
    if constexpr (F::ID == 0) tw.hb = li.hb;
    round_regs<F, DIR, K, B0>(v, pbase, tw);
    if constexpr (LAST) {
      if (!li.valid) continue;
      if constexpr (F::ID == 1 && DIR == 0) {
        if (io.twist_out) {
#pragma unroll
          for (int i = 0; i < R; ++i)
            v[i] = BabyBear::mul_tw(v[i], io.twist_out[((uint64_t)(pbase | (i << B0)) << io.lo) | li.low]);
        }
      }
      if (io.has_scale) {
#pragma unroll
        for (int i = 0; i < R; ++i) {
          if constexpr (F::ID == 1) v[i] = BabyBear::mul_tw(v[i], make_uint2(io.scale_w, io.scale_q));
          else v[i] = M31::mul(v[i], io.scale_w);
        }
      }
      if constexpr (!CONTIG && B0 == 0 && K >= 2) {
        uint4 *d4 = reinterpret_cast<uint4 *>(io.dst + io.dst_off + li.dbase + pbase);
#pragma unroll
        for (int i = 0; i < R / 4; ++i) d4[i] = make_uint4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
      } else {
#pragma unroll
        for (int i = 0; i < R; ++i) io.dst[io.dst_off + li.dbase + (uint64_t)(pbase | (i << B0)) * io.dps] = v[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < R; ++i) tile[lidx<LOG_T, CONTIG>(lane, pbase | (i << B0), log_w)] = v[i];
    }
  }
}

template <class F, int DIR, int LOG_T, bool CONTIG, int r, int rend, class TW>
__device__ __forceinline__ void rounds(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes, const TW &tw,
                                       const RoundIO &io) {
  using S = Sched<LOG_T, KR>;
  if constexpr (r < rend) {
    constexpr int K = S::template k<DIR, r>();
    constexpr int B0 = S::template b0<DIR, r>();
    run_round<F, DIR, LOG_T, CONTIG, K, B0, r == 0, r == S::NR - 1>(tile, log_w, lanes, tw, io);
    if constexpr (r + 1 < rend) __syncthreads();
    rounds<F, DIR, LOG_T, CONTIG, r + 1, rend>(tile, log_w, lanes, tw, io);
  }
}

// LDS: [LaneInfo x W][root table T x uint2 (BabyBear)][tile W x (T + T/32)]
__host__ __device__ constexpr size_t lds_bytes(int field, int log_t, uint32_t log_w) {
  return (kLaneInfoBytes << log_w) + (field == 1 ? ((size_t)8 << log_t) : 0) +
         ((((size_t)1 << log_t) + ((size_t)1 << log_t) / 32) << log_w) * 4;
}

template <class F, int DIR, int LOG_T>
__device__ __forceinline__ typename TwSel<F, DIR, LOG_T>::type make_tw(const PassArgs &a, unsigned char *tabp) {
  typename TwSel<F, DIR, LOG_T>::type tw;
  if constexpr (F::ID == 1) {
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

// a.tw (BabyBear) = root table g_{2^11}^e, e < 2^11 (Shoup pairs)
template <int LOG_T>
__device__ __forceinline__ void load_root(const PassArgs &a, uint2 *tab) {
  const uint2 *root = reinterpret_cast<const uint2 *>(a.tw);
  for (uint32_t e = threadIdx.x; e < (1u << LOG_T); e += blockDim.x) tab[e] = root[e << (kMaxLogTile - LOG_T)];
}

template <class F, int DIR, int LOG_T, bool CONTIG>
__global__ __launch_bounds__(512) void k_fast(const PassArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t log_w = a.log_w;
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(smem);
  unsigned char *tabp = smem + (kLaneInfoBytes << log_w);
  uint32_t *tile = reinterpret_cast<uint32_t *>(tabp + (F::ID == 1 ? ((size_t)8 << LOG_T) : 0));
  const uint64_t g0 = (uint64_t)blockIdx.x << log_w;
  for (uint32_t l = threadIdx.x; l < (1u << log_w); l += blockDim.x) decode_lane(a, LOG_T, g0 + l, lanes[l]);
  if constexpr (F::ID == 1) load_root<LOG_T>(a, reinterpret_cast<uint2 *>(tabp));
  __syncthreads();
  RoundIO io;
  io.src = a.src;
  io.dst = a.dst;
  io.sps = ((uint64_t)1 << a.lo) * a.src_rs;
  io.dps = ((uint64_t)1 << a.lo) * a.dst_rs;
  io.dst_off = 0;
  io.twist_in = (F::ID == 1 && DIR == 1) ? a.twist : nullptr;
  io.twist_out = (F::ID == 1 && DIR == 0) ? a.twist : nullptr;
  io.lo = a.lo;
  io.has_scale = a.has_scale;
  io.scale_w = a.scale_w;
  io.scale_q = a.scale_q;
  const auto tw = make_tw<F, DIR, LOG_T>(a, tabp);
  rounds<F, DIR, LOG_T, CONTIG, 0, Sched<LOG_T, KR>::NR>(tile, log_w, lanes, tw, io);
}

// ---- coset LDE middle pass ----
// inverse layers [lo, log_n) (pre-twist on load), x n^-1 (shift*zeta_b)^row,
// forward layers [log_n-1, lo] (post-twist on store) for every output block.
// The inverse's last round and the forward's first round own the same points
// (one item per thread), so the coefficients stay in registers across blocks.
template <int LOG_T>
__global__ __launch_bounds__(512) void k_mid_fast(const PassArgs a) {
  using F = BabyBear;
  using S = Sched<LOG_T, KR>;
  static_assert(S::NR >= 2, "middle pass needs two rounds");
  constexpr int K = KR;
  constexpr int B0 = LOG_T - KR;  // top chunk
  static_assert(S::template k<1, S::NR - 1>() == K && S::template b0<1, S::NR - 1>() == B0, "inverse last round");
  static_assert(S::template k<0, 0>() == K && S::template b0<0, 0>() == B0, "forward first round");
  constexpr int R = 1 << K;
  extern __shared__ __align__(16) unsigned char smem[];
  const uint32_t log_w = a.log_w;
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(smem);
  unsigned char *tabp = smem + (kLaneInfoBytes << log_w);
  uint32_t *tile = reinterpret_cast<uint32_t *>(tabp + ((size_t)8 << LOG_T));
  const uint64_t g0 = (uint64_t)blockIdx.x << log_w;
  for (uint32_t l = threadIdx.x; l < (1u << log_w); l += blockDim.x) decode_lane(a, LOG_T, g0 + l, lanes[l]);
  load_root<LOG_T>(a, reinterpret_cast<uint2 *>(tabp));
  __syncthreads();
  RootTw<LOG_T, 1> twi;
  RootTw<LOG_T, 0> twf;
  twi.tab = twf.tab = reinterpret_cast<const uint2 *>(tabp);
  RoundIO io{};
  io.src = a.src;
  io.dst = a.dst;
  io.sps = ((uint64_t)1 << a.lo) * a.src_rs;
  io.dps = ((uint64_t)1 << a.lo) * a.dst_rs;
  io.twist_in = a.twist;
  io.twist_out = a.twist_fwd;
  io.lo = a.lo;
  // inverse rounds 0 .. NR-2 (round 0 loads from HBM); they end with an LDS write
  rounds<F, 1, LOG_T, true, 0, S::NR - 1>(tile, log_w, lanes, twi, io);
  __syncthreads();
  // inverse last round: one item per thread (blockDim == W << (LOG_T - K))
  const uint32_t W = 1u << log_w;
  const uint32_t it = threadIdx.x;
  const uint32_t lane = it & (W - 1), rb = it >> log_w;
  const uint32_t pbase = round_pbase<K, B0>(rb);
  uint32_t c[R];
#pragma unroll
  for (int i = 0; i < R; ++i) c[i] = tile[lidx<LOG_T, true>(lane, pbase | (i << B0), log_w)];
  round_regs<F, 1, K, B0>(c, pbase, twi);
  const uint32_t low = lanes[lane].low;
  const uint64_t n = 1ull << a.log_n;
  for (uint32_t b = 0; b < a.blocks; ++b) {
    const uint2 *mid = a.mid + b * n;
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) v[i] = F::mul_tw(c[i], mid[((uint64_t)(pbase | (i << B0)) << a.lo) | low]);
    round_regs<F, 0, K, B0>(v, pbase, twf);
    __syncthreads();  // every thread is done reading the tile
#pragma unroll
    for (int i = 0; i < R; ++i) tile[lidx<LOG_T, true>(lane, pbase | (i << B0), log_w)] = v[i];
    __syncthreads();
    io.dst_off = b * a.dst_block_off;
    rounds<F, 0, LOG_T, true, 1, S::NR>(tile, log_w, lanes, twf, io);
  }
}

template <class F, int DIR, int LOG_T, bool CONTIG>
static hipError_t launch_t(const PassArgs &a, uint64_t tiles, hipStream_t s) {
  const size_t lds = lds_bytes(F::ID, LOG_T, a.log_w);
  const uint64_t threads = (uint64_t)1 << (a.log_w + LOG_T - KR);
  if (lds > kMaxLdsBytes || threads > 512 || tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  static bool configured = false;
  auto kern = k_fast<F, DIR, LOG_T, CONTIG>;
  if (!configured) {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    configured = true;
  }
  hipLaunchKernelGGL(kern, dim3((uint32_t)tiles), dim3((uint32_t)threads), lds, s, a);
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

#if SFN_IN_PART(4)
hipError_t launch_fast_m31_fwd(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_m31_inv(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_bb_fwd(int, bool, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_fast_bb_inv(int, bool, const PassArgs &, uint64_t, hipStream_t);

size_t fast_lds_bytes(int field, int log_t, uint32_t log_w) { return fast::lds_bytes(field, log_t, log_w); }

hipError_t launch_fast(int field, int dir, int log_t, bool contig, const PassArgs &a, uint64_t tiles,
                       hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  if (field == 0) return dir == 0 ? launch_fast_m31_fwd(log_t, contig, a, tiles, s)
                                  : launch_fast_m31_inv(log_t, contig, a, tiles, s);
  return dir == 0 ? launch_fast_bb_fwd(log_t, contig, a, tiles, s) : launch_fast_bb_inv(log_t, contig, a, tiles, s);
}

template <int LOG_T>
static hipError_t launch_mid_fast_t(const PassArgs &a, uint64_t tiles, hipStream_t s) {
  const size_t lds = fast::lds_bytes(1, LOG_T, a.log_w);
  const uint64_t threads = (uint64_t)1 << (a.log_w + LOG_T - fast::KR);
  if (lds > kMaxLdsBytes || threads > 512 || tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  static bool configured = false;
  if (!configured) {
    (void)hipFuncSetAttribute((const void *)fast::k_mid_fast<LOG_T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)kMaxLdsBytes);
    configured = true;
  }
  hipLaunchKernelGGL(fast::k_mid_fast<LOG_T>, dim3((uint32_t)tiles), dim3((uint32_t)threads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_lde_mid_fast(int log_t, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  switch (log_t) {
    case 6: return launch_mid_fast_t<6>(a, tiles, s);
    case 7: return launch_mid_fast_t<7>(a, tiles, s);
    case 8: return launch_mid_fast_t<8>(a, tiles, s);
    case 9: return launch_mid_fast_t<9>(a, tiles, s);
    case 10: return launch_mid_fast_t<10>(a, tiles, s);
    case 11: return launch_mid_fast_t<11>(a, tiles, s);
    default: return hipErrorInvalidValue;
  }
}
#endif  // SFN_IN_PART(4)

}  // namespace sfn
