// Hand-written gfx950 kernels for the M31 circle NTT, the BabyBear and
// KoalaBear NTTs and the BabyBear coset LDE.  KoalaBear (p = 2^31 - 2^24 + 1)
// shares every BabyBear kernel through the MontyField template: same
// Montgomery form, same layer order and twiddle structure, its own constants.
//
// Pass kernel: one workgroup owns 2^log_w independent sub-transforms ("lanes")
// of T = 2^LOG_T points each, i.e. layers [lo, lo + LOG_T) of every column.
// The tile is staged through LDS with coalesced u32 loads; inside, each thread
// takes 2^K points of one lane into registers and runs K layers there
// (radix-2^K register block) between LDS exchanges.  Layer order and
// butterflies follow the reference CPU backends:
//   M31 forward  = CT, layers high -> low   (src/ntt/m31/cpu_reference.rs:50-79)
//   M31 inverse  = GS, layers low -> high   (src/ntt/m31/cpu_reference.rs:82-120)
//   BB forward   = GS, layers high -> low   (src/ntt/babybear/cpu_reference.rs:34-62)
//   BB inverse   = CT, layers low -> high   (src/ntt/babybear/cpu_reference.rs:69-103)
//
// BabyBear twiddles of layer s at row r are g_{s+1}^(r mod 2^s).  For a layer
// range [lo, hi) and r = (p << lo) | low this factors as
// g_{s-lo+1}^(p mod 2^(s-lo)) * g_{s+1}^low, and the second factor, constant on
// each half that a butterfly produces, collects into one multiply per point:
// forward = standard 2^(hi-lo) GS sub-transform, then x[p] *= g_hi^(low*bitrev(p));
// inverse = x[p] *= g_hi^(-low*bitrev(p)), then the standard CT sub-transform.
// The result is the same linear map as the layer-by-layer network, so the
// outputs are bit-identical, and the in-tile twiddles come from one small
// table shared by every lane.
//
// M31 twiddles are block-indexed (tw[l][r >> (l+1)]) and have no such
// structure; they are read from the transform's table.
#include "pass_common.hpp"

namespace sfn {

// The pass instantiations are split over translation units (SFN_PART 0-3
// and 6-7, one per field x direction; 4 = everything else, 5 = the
// many-job launch) so the build compiles them in parallel.  Without SFN_PART one unit holds everything.
#ifndef SFN_PART
#define SFN_PART -1
#endif
#define SFN_IN_PART(k) (SFN_PART < 0 || SFN_PART == (k))

constexpr int NT = kPassThreads;

// LDS -> registers -> LDS for every (lane, block) item of round r.
template <class F, int DIR, int LOG_T, int LOGR, int r>
__device__ __forceinline__ void lds_round(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes,
                                          TwCtx<F> tw) {
  using S = Sched<LOG_T, LOGR>;
  constexpr int K = S::template k<DIR, r>();
  constexpr int B0 = S::template b0<DIR, r>();
  constexpr int R = 1 << K;
  constexpr int LOG_NB = LOG_T - K;
  const uint32_t items = 1u << (log_w + LOG_NB);
  for (uint32_t it = threadIdx.x; it < items; it += blockDim.x) {
    const uint32_t lane = it >> LOG_NB;
    const uint32_t rb = it & ((1u << LOG_NB) - 1);
    const uint32_t pbase = round_pbase<K, B0>(rb);
    uint32_t *lt = tile + lane * Tile<LOG_T>::TP;
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int p = pbase | (i << B0);
      v[i] = lt[Tile<LOG_T>::idx(p)];
    }
    if constexpr (!F::MONTY) tw.hb = lanes[lane].hb;
    round_regs<F, DIR, K, B0>(v, pbase, tw);
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const int p = pbase | (i << B0);
      lt[Tile<LOG_T>::idx(p)] = v[i];
    }
  }
}

template <class F, int DIR, int LOG_T, int LOGR, int r, int rend>
__device__ __forceinline__ void rounds_range(uint32_t *tile, uint32_t log_w, const LaneInfo *lanes,
                                             const TwCtx<F> &tw) {
  if constexpr (r < rend) {
    lds_round<F, DIR, LOG_T, LOGR, r>(tile, log_w, lanes, tw);
    __syncthreads();
    rounds_range<F, DIR, LOG_T, LOGR, r + 1, rend>(tile, log_w, lanes, tw);
  }
}

template <class F>
__device__ __forceinline__ uint32_t scale_mul(uint32_t v, uint32_t w, uint32_t q);
template <>
__device__ __forceinline__ uint32_t scale_mul<M31>(uint32_t v, uint32_t w, uint32_t) { return M31::mul(v, w); }
template <>
__device__ __forceinline__ uint32_t scale_mul<BabyBear>(uint32_t v, uint32_t w, uint32_t q) {
  return BabyBear::mul_tw(v, make_uint2(w, q));
}
template <>
__device__ __forceinline__ uint32_t scale_mul<KoalaBear>(uint32_t v, uint32_t w, uint32_t q) {
  return KoalaBear::mul_tw(v, make_uint2(w, q));
}

// Dynamic LDS layout: [LaneInfo x W][BB small table, T uint2][tile W x TP u32]
static_assert(sizeof(LaneInfo) == kLaneInfoBytes, "LaneInfo size");
__host__ __device__ constexpr size_t lane_bytes(uint32_t log_w) { return kLaneInfoBytes << log_w; }

// One tile (2^log_w lanes) of a pass; shared by k_pass and the multi-job kernel.
template <class F, int DIR, int LOG_T, int LOGR>
__device__ __forceinline__ void pass_tile(const PassArgs &a, uint64_t tile_idx, unsigned char *smem) {
  constexpr int T = 1 << LOG_T;
  constexpr int TP = Tile<LOG_T>::TP;
  const uint32_t log_w = a.log_w;
  const uint32_t W = 1u << log_w;
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(smem);
  uint2 *stw = reinterpret_cast<uint2 *>(smem + lane_bytes(log_w));
  uint32_t *tile = reinterpret_cast<uint32_t *>(smem + lane_bytes(log_w) + (F::MONTY ? T * sizeof(uint2) : 0));

  const uint64_t g0 = tile_idx << log_w;
  for (uint32_t l = threadIdx.x; l < W; l += NT) decode_lane(a, LOG_T, g0 + l, lanes[l]);
  if constexpr (F::MONTY) {
    const uint2 *gtw = reinterpret_cast<const uint2 *>(a.tw);
    for (uint32_t i = threadIdx.x; i < T - 1; i += NT) stw[i] = gtw[i];
  }
  __syncthreads();

  // ---- load (coalesced), inverse four-step twist on the way in ----
  const uint64_t sps = ((uint64_t)1 << a.lo) * a.src_rs;
  const bool twist_in = F::MONTY && DIR == 1 && a.twist != nullptr;
  for (uint32_t e = threadIdx.x; e < (W << LOG_T); e += NT) {
    uint32_t lane, p;
    if (a.lane_contig) { lane = e & (W - 1); p = e >> log_w; }
    else { lane = e >> LOG_T; p = e & (T - 1); }
    const LaneInfo &li = lanes[lane];
    uint32_t v = 0;
    if (li.valid) {
      v = a.src[li.sbase + p * sps];
      if constexpr (F::MONTY && DIR == 1) {
        if (twist_in) v = F::mul_tw(v, a.twist[((uint64_t)p << a.lo) | li.low]);
      }
    }
    tile[lane * TP + Tile<LOG_T>::idx(p)] = v;
  }
  __syncthreads();

  TwCtx<F> tw;
  if constexpr (F::MONTY) {
    tw.stw = stw;
  } else {
    tw.tbl = reinterpret_cast<const uint32_t *>(a.tw);
    tw.n_log = a.log_n;
    tw.lo = a.lo;
    tw.hi = a.lo + LOG_T;
    tw.hb = 0;
  }
  rounds_range<F, DIR, LOG_T, LOGR, 0, Sched<LOG_T, LOGR>::NR>(tile, log_w, lanes, tw);

  // ---- store (coalesced), forward twist / 1/n scale on the way out ----
  const uint64_t dps = ((uint64_t)1 << a.lo) * a.dst_rs;
  const bool twist_out = F::MONTY && DIR == 0 && a.twist != nullptr;
  for (uint32_t e = threadIdx.x; e < (W << LOG_T); e += NT) {
    uint32_t lane, p;
    if (a.lane_contig) { lane = e & (W - 1); p = e >> log_w; }
    else { lane = e >> LOG_T; p = e & (T - 1); }
    const LaneInfo &li = lanes[lane];
    if (!li.valid) continue;
    uint32_t v = tile[lane * TP + Tile<LOG_T>::idx(p)];
    if constexpr (F::MONTY && DIR == 0) {
      if (twist_out) v = F::mul_tw(v, a.twist[((uint64_t)p << a.lo) | li.low]);
    }
    if (a.has_scale) v = scale_mul<F>(v, a.scale_w, a.scale_q);
    a.dst[li.dbase + p * dps] = v;
  }
}

template <class F, int DIR, int LOG_T, int LOGR>
__global__ __launch_bounds__(NT) void k_pass(const PassArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  pass_tile<F, DIR, LOG_T, LOGR>(a, blockIdx.x, smem);
}

#if SFN_IN_PART(4)
// ---- coset LDE middle pass (BabyBear) ----
// Layers [lo, log_n) of the inverse, x n^-1 * (shift*zeta_b)^row, then layers
// [log_n-1, lo] of the forward for every output block b, without leaving the
// workgroup.  The inverse's last register round and the forward's first round
// own the same points, so the coefficients stay in registers across blocks.
template <int LOG_T, int LOGR>
__global__ __launch_bounds__(1024) void k_lde_mid(const PassArgs a) {
  using F = BabyBear;
  using S = Sched<LOG_T, LOGR>;
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int T = 1 << LOG_T;
  constexpr int TP = Tile<LOG_T>::TP;
  // The inverse's last round and the forward's first round cover the same
  // top chunk of point bits (see Sched).
  constexpr int K = S::template k<1, S::NR - 1>();
  constexpr int B0 = S::template b0<1, S::NR - 1>();
  static_assert(K == S::template k<0, 0>() && B0 == S::template b0<0, 0>(), "round mismatch");
  constexpr int R = 1 << K;
  constexpr int LOG_NB = LOG_T - K;
  const uint32_t log_w = a.log_w;
  const uint32_t W = 1u << log_w;
  const uint32_t nthreads = blockDim.x;  // == W << LOG_NB (checked by the launcher)
  LaneInfo *lanes = reinterpret_cast<LaneInfo *>(smem);
  uint2 *stw_inv = reinterpret_cast<uint2 *>(smem + lane_bytes(log_w));
  uint2 *stw_fwd = stw_inv + T;
  uint32_t *tile = reinterpret_cast<uint32_t *>(stw_fwd + T);

  const uint64_t g0 = (uint64_t)blockIdx.x << log_w;
  for (uint32_t l = threadIdx.x; l < W; l += nthreads) decode_lane(a, LOG_T, g0 + l, lanes[l]);
  {
    // a.tw = [inverse position table | forward position table], 2^kMaxLogTile each
    const uint2 *gt = reinterpret_cast<const uint2 *>(a.tw);
    for (uint32_t i = threadIdx.x; i < T - 1; i += nthreads) {
      stw_inv[i] = gt[i];
      stw_fwd[i] = gt[(1u << kMaxLogTile) + i];
    }
  }
  __syncthreads();

  const uint64_t sps = ((uint64_t)1 << a.lo) * a.src_rs;
  for (uint32_t e = threadIdx.x; e < (W << LOG_T); e += nthreads) {
    uint32_t lane, p;
    if (a.lane_contig) { lane = e & (W - 1); p = e >> log_w; }
    else { lane = e >> LOG_T; p = e & (T - 1); }
    const LaneInfo &li = lanes[lane];
    uint32_t v = 0;
    if (li.valid) {
      v = a.src[li.sbase + p * sps];
      v = F::mul_tw(v, a.twist[((uint64_t)p << a.lo) | li.low]);
    }
    tile[lane * TP + Tile<LOG_T>::idx(p)] = v;
  }
  __syncthreads();

  TwCtx<F> twi, twf;
  twi.stw = stw_inv;
  twf.stw = stw_fwd;
  rounds_range<F, 1, LOG_T, LOGR, 0, S::NR - 1>(tile, log_w, lanes, twi);

  // last inverse round: exactly one item per thread; coefficients stay in registers
  const uint32_t it = threadIdx.x;
  const uint32_t lane = it >> LOG_NB;
  const uint32_t rb = it & ((1u << LOG_NB) - 1);
  const uint32_t pbase = round_pbase<K, B0>(rb);
  uint32_t *lt = tile + lane * TP;
  uint32_t c[R];
#pragma unroll
  for (int i = 0; i < R; ++i) c[i] = lt[Tile<LOG_T>::idx(pbase | (i << B0))];
  round_regs<F, 1, K, B0>(c, pbase, twi);
  __syncthreads();

  const uint32_t low = lanes[lane].low;
  const uint64_t n = 1ull << a.log_n;
  const uint64_t dps = ((uint64_t)1 << a.lo) * a.dst_rs;
  for (uint32_t b = 0; b < a.blocks; ++b) {
    const uint2 *mid = a.mid + b * n;
    uint32_t v[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const uint32_t p = pbase | (i << B0);
      v[i] = F::mul_tw(c[i], mid[((uint64_t)p << a.lo) | low]);
    }
    round_regs<F, 0, K, B0>(v, pbase, twf);
#pragma unroll
    for (int i = 0; i < R; ++i) lt[Tile<LOG_T>::idx(pbase | (i << B0))] = v[i];
    __syncthreads();
    rounds_range<F, 0, LOG_T, LOGR, 1, S::NR>(tile, log_w, lanes, twf);
    for (uint32_t e = threadIdx.x; e < (W << LOG_T); e += nthreads) {
      uint32_t ln, p;
      if (a.lane_contig) { ln = e & (W - 1); p = e >> log_w; }
      else { ln = e >> LOG_T; p = e & (T - 1); }
      const LaneInfo &li = lanes[ln];
      if (!li.valid) continue;
      uint32_t x = tile[ln * TP + Tile<LOG_T>::idx(p)];
      x = F::mul_tw(x, a.twist_fwd[((uint64_t)p << a.lo) | li.low]);
      a.dst[li.dbase + b * a.dst_block_off + p * dps] = x;
    }
    __syncthreads();
  }
}

// ---- V1: one butterfly per thread, one launch per layer ----
template <class F, int DIR>
__global__ __launch_bounds__(256) void k_layer(const LayerArgs a) {
  const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= a.butterflies) return;
  const uint32_t l = a.layer;
  const uint32_t half_log = a.log_n - 1;
  const uint32_t c = (uint32_t)(g >> half_log);
  const uint64_t k = g & ((1ull << half_log) - 1);
  const uint64_t r0 = ((k >> l) << (l + 1)) | (k & ((1ull << l) - 1));
  const uint64_t r1 = r0 + (1ull << l);
  uint32_t x = a.src[c * a.src_cs + r0 * a.src_rs];
  uint32_t y = a.src[c * a.src_cs + r1 * a.src_rs];
  typename F::Tw w;
  if constexpr (F::MONTY) {
    w = reinterpret_cast<const uint2 *>(a.tw)[(1ull << l) - 1 + (r0 & ((1ull << l) - 1))];
  } else {
    const uint64_t off = (1ull << a.log_n) - (1ull << (a.log_n - l));
    w = reinterpret_cast<const uint32_t *>(a.tw)[off + (r0 >> (l + 1))];
  }
  constexpr bool GS = F::MONTY == (DIR == 0);
  if constexpr (GS) bfly_gs<F>(x, y, w);
  else bfly_ct<F>(x, y, w);
  if (a.has_scale) {
    x = scale_mul<F>(x, a.scale_w, a.scale_q);
    y = scale_mul<F>(y, a.scale_w, a.scale_q);
  }
  a.dst[c * a.dst_cs + r0 * a.dst_rs] = x;
  a.dst[c * a.dst_cs + r1 * a.dst_rs] = y;
}

template <class F>
__global__ __launch_bounds__(256) void k_pointwise(const uint32_t *a, const uint32_t *b, uint32_t *out,
                                                   uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = F::mul(a[i], b[i]);
}

// Segment k copies words [begin_k, begin_{k+1}) of one contiguous range (the
// last segment ends at `total`) between its own source and destination.
__global__ __launch_bounds__(256) void k_copy_segments(const CopySeg *segs, uint32_t nseg, uint64_t total) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t lo = 0, hi = nseg;  // segs[lo].begin <= i < segs[hi].begin
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (segs[mid].begin <= i) lo = mid;
      else hi = mid;
    }
    const uint64_t k = i - segs[lo].begin;
    segs[lo].dst[k] = segs[lo].src[k];
  }
}

__global__ __launch_bounds__(256) void k_copy(const uint32_t *src, uint32_t *dst, uint64_t len) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// Row order of a 2^L-row array of `batch` columns between natural and
// bit-reversed: dst row r = src row bitrev_L(r).  Consecutive threads walk
// rows (SoA, rs == 1) or columns (row-major) so one side of the copy is
// coalesced.  In place (src == dst) every pair r < bitrev(r) is swapped once.
__device__ __forceinline__ void rows_index(uint64_t i, uint32_t L, uint32_t batch, uint64_t rs, uint64_t &r,
                                           uint64_t &c) {
  if (rs == 1) {
    c = i >> L;
    r = i & ((1ull << L) - 1);
  } else {
    r = i / batch;
    c = i - r * batch;
  }
}

__global__ __launch_bounds__(256) void k_rows_bitrev(const uint32_t *src, uint32_t *dst, uint32_t L, uint32_t batch,
                                                     uint64_t scs, uint64_t srs, uint64_t dcs, uint64_t drs,
                                                     uint64_t total) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r, c;
    rows_index(i, L, batch, drs, r, c);
    const uint64_t rr = L ? (uint64_t)(__brev((uint32_t)r) >> (32 - L)) : 0;
    dst[c * dcs + r * drs] = src[c * scs + rr * srs];
  }
}

__global__ __launch_bounds__(256) void k_rows_bitrev_inplace(uint32_t *data, uint32_t L, uint32_t batch, uint64_t cs,
                                                             uint64_t rs, uint64_t total) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t r, c;
    rows_index(i, L, batch, rs, r, c);
    const uint64_t rr = L ? (uint64_t)(__brev((uint32_t)r) >> (32 - L)) : 0;
    if (r < rr) {
      uint32_t *a = data + c * cs + r * rs, *b = data + c * cs + rr * rs;
      const uint32_t t = *a;
      *a = *b;
      *b = t;
    }
  }
}

#endif  // SFN_IN_PART(4)

// ------------------------------------------------------------ launchers
template <class F, int DIR, int LOG_T, int LOGR>
static hipError_t launch_pass_t(const PassArgs &a, uint64_t tiles, hipStream_t s) {
  const size_t lds = pass_lds_bytes(F::ID, LOG_T, a.log_w);
  if (lds > kMaxLdsBytes) return hipErrorInvalidConfiguration;
  auto kern = k_pass<F, DIR, LOG_T, LOGR>;
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  hipLaunchKernelGGL(kern, dim3((uint32_t)tiles), dim3(NT), lds, s, a);
  return hipGetLastError();
}

template <class F, int DIR, int LOGR>
static hipError_t launch_pass_lt(int log_t, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  switch (log_t) {
#define SFN_CASE(L) \
  case L: return launch_pass_t<F, DIR, L, (LOGR < L ? LOGR : L)>(a, tiles, s);
    SFN_CASE(1) SFN_CASE(2) SFN_CASE(3) SFN_CASE(4) SFN_CASE(5) SFN_CASE(6)
    SFN_CASE(7) SFN_CASE(8) SFN_CASE(9) SFN_CASE(10) SFN_CASE(11)
#undef SFN_CASE
    default: return hipErrorInvalidValue;
  }
}

template <class F, int DIR>
static hipError_t launch_pass_r(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  switch (logr) {
    case 1: return launch_pass_lt<F, DIR, 1>(log_t, a, tiles, s);
    case 2: return launch_pass_lt<F, DIR, 2>(log_t, a, tiles, s);
    case 3: return launch_pass_lt<F, DIR, 3>(log_t, a, tiles, s);
    case 4: return launch_pass_lt<F, DIR, 4>(log_t, a, tiles, s);
    case 5: return launch_pass_lt<F, DIR, 5>(log_t, a, tiles, s);
    default: return hipErrorInvalidValue;
  }
}


#if SFN_IN_PART(0)
hipError_t launch_pass_m31_fwd(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<M31, 0>(log_t, logr, a, tiles, s);
}
#endif
#if SFN_IN_PART(1)
hipError_t launch_pass_m31_inv(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<M31, 1>(log_t, logr, a, tiles, s);
}
#endif
#if SFN_IN_PART(2)
hipError_t launch_pass_bb_fwd(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<BabyBear, 0>(log_t, logr, a, tiles, s);
}
#endif
#if SFN_IN_PART(3)
hipError_t launch_pass_bb_inv(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<BabyBear, 1>(log_t, logr, a, tiles, s);
}
#endif
#if SFN_IN_PART(6)
hipError_t launch_pass_kb_fwd(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<KoalaBear, 0>(log_t, logr, a, tiles, s);
}
#endif
#if SFN_IN_PART(7)
hipError_t launch_pass_kb_inv(int log_t, int logr, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  return launch_pass_r<KoalaBear, 1>(log_t, logr, a, tiles, s);
}
#endif

#if SFN_IN_PART(4)
hipError_t launch_pass_m31_fwd(int, int, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_pass_m31_inv(int, int, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_pass_bb_fwd(int, int, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_pass_bb_inv(int, int, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_pass_kb_fwd(int, int, const PassArgs &, uint64_t, hipStream_t);
hipError_t launch_pass_kb_inv(int, int, const PassArgs &, uint64_t, hipStream_t);

hipError_t launch_pass(int field, int dir, int log_t, int logr, const PassArgs &a, uint64_t tiles,
                       hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  if (tiles > 0x7FFFFFFFull) return hipErrorInvalidConfiguration;
  if (field == 0) return dir == 0 ? launch_pass_m31_fwd(log_t, logr, a, tiles, s)
                                  : launch_pass_m31_inv(log_t, logr, a, tiles, s);
  if (field == 2) return dir == 0 ? launch_pass_kb_fwd(log_t, logr, a, tiles, s)
                                  : launch_pass_kb_inv(log_t, logr, a, tiles, s);
  return dir == 0 ? launch_pass_bb_fwd(log_t, logr, a, tiles, s) : launch_pass_bb_inv(log_t, logr, a, tiles, s);
}

template <int LOG_T, int LOGR>
static hipError_t launch_mid_t(const PassArgs &a, uint64_t tiles, hipStream_t s) {
  using S = Sched<LOG_T, LOGR>;
  constexpr int K = S::template k<1, S::NR - 1>();
  const size_t lds = lde_mid_lds_bytes(LOG_T, a.log_w);
  if (lds > kMaxLdsBytes) return hipErrorInvalidConfiguration;
  auto kern = k_lde_mid<LOG_T, LOGR>;
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  const uint64_t threads = (uint64_t)(1u << a.log_w) << (LOG_T - K);
  if (threads > 1024 || threads < 1) return hipErrorInvalidConfiguration;
  hipLaunchKernelGGL(kern, dim3((uint32_t)tiles), dim3((uint32_t)threads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_lde_mid(int log_t, const PassArgs &a, uint64_t tiles, hipStream_t s) {
  if (tiles == 0) return hipSuccess;
  switch (log_t) {
#define SFN_CASE(L) \
  case L: return launch_mid_t<L, (L < 5 ? L : 5)>(a, tiles, s);
    SFN_CASE(1) SFN_CASE(2) SFN_CASE(3) SFN_CASE(4) SFN_CASE(5) SFN_CASE(6)
    SFN_CASE(7) SFN_CASE(8) SFN_CASE(9) SFN_CASE(10) SFN_CASE(11)
#undef SFN_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_layer(int field, int dir, const LayerArgs &a, hipStream_t s) {
  const uint64_t blocks = (a.butterflies + 255) / 256;
  if (blocks == 0) return hipSuccess;
  if (field == 0) {
    if (dir == 0) hipLaunchKernelGGL((k_layer<M31, 0>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_layer<M31, 1>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
  } else if (field == 2) {
    if (dir == 0) hipLaunchKernelGGL((k_layer<KoalaBear, 0>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_layer<KoalaBear, 1>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
  } else {
    if (dir == 0) hipLaunchKernelGGL((k_layer<BabyBear, 0>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((k_layer<BabyBear, 1>), dim3((uint32_t)blocks), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_pointwise(int field, const uint32_t *a, const uint32_t *b, uint32_t *out, uint64_t len,
                            hipStream_t s) {
  if (len == 0) return hipSuccess;
  uint64_t blocks = (len + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (field == 0) hipLaunchKernelGGL((k_pointwise<M31>), dim3((uint32_t)blocks), dim3(256), 0, s, a, b, out, len);
  else if (field == 2) hipLaunchKernelGGL((k_pointwise<KoalaBear>), dim3((uint32_t)blocks), dim3(256), 0, s, a, b, out, len);
  else hipLaunchKernelGGL((k_pointwise<BabyBear>), dim3((uint32_t)blocks), dim3(256), 0, s, a, b, out, len);
  return hipGetLastError();
}

hipError_t launch_rows_bitrev(const uint32_t *src, uint32_t *dst, uint32_t L, uint32_t batch, uint64_t scs,
                              uint64_t srs, uint64_t dcs, uint64_t drs, hipStream_t s) {
  const uint64_t total = (uint64_t)batch << L;
  if (total == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (src == dst)
    hipLaunchKernelGGL(k_rows_bitrev_inplace, dim3((uint32_t)blocks), dim3(256), 0, s, dst, L, batch, dcs, drs, total);
  else
    hipLaunchKernelGGL(k_rows_bitrev, dim3((uint32_t)blocks), dim3(256), 0, s, src, dst, L, batch, scs, srs, dcs, drs,
                       total);
  return hipGetLastError();
}

hipError_t launch_copy_segments(const CopySeg *d_segs, uint32_t nseg, uint64_t total, hipStream_t s) {
  if (total == 0 || nseg == 0) return hipSuccess;
  uint64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy_segments, dim3((uint32_t)blocks), dim3(256), 0, s, d_segs, nseg, total);
  return hipGetLastError();
}

hipError_t launch_copy(const uint32_t *src, uint32_t *dst, uint64_t len, hipStream_t s) {
  if (len == 0) return hipSuccess;
  uint64_t blocks = (len + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy, dim3((uint32_t)blocks), dim3(256), 0, s, src, dst, len);
  return hipGetLastError();
}

#endif  // SFN_IN_PART(4)

#if SFN_IN_PART(5)
// ---- many jobs in one launch ----
// Segments (one per job pass) own consecutive ranges of a global tile index;
// a persistent grid walks the tiles with a grid stride, so every wave reaches
// the end of the list and the launch drains.  Field, direction and tile size
// are selected per segment (uniform across the workgroup).
template <class F, int DIR>
__device__ __forceinline__ void multi_dispatch(uint32_t log_t, const PassArgs &a, uint64_t t, unsigned char *smem) {
  switch (log_t) {
#define SFN_CASE(L) \
  case L: pass_tile<F, DIR, L, (L < 5 ? L : 5)>(a, t, smem); break;
    SFN_CASE(1) SFN_CASE(2) SFN_CASE(3) SFN_CASE(4) SFN_CASE(5) SFN_CASE(6)
    SFN_CASE(7) SFN_CASE(8) SFN_CASE(9) SFN_CASE(10) SFN_CASE(11)
#undef SFN_CASE
    default: break;
  }
}

__global__ __launch_bounds__(NT) void k_multi(const SegDesc *segs, uint32_t nseg, uint64_t total_tiles) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint32_t s = 0;
  for (uint64_t t = blockIdx.x; t < total_tiles; t += gridDim.x) {
    while (s + 1 < nseg && segs[s + 1].tile_begin <= t) ++s;
    const PassArgs a = segs[s].a;
    const uint64_t local = t - segs[s].tile_begin;
    const uint32_t kind = segs[s].kind, log_t = segs[s].log_t;
    switch (kind) {
      case 0: multi_dispatch<M31, 0>(log_t, a, local, smem); break;
      case 1: multi_dispatch<M31, 1>(log_t, a, local, smem); break;
      case 2: multi_dispatch<BabyBear, 0>(log_t, a, local, smem); break;
      case 3: multi_dispatch<BabyBear, 1>(log_t, a, local, smem); break;
      case 4: multi_dispatch<KoalaBear, 0>(log_t, a, local, smem); break;
      case 5: multi_dispatch<KoalaBear, 1>(log_t, a, local, smem); break;
      default: break;
    }
    __syncthreads();  // the next tile reuses the LDS
  }
}

hipError_t launch_multi(const SegDesc *d_segs, uint32_t nseg, uint64_t total_tiles, size_t lds, uint32_t grid,
                        hipStream_t s) {
  if (total_tiles == 0) return hipSuccess;
  if (lds > kMaxLdsBytes || grid == 0) return hipErrorInvalidConfiguration;
  // one-time attribute set-up; a function-local static initialises once,
  // also when several host threads launch at the same time
  static const bool configured = [&] {
    (void)hipFuncSetAttribute((const void *)k_multi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLdsBytes);
    return true;
  }();
  (void)configured;
  hipLaunchKernelGGL(k_multi, dim3(grid), dim3(NT), lds, s, d_segs, nseg, total_tiles);
  return hipGetLastError();
}
#endif  // SFN_IN_PART(5)

}  // namespace sfn
