// Device field arithmetic for gfx950 (VALU only; 32-bit lanes).
//
// M31 (p = 2^31 - 1): canonical u32 values, Mersenne folding
//   (src/field/m31.rs:101-115, shaders/fields/m31_field.metal:18-37).
// Montgomery fields, R = 2^32, raw Montgomery values:
//   BabyBear  p = 2^31 - 2^27 + 1 (src/field/babybear.rs:40-51,
//             shaders/fields/babybear_field.metal:23-43);
//   KoalaBear p = 2^31 - 2^24 + 1 (same arithmetic, two-adicity 24; not in
//             the reference, which lists it as future work in README.md:20).
//
// Twiddle multiplications use a precomputed companion (Shoup, "NTL: A Library
// for doing Number Theory", MulModPrecon): for a constant w in [0, p) and
// wq = floor(w * 2^32 / p), q = mulhi(x, wq) and r = x*w - q*p (mod 2^32)
// lies in [0, 2p) for every 32-bit x.  For a Montgomery field the raw value
// x = X*R is multiplied by the *standard* value of the twiddle, which gives
// the raw Montgomery form of X*W directly.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sfn {

// a in [0, 2p) -> [0, p): when a < p the subtraction wraps above a.
template <uint32_t P>
__device__ __forceinline__ uint32_t canon(uint32_t a) {
  return min(a, a - P);
}

struct M31 {
  static constexpr uint32_t P = 0x7FFFFFFFu;
  static constexpr int ID = 0;
  static constexpr bool MONTY = false;
  using Tw = uint32_t;  // twiddles are plain canonical values

  __device__ static __forceinline__ uint32_t add(uint32_t a, uint32_t b) { return canon<P>(a + b); }
  __device__ static __forceinline__ uint32_t sub(uint32_t a, uint32_t b) { return canon<P>(a + P - b); }
  // a, b < p: t < p^2, (t & p) + (t >> 31) < 2p.
  __device__ static __forceinline__ uint32_t mul(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t lo = (uint32_t)t & P;
    uint32_t hi = (uint32_t)(t >> 31);
    return canon<P>(lo + hi);
  }
  __device__ static __forceinline__ uint32_t mul_tw(uint32_t a, Tw w) { return mul(a, w); }
};

// PINV = p^-1 mod 2^32; both moduli are below 2^31, so sums of two canonical
// values and the Shoup remainders fit in 32 bits.
template <uint32_t P_, uint32_t PINV_, int ID_, uint32_t ROOT_, int LOG_ROOT_>
struct MontyField {
  static constexpr uint32_t P = P_;
  static constexpr uint32_t PINV = PINV_;
  static constexpr uint32_t PINV_NEG = 0u - PINV_;  // -p^-1 mod 2^32
  static constexpr int ID = ID_;
  // a primitive 2^LOG_ROOT-th root of unity (standard form): the field's
  // two-adic generator, from which every root table is derived
  static constexpr uint32_t ROOT = ROOT_;
  static constexpr int LOG_ROOT = LOG_ROOT_;
  static constexpr bool MONTY = true;
  using Tw = uint2;  // .x = standard value w, .y = floor(w * 2^32 / p)
  static_assert((uint32_t)(P_ * PINV_) == 1u, "PINV is p^-1 mod 2^32");

  __device__ static __forceinline__ uint32_t add(uint32_t a, uint32_t b) { return canon<P>(a + b); }
  __device__ static __forceinline__ uint32_t sub(uint32_t a, uint32_t b) { return canon<P>(a + P - b); }
  __device__ static __forceinline__ uint32_t mulp(uint32_t q) { return q * P; }
  // x any 32-bit value; result canonical.
  __device__ static __forceinline__ uint32_t mul_tw(uint32_t x, Tw w) {
    uint32_t q = __umulhi(x, w.y);
    uint32_t r = x * w.x - mulp(q);
    return canon<P>(r);
  }
  // Lazy variant: result in [0, 2p).
  __device__ static __forceinline__ uint32_t mul_tw_lazy(uint32_t x, Tw w) {
    uint32_t q = __umulhi(x, w.y);
    return x * w.x - mulp(q);
  }
  // x * w for a twiddle held in Montgomery form wm = w * 2^32 mod p, any 32-bit
  // x: t = x*wm < p*2^32, m = t_lo * p^-1 (mod 2^32) makes t - m*p divisible by
  // 2^32 with no borrow, so (t - m*p) / 2^32 = t_hi - mulhi(m, p) in (-p, p).
  __device__ static __forceinline__ uint32_t mul_mont(uint32_t x, uint32_t wm) {
    const uint64_t t = (uint64_t)x * wm;
    const uint32_t m = (uint32_t)t * PINV;
    const uint32_t r = (uint32_t)(t >> 32) - __umulhi(m, P);
    return min(r, r + P);
  }
  // Montgomery product without the final reduction: for a * b < p * 2^32 the
  // sum t + m*p < 2^33 * p fits 64 bits and the result lies in [0, 2p).
  __device__ static __forceinline__ uint32_t mul_lazy(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * PINV_NEG;
    return (uint32_t)((t + (uint64_t)m * P) >> 32);
  }
  // Montgomery product of two raw values (src/field/babybear.rs:133-137);
  // canonical for a * b < p * 2^32 (a, b < p, or one of them below 2p).
  __device__ static __forceinline__ uint32_t mul(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t m = (uint32_t)t * PINV_NEG;
    uint64_t s = t + (uint64_t)m * P;  // < 2^64 for a, b < p
    return canon<P>((uint32_t)(s >> 32));
  }
};

// BabyBear's 2^27-th root: src/field/babybear.rs:31-33; KoalaBear's 2^24-th
// root: 3^127 (3 generates the multiplicative group, p - 1 = 2^24 * 127)
using BabyBear = MontyField<0x78000001u, 0x88000001u, 1, 440564289u, 27>;
using KoalaBear = MontyField<0x7F000001u, 0x81000001u, 2, 1791270792u, 24>;

// Compile-time root table g^e, e < 2^LOG, g = a primitive 2^LOG-th root of F,
// as Shoup pairs (w, floor(w * 2^32 / p)): the values the host's root tables
// hold.  Indexed with constants (after unrolling), its entries fold into
// instruction operands and occupy no memory.
template <class F, int LOG>
struct ConstRoots {
  uint32_t w[1 << LOG], q[1 << LOG];
  constexpr ConstRoots() : w{}, q{} {
    uint64_t g = F::ROOT;
    for (int k = LOG; k < F::LOG_ROOT; ++k) g = g * g % F::P;
    uint64_t x = 1;
    for (int e = 0; e < (1 << LOG); ++e) {
      w[e] = (uint32_t)x;
      q[e] = (uint32_t)((x << 32) / F::P);
      x = x * g % F::P;
    }
  }
};
template <class F, int LOG>
__device__ constexpr ConstRoots<F, LOG> kConstRoots{};

// ---- butterflies ----
// CT (decimation in time):  (x, y) -> (x + w*y, x - w*y)
// GS (decimation in freq.): (x, y) -> (x + y, (x - y)*w)
template <class F>
__device__ __forceinline__ void bfly_ct(uint32_t &x, uint32_t &y, typename F::Tw w) {
  uint32_t t = F::mul_tw(y, w);
  uint32_t a = x;
  x = F::add(a, t);
  y = F::sub(a, t);
}

// CT with the negated twiddle: given w' = -w, (x, y) -> (x - w'y, x + w'y)
// = (x + wy, x - wy).
template <class F>
__device__ __forceinline__ void bfly_ct_neg(uint32_t &x, uint32_t &y, typename F::Tw wn) {
  uint32_t t = F::mul_tw(y, wn);
  uint32_t a = x;
  x = F::sub(a, t);
  y = F::add(a, t);
}

// GS with both outputs unreduced, in [0, 2p): a + b and the lazy twiddle product.
template <class F>
__device__ __forceinline__ void bfly_gs_lazy(uint32_t &x, uint32_t &y, typename F::Tw w) {
  const uint32_t a = x, b = y;
  x = a + b;
  y = F::mul_tw_lazy(a + F::P - b, w);
}

template <class F>
__device__ __forceinline__ void bfly_gs(uint32_t &x, uint32_t &y, typename F::Tw w) {
  uint32_t a = x, b = y;
  x = F::add(a, b);
  if constexpr (F::MONTY) y = F::mul_tw(a + F::P - b, w);  // (0, 2p): Shoup accepts any 32-bit input
  else y = F::mul(F::sub(a, b), w);
}

// Either butterfly with twiddle 1: (x, y) -> (x + y, x - y).
template <class F>
__device__ __forceinline__ void bfly_unit(uint32_t &x, uint32_t &y) {
  const uint32_t a = x, b = y;
  x = F::add(a, b);
  y = F::sub(a, b);
}

// Montgomery-field CT butterflies whose outputs may stay unreduced: with lx
// (ly) the x (y) output is left in [0, 2p) as a + b or a - b + p (a, b < p),
// for a value whose next use is the multiplied operand of a twiddle product
// (Shoup's product takes any 32-bit value).  lx and ly are constants once the
// caller's loops are unrolled.
template <class F>
__device__ __forceinline__ void bfly_unit_l(uint32_t &x, uint32_t &y, bool lx, bool ly) {
  const uint32_t a = x, b = y;
  x = lx ? a + b : F::add(a, b);
  y = ly ? a + F::P - b : F::sub(a, b);
}
template <class F>
__device__ __forceinline__ void bfly_ct_l(uint32_t &x, uint32_t &y, typename F::Tw w, bool lx, bool ly) {
  const uint32_t t = F::mul_tw(y, w), a = x;
  x = lx ? a + t : F::add(a, t);
  y = ly ? a + F::P - t : F::sub(a, t);
}
template <class F>
__device__ __forceinline__ void bfly_ct_neg_l(uint32_t &x, uint32_t &y, typename F::Tw wn, bool lx, bool ly) {
  const uint32_t t = F::mul_tw(y, wn), a = x;
  x = lx ? a + F::P - t : F::sub(a, t);
  y = ly ? a + t : F::add(a, t);
}

}  // namespace sfn
