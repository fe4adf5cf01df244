// Host-side generation of twiddle, twist and coset-factor tables.
// All arithmetic is exact modular arithmetic, so any evaluation order gives
// the same values as the reference's generators.  Large tables are filled in
// chunks on up to 16 host threads; each chunk starts from a power (or circle
// point) computed directly, so the result does not depend on the split.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace sfn {
namespace host {

constexpr uint32_t kM31P = 0x7FFFFFFFu;
constexpr uint32_t kBBP = 0x78000001u;
constexpr uint32_t kBBRoot27 = 440564289u;  // primitive 2^27-th root (src/field/babybear.rs:31-33)
// KoalaBear, p = 2^31 - 2^24 + 1 = 2^24 * 127 + 1 (README.md:20, planned in the
// reference): 3 generates F_p^*, so 3^127 is a primitive 2^24-th root.
constexpr uint32_t kKBP = 0x7F000001u;
constexpr uint32_t kKBRoot24 = 1791270792u;  // 3^127 mod p

inline uint32_t mulmod(uint32_t a, uint32_t b, uint32_t p) { return (uint32_t)((uint64_t)a * b % p); }
inline uint32_t powmod(uint32_t a, uint64_t e, uint32_t p) {
  uint32_t r = 1 % p;
  while (e) {
    if (e & 1) r = mulmod(r, a, p);
    a = mulmod(a, a, p);
    e >>= 1;
  }
  return r;
}
inline uint32_t invmod(uint32_t a, uint32_t p) { return a ? powmod(a, p - 2, p) : 0; }

// BabyBear standard <-> Montgomery raw (R = 2^32)
inline uint32_t bb_to_monty(uint32_t v) { return (uint32_t)(((uint64_t)v << 32) % kBBP); }
inline uint32_t bb_from_monty(uint32_t m) {
  // m * R^-1 mod p; R^-1 = (2^32)^-1 mod p
  static const uint32_t rinv = invmod((uint32_t)((1ull << 32) % kBBP), kBBP);
  return mulmod(m, rinv, kBBP);
}
// Primitive 2^k-th root of unity, standard form (two_adic_generator, src/field/babybear.rs:69-76).
inline uint32_t bb_root(uint32_t k) { return powmod(kBBRoot27, 1ull << (27 - k), kBBP); }

// A Montgomery field (BabyBear or KoalaBear): modulus, primitive 2^max_log-th
// root (standard form) and two-adicity.  The NTT tables below take one.
struct Monty {
  uint32_t p, root_max, max_log;
  uint32_t root(uint32_t k) const { return powmod(root_max, 1ull << (max_log - k), p); }
  uint32_t to_monty(uint32_t v) const { return (uint32_t)(((uint64_t)v << 32) % p); }
};
constexpr Monty kBabyBear{kBBP, kBBRoot27, 27};
constexpr Monty kKoalaBear{kKBP, kKBRoot24, 24};
// field id of the C ABI: 1 = BabyBear, 2 = KoalaBear
inline const Monty &monty_field(int field) { return field == 2 ? kKoalaBear : kBabyBear; }

// Shoup pair: (w, floor(w * 2^32 / p))
struct Pair {
  uint32_t w, q;
};
inline Pair shoup(uint32_t w, uint32_t p) { return Pair{w, (uint32_t)(((uint64_t)w << 32) / p)}; }

// Low `bits` bits of v in reverse order (bits <= 64; 0 -> 0).
inline uint64_t bitrev(uint64_t v, uint32_t bits) {
  if (bits == 0) return 0;
  v = ((v >> 1) & 0x5555555555555555ull) | ((v & 0x5555555555555555ull) << 1);
  v = ((v >> 2) & 0x3333333333333333ull) | ((v & 0x3333333333333333ull) << 2);
  v = ((v >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((v & 0x0F0F0F0F0F0F0F0Full) << 4);
  v = ((v >> 8) & 0x00FF00FF00FF00FFull) | ((v & 0x00FF00FF00FF00FFull) << 8);
  v = ((v >> 16) & 0x0000FFFF0000FFFFull) | ((v & 0x0000FFFF0000FFFFull) << 16);
  v = (v >> 32) | (v << 32);
  return v >> (64 - bits);
}

// Position table (BabyBear: src/field/babybear.rs:83-97; inverse:
// src/ntt/babybear/twiddles.rs:13-19): layer s at offset 2^s - 1 holds
// g_{s+1}^(+-j), j < 2^s, as Shoup pairs.  Independent of the transform size.
std::vector<Pair> position_table(const Monty &f, uint32_t log_max, bool inverse);

// Root table g_{2^log}^e, e < 2^log (Shoup pairs): one table for every layer
// and both directions of a sub-transform of up to 2^log points.
std::vector<Pair> root_table(const Monty &f, uint32_t log);

// Montgomery-form powers of f's primitive 2^max_log-th root g in two levels:
// [g^j, j < 2^lo_bits][g^(j << lo_bits), j < 2^(max_log - lo_bits)], the second
// level padded with zeros to `words` in total.
std::vector<uint32_t> root_powers_m(const Monty &f, uint32_t lo_bits, size_t words);

// Four-step twist of layers [lo, hi): entry (p << lo) | low holds
// g_hi^(+-low * bitrev(p, hi - lo)).
std::vector<Pair> twist_table(const Monty &f, uint32_t lo, uint32_t hi, bool inverse);

// Reorders a table of 2^hi entries indexed (p << lo) | low into (low << (hi - lo)) | p.
std::vector<uint32_t> low_major(const std::vector<uint32_t> &t, uint32_t lo, uint32_t hi);

// Single-word Montgomery form (w * 2^32 mod p) of a table of standard values.
std::vector<uint32_t> to_monty_words(const Monty &f, const std::vector<Pair> &t);

// Coset-LDE tuned middle pass, sigma_b = shift * g_{L+a}^bitrev(b, a):
//  post:   block b entry (p << lo) | low = g_L^(low*bitrev(p, L-lo)) * n^-1 * sigma_b^low
//  pointf: block b holds sigma_b^(p << lo) for p < 2^(L-lo), item-major for
//          radix-32 items (L - lo >= 5): point p = r | (i << (L-lo-5)),
//          r < 2^(L-lo-5), i < 32, sits at entry r * 32 + i
// both in Montgomery form.
void bb_lde_fast_tables(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo,
                        std::vector<uint32_t> &post, std::vector<uint32_t> &pointf);

// The per-lane part of the post factors: block b entry low holds
// n^-1 * sigma_b^low (Montgomery form), low < 2^lo, sigma_b = shift * g_{L+a}^bitrev(b, a).
std::vector<uint32_t> bb_lde_lane_factors(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo);

// Coset-LDE middle factors: block b (of 2^a) entry j holds
// n^-1 * (shift * g_{L+a}^bitrev(b, a))^j, j < 2^L.
std::vector<Pair> bb_lde_mid_table(uint32_t log_n, uint32_t added_bits, uint32_t shift_std);

// M31 circle twiddles (src/ntt/m31/twiddles.rs:19-68): layer l at offset
// n - 2^(L-l); inverse = element-wise inverse.
std::vector<uint32_t> m31_block_table(uint32_t log_n, bool inverse);

}  // namespace host
}  // namespace sfn
