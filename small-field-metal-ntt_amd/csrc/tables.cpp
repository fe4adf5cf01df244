#include "tables.hpp"

namespace sfn {
namespace host {

std::vector<Pair> position_table(const Monty &f, uint32_t log_max, bool inverse) {
  std::vector<Pair> t(((size_t)1 << log_max) > 0 ? ((size_t)1 << log_max) : 1);
  for (uint32_t s = 0; s < log_max; ++s) {
    uint32_t g = f.root(s + 1);
    if (inverse) g = invmod(g, f.p);
    uint32_t w = 1;
    Pair *layer = t.data() + ((size_t)1 << s) - 1;
    for (size_t j = 0; j < ((size_t)1 << s); ++j) {
      layer[j] = shoup(w, f.p);
      w = mulmod(w, g, f.p);
    }
  }
  return t;
}

std::vector<Pair> root_table(const Monty &f, uint32_t log) {
  std::vector<Pair> t((size_t)1 << log);
  const uint32_t g = f.root(log);
  uint32_t w = 1;
  for (size_t e = 0; e < t.size(); ++e) {
    t[e] = shoup(w, f.p);
    w = mulmod(w, g, f.p);
  }
  return t;
}

std::vector<Pair> twist_table(const Monty &f, uint32_t lo, uint32_t hi, bool inverse) {
  const uint32_t lt = hi - lo;
  const uint64_t mask = ((uint64_t)1 << hi) - 1;
  std::vector<uint32_t> pw((size_t)1 << hi);
  const uint32_t g = f.root(hi);
  uint32_t w = 1;
  for (size_t e = 0; e < pw.size(); ++e) {
    pw[e] = w;
    w = mulmod(w, g, f.p);
  }
  std::vector<Pair> t((size_t)1 << hi);
  for (uint64_t p = 0; p < ((uint64_t)1 << lt); ++p) {
    const uint64_t br = bitrev(p, lt);
    for (uint64_t low = 0; low < ((uint64_t)1 << lo); ++low) {
      uint64_t e = (low * br) & mask;
      if (inverse) e = (((uint64_t)1 << hi) - e) & mask;
      t[(p << lo) | low] = shoup(pw[e], f.p);
    }
  }
  return t;
}

std::vector<Pair> bb_lde_mid_table(uint32_t log_n, uint32_t added_bits, uint32_t shift_std) {
  const size_t n = (size_t)1 << log_n;
  const uint32_t blocks = 1u << added_bits;
  const uint32_t inv_n = invmod((uint32_t)(n % kBBP), kBBP);
  const uint32_t gN = bb_root(log_n + added_bits);
  std::vector<Pair> t(n * blocks);
  for (uint32_t b = 0; b < blocks; ++b) {
    const uint32_t zeta = powmod(gN, bitrev(b, added_bits), kBBP);
    const uint32_t step = mulmod(shift_std, zeta, kBBP);
    uint32_t w = inv_n;
    for (size_t j = 0; j < n; ++j) {
      t[b * n + j] = shoup(w, kBBP);
      w = mulmod(w, step, kBBP);
    }
  }
  return t;
}

std::vector<uint32_t> low_major(const std::vector<uint32_t> &t, uint32_t lo, uint32_t hi) {
  const uint32_t lt = hi - lo;
  std::vector<uint32_t> r(t.size());
  for (uint64_t p = 0; p < ((uint64_t)1 << lt); ++p)
    for (uint64_t low = 0; low < ((uint64_t)1 << lo); ++low) r[(low << lt) | p] = t[(p << lo) | low];
  return r;
}

std::vector<uint32_t> to_monty_words(const Monty &f, const std::vector<Pair> &t) {
  std::vector<uint32_t> m(t.size());
  for (size_t i = 0; i < t.size(); ++i) m[i] = f.to_monty(t[i].w);
  return m;
}

void bb_lde_fast_tables(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo,
                        std::vector<uint32_t> &post, std::vector<uint32_t> &pointf) {
  const size_t n = (size_t)1 << log_n;
  const uint32_t lt = log_n - lo;
  const size_t T = (size_t)1 << lt;
  const uint32_t blocks = 1u << added_bits;
  const uint32_t inv_n = invmod((uint32_t)(n % kBBP), kBBP);
  const uint32_t gN = bb_root(log_n + added_bits);
  const std::vector<Pair> tw = twist_table(kBabyBear, lo, log_n, false);
  post.assign(n * blocks, 0);
  pointf.assign(T * blocks, 0);
  for (uint32_t b = 0; b < blocks; ++b) {
    const uint32_t sigma = mulmod(shift_std, powmod(gN, bitrev(b, added_bits), kBBP), kBBP);
    std::vector<uint32_t> lane((size_t)1 << lo);
    uint32_t w = inv_n;
    for (size_t low = 0; low < lane.size(); ++low) {
      lane[low] = w;  // n^-1 * sigma^low
      w = mulmod(w, sigma, kBBP);
    }
    for (size_t p = 0; p < T; ++p)
      for (size_t low = 0; low < lane.size(); ++low) {
        const size_t i = (p << lo) | low;
        post[b * n + i] = bb_to_monty(mulmod(tw[i].w, lane[low], kBBP));
      }
    const uint32_t step = powmod(sigma, (uint64_t)1 << lo, kBBP);  // sigma^(2^lo)
    uint32_t f = 1;
    for (size_t p = 0; p < T; ++p) {
      const size_t at = lt >= 5 ? ((p & ((T >> 5) - 1)) << 5) | (p >> (lt - 5)) : p;
      pointf[b * T + at] = bb_to_monty(f);
      f = mulmod(f, step, kBBP);
    }
  }
}

namespace {
struct Pt {
  uint32_t x, y;
};
inline uint32_t m_add(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a + b) % kM31P); }
inline uint32_t m_sub(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a + kM31P - b) % kM31P); }
inline uint32_t m_mul(uint32_t a, uint32_t b) { return mulmod(a, b, kM31P); }
// circle group law (src/field/circle.rs:34-40)
inline Pt p_add(Pt a, Pt b) {
  return Pt{m_sub(m_mul(a.x, b.x), m_mul(a.y, b.y)), m_add(m_mul(a.x, b.y), m_mul(a.y, b.x))};
}
inline Pt p_double(Pt a) { return p_add(a, a); }
// generator of the subgroup of order 2^k: CIRCLE_GEN doubled 31-k times (src/field/circle.rs:78-85)
inline Pt p_gen(uint32_t k) {
  Pt g{2u, 1268011823u};
  for (uint32_t i = 0; i < 31 - k; ++i) g = p_double(g);
  return g;
}
}  // namespace

std::vector<uint32_t> m31_block_table(uint32_t log_n, bool inverse) {
  const size_t n = (size_t)1 << log_n;
  std::vector<uint32_t> t(n > 1 ? n - 1 : 1, 0);
  if (log_n == 0) return t;
  // coset odds(L): initial = G_{2n}, step = G_n (src/field/circle.rs:108-117)
  Pt init = p_gen(log_n + 1), step = p_gen(log_n);
  std::vector<Pt> pts;
  for (uint32_t l = 0; l < log_n; ++l) {
    const uint32_t ls = log_n - l;                 // log size of the current coset
    const size_t half = (size_t)1 << (ls - 1);
    pts.resize(half);
    Pt cur = init;
    for (size_t m = 0; m < half; ++m) {  // at(m) = initial + m * step
      pts[m] = cur;
      cur = p_add(cur, step);
    }
    uint32_t *layer = t.data() + (n - ((size_t)1 << ls));
    const bool last = (l == log_n - 1);
    for (size_t i = 0; i < half; ++i) {
      const Pt &p = pts[bitrev(i, ls - 1)];
      layer[i] = last ? p.y : p.x;
    }
    init = p_double(init);  // coset doubling (src/field/circle.rs:143-150)
    step = p_double(step);
  }
  if (inverse) {
    // batch inversion (Montgomery's trick); zero maps to zero like a^(p-2)
    std::vector<uint32_t> pre(t.size());
    uint32_t acc = 1;
    for (size_t i = 0; i < t.size(); ++i) {
      pre[i] = acc;
      if (t[i]) acc = m_mul(acc, t[i]);
    }
    uint32_t inv = invmod(acc, kM31P);
    for (size_t i = t.size(); i-- > 0;) {
      if (!t[i]) continue;
      const uint32_t v = t[i];
      t[i] = m_mul(inv, pre[i]);
      inv = m_mul(inv, v);
    }
  }
  return t;
}

}  // namespace host
}  // namespace sfn
