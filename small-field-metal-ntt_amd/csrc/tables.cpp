#include "tables.hpp"

#include <algorithm>
#include <thread>

namespace sfn {
namespace host {
namespace {

// Runs fn(begin, end) over [0, n) in contiguous chunks on up to 16 host
// threads (inline below 2^16 items).  A chunk whose thread cannot be started
// runs on the calling thread.  fn must not throw.
template <class Fn>
void parallel_chunks(size_t n, const Fn &fn) {
  constexpr size_t kMinChunk = (size_t)1 << 16;
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t nt = std::min<size_t>(std::min<unsigned>(hw ? hw : 1u, 16u), (n + kMinChunk - 1) / kMinChunk);
  if (nt <= 1) {
    if (n) fn((size_t)0, n);
    return;
  }
  const size_t chunk = (n + nt - 1) / nt;
  std::vector<std::thread> th;
  th.reserve(nt);
  for (size_t b = chunk; b < n; b += chunk) {
    const size_t e = std::min(n, b + chunk);
    try {
      th.emplace_back([&fn, b, e] { fn(b, e); });
    } catch (...) {
      fn(b, e);
    }
  }
  fn((size_t)0, std::min(n, chunk));
  for (auto &x : th) x.join();
}

// out[e] = g^e (mod p) for e in [b, e_end)
inline void fill_powers(uint32_t *out, size_t b, size_t e_end, uint32_t g, uint32_t p) {
  uint32_t w = powmod(g, b, p);
  for (size_t e = b; e < e_end; ++e) {
    out[e] = w;
    w = mulmod(w, g, p);
  }
}

}  // namespace

std::vector<Pair> position_table(const Monty &f, uint32_t log_max, bool inverse) {
  std::vector<Pair> t((size_t)1 << log_max);
  for (uint32_t s = 0; s < log_max; ++s) {
    uint32_t g = f.root(s + 1);
    if (inverse) g = invmod(g, f.p);
    Pair *layer = t.data() + ((size_t)1 << s) - 1;
    parallel_chunks((size_t)1 << s, [&](size_t b, size_t e) {
      uint32_t w = powmod(g, b, f.p);
      for (size_t j = b; j < e; ++j) {
        layer[j] = shoup(w, f.p);
        w = mulmod(w, g, f.p);
      }
    });
  }
  return t;
}

std::vector<Pair> root_table(const Monty &f, uint32_t log) {
  std::vector<Pair> t((size_t)1 << log);
  const uint32_t g = f.root(log);
  parallel_chunks(t.size(), [&](size_t b, size_t e) {
    uint32_t w = powmod(g, b, f.p);
    for (size_t j = b; j < e; ++j) {
      t[j] = shoup(w, f.p);
      w = mulmod(w, g, f.p);
    }
  });
  return t;
}

std::vector<uint32_t> root_powers_m(const Monty &f, uint32_t lo_bits, size_t words) {
  const size_t n_lo = (size_t)1 << lo_bits, n_hi = (size_t)1 << (f.max_log - lo_bits);
  std::vector<uint32_t> t(words, 0);
  if (n_lo + n_hi > words) return {};
  std::vector<uint32_t> p(n_lo);
  fill_powers(p.data(), 0, n_lo, f.root_max, f.p);
  for (size_t j = 0; j < n_lo; ++j) t[j] = f.to_monty(p[j]);
  const uint32_t g_hi = powmod(f.root_max, n_lo, f.p);
  p.assign(n_hi, 0);
  fill_powers(p.data(), 0, n_hi, g_hi, f.p);
  for (size_t j = 0; j < n_hi; ++j) t[n_lo + j] = f.to_monty(p[j]);
  return t;
}

std::vector<Pair> twist_table(const Monty &f, uint32_t lo, uint32_t hi, bool inverse) {
  const uint32_t lt = hi - lo;
  const size_t size = (size_t)1 << hi;
  const uint64_t mask = size - 1, low_mask = ((uint64_t)1 << lo) - 1;
  std::vector<uint32_t> pw(size);
  const uint32_t g = f.root(hi);
  parallel_chunks(size, [&](size_t b, size_t e) { fill_powers(pw.data(), b, e, g, f.p); });
  std::vector<Pair> t(size);
  parallel_chunks(size, [&](size_t b, size_t e_end) {
    for (size_t i = b; i < e_end; ++i) {  // i = (p << lo) | low
      uint64_t e = ((i & low_mask) * bitrev(i >> lo, lt)) & mask;
      if (inverse) e = (size - e) & mask;
      t[i] = shoup(pw[e], f.p);
    }
  });
  return t;
}

std::vector<Pair> bb_lde_mid_table(uint32_t log_n, uint32_t added_bits, uint32_t shift_std) {
  const size_t n = (size_t)1 << log_n;
  const uint32_t blocks = 1u << added_bits;
  const uint32_t inv_n = invmod((uint32_t)(n % kBBP), kBBP);
  const uint32_t gN = bb_root(log_n + added_bits);
  std::vector<Pair> t(n * blocks);
  for (uint32_t blk = 0; blk < blocks; ++blk) {
    const uint32_t zeta = powmod(gN, bitrev(blk, added_bits), kBBP);
    const uint32_t step = mulmod(shift_std, zeta, kBBP);
    Pair *row = t.data() + blk * n;
    parallel_chunks(n, [&](size_t b, size_t e) {
      uint32_t w = mulmod(inv_n, powmod(step, b, kBBP), kBBP);
      for (size_t j = b; j < e; ++j) {
        row[j] = shoup(w, kBBP);
        w = mulmod(w, step, kBBP);
      }
    });
  }
  return t;
}

std::vector<uint32_t> low_major(const std::vector<uint32_t> &t, uint32_t lo, uint32_t hi) {
  const uint32_t lt = hi - lo;
  const uint64_t low_mask = ((uint64_t)1 << lo) - 1;
  std::vector<uint32_t> r(t.size());
  parallel_chunks(t.size(), [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) r[((i & low_mask) << lt) | (i >> lo)] = t[i];  // i = (p << lo) | low
  });
  return r;
}

std::vector<uint32_t> to_monty_words(const Monty &f, const std::vector<Pair> &t) {
  std::vector<uint32_t> m(t.size());
  parallel_chunks(t.size(), [&](size_t b, size_t e) {
    for (size_t i = b; i < e; ++i) m[i] = f.to_monty(t[i].w);
  });
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
  const size_t low_mask = ((size_t)1 << lo) - 1;
  post.assign(n * blocks, 0);
  pointf.assign(T * blocks, 0);
  std::vector<uint32_t> lane((size_t)1 << lo);
  for (uint32_t blk = 0; blk < blocks; ++blk) {
    const uint32_t sigma = mulmod(shift_std, powmod(gN, bitrev(blk, added_bits), kBBP), kBBP);
    parallel_chunks(lane.size(), [&](size_t b, size_t e) {  // n^-1 * sigma^low
      uint32_t w = mulmod(inv_n, powmod(sigma, b, kBBP), kBBP);
      for (size_t low = b; low < e; ++low) {
        lane[low] = w;
        w = mulmod(w, sigma, kBBP);
      }
    });
    uint32_t *po = post.data() + blk * n;
    parallel_chunks(n, [&](size_t b, size_t e) {  // i = (p << lo) | low
      for (size_t i = b; i < e; ++i) po[i] = bb_to_monty(mulmod(tw[i].w, lane[i & low_mask], kBBP));
    });
    const uint32_t step = powmod(sigma, (uint64_t)1 << lo, kBBP);  // sigma^(2^lo)
    uint32_t *pf = pointf.data() + blk * T;
    parallel_chunks(T, [&](size_t b, size_t e) {
      uint32_t f = powmod(step, b, kBBP);
      for (size_t p = b; p < e; ++p) {
        const size_t at = lt >= 5 ? ((p & ((T >> 5) - 1)) << 5) | (p >> (lt - 5)) : p;
        pf[at] = bb_to_monty(f);
        f = mulmod(f, step, kBBP);
      }
    });
  }
}

std::vector<uint32_t> bb_lde_lane_factors(uint32_t log_n, uint32_t added_bits, uint32_t shift_std, uint32_t lo) {
  const size_t n = (size_t)1 << log_n, L = (size_t)1 << lo;
  const uint32_t blocks = 1u << added_bits;
  const uint32_t inv_n = invmod((uint32_t)(n % kBBP), kBBP);
  const uint32_t gN = bb_root(log_n + added_bits);
  std::vector<uint32_t> t(L * blocks);
  for (uint32_t blk = 0; blk < blocks; ++blk) {
    const uint32_t sigma = mulmod(shift_std, powmod(gN, bitrev(blk, added_bits), kBBP), kBBP);
    uint32_t w = inv_n;
    for (size_t low = 0; low < L; ++low) {
      t[blk * L + low] = bb_to_monty(w);
      w = mulmod(w, sigma, kBBP);
    }
  }
  return t;
}

namespace {
struct Pt {
  uint32_t x, y;
};
inline uint32_t m_add(uint32_t a, uint32_t b) {
  const uint32_t s = a + b;  // < 2^32 for a, b < p
  return s >= kM31P ? s - kM31P : s;
}
inline uint32_t m_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + kM31P - b; }
inline uint32_t m_mul(uint32_t a, uint32_t b) { return mulmod(a, b, kM31P); }
// circle group law (src/field/circle.rs:34-40); identity (1, 0)
inline Pt p_add(Pt a, Pt b) {
  return Pt{m_sub(m_mul(a.x, b.x), m_mul(a.y, b.y)), m_add(m_mul(a.x, b.y), m_mul(a.y, b.x))};
}
inline Pt p_double(Pt a) { return p_add(a, a); }
// k * a by double-and-add (the group is abelian, so chunks may start anywhere)
inline Pt p_mul(Pt a, uint64_t k) {
  Pt r{1u, 0u};
  while (k) {
    if (k & 1) r = p_add(r, a);
    a = p_double(a);
    k >>= 1;
  }
  return r;
}
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
  for (uint32_t l = 0; l < log_n; ++l) {
    const uint32_t ls = log_n - l;  // log size of the current coset
    const size_t half = (size_t)1 << (ls - 1);
    uint32_t *layer = t.data() + (n - ((size_t)1 << ls));
    const bool last = (l == log_n - 1);
    // entry i holds point bitrev(i) of the coset, at(m) = initial + m * step;
    // bitrev is an involution, so point m goes to entry bitrev(m)
    parallel_chunks(half, [&](size_t b, size_t e) {
      Pt cur = p_add(init, p_mul(step, b));
      for (size_t m = b; m < e; ++m) {
        layer[bitrev(m, ls - 1)] = last ? cur.y : cur.x;
        cur = p_add(cur, step);
      }
    });
    init = p_double(init);  // coset doubling (src/field/circle.rs:143-150)
    step = p_double(step);
  }
  if (inverse) {
    // batch inversion (Montgomery's trick) per chunk; zero maps to zero like a^(p-2)
    std::vector<uint32_t> pre(t.size());
    parallel_chunks(t.size(), [&](size_t b, size_t e) {
      uint32_t acc = 1;
      for (size_t i = b; i < e; ++i) {
        pre[i] = acc;
        if (t[i]) acc = m_mul(acc, t[i]);
      }
      uint32_t inv = invmod(acc, kM31P);
      for (size_t i = e; i-- > b;) {
        if (!t[i]) continue;
        const uint32_t v = t[i];
        t[i] = m_mul(inv, pre[i]);
        inv = m_mul(inv, v);
      }
    });
  }
  return t;
}

}  // namespace host
}  // namespace sfn
