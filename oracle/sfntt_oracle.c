/*
 * sfntt_oracle.c -- CPU restatement of the reference's CPU NTT / LDE paths.
 *
 * TEST INFRASTRUCTURE ONLY (see sfntt_oracle.h).  The product HIP library
 * never links this file.  Written from the reference's algorithm
 * descriptions; each function cites the reference function it restates.
 *
 * Parity pinning (what checks each part independently of this restatement)
 * lives in tests/test_oracle.py:
 *   - field ops: big-integer closed forms and the reference's known values;
 *   - BabyBear NTT: brute-force DFT closed form out[i] = X[bitrev(i)];
 *   - M31 circle NTT: brute-force closed form out[i] = sum_k a[k] *
 *     prod_{l in bits(k)} phi_l(P_i), P_i = odds(L).at(bitrev(i));
 *   - coset LDE: brute-force interpolate-and-evaluate on the shifted coset.
 */
#include "sfntt_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ M31 */

/* src/field/m31.rs:101-115 (M31::reduce): lo + hi fold, one conditional subtract. */
uint32_t orc_m31_reduce(uint64_t v) {
    uint32_t lo = (uint32_t)(v & 0x7FFFFFFFu);
    uint32_t hi = (uint32_t)(v >> 31);
    uint32_t s = lo + hi;
    return s >= ORC_M31_P ? s - ORC_M31_P : s;
}

/* src/field/m31.rs:79-83 */
uint32_t orc_m31_add(uint32_t a, uint32_t b) { return orc_m31_reduce((uint64_t)a + b); }

/* src/field/m31.rs:85-93 */
uint32_t orc_m31_sub(uint32_t a, uint32_t b) {
    return a >= b ? a - b : ORC_M31_P - (b - a);
}

/* src/field/m31.rs:95-99 */
uint32_t orc_m31_mul(uint32_t a, uint32_t b) { return orc_m31_reduce((uint64_t)a * b); }

/* src/field/m31.rs:117-131: Fermat a^(p-2). */
uint32_t orc_m31_inv(uint32_t a) {
    uint32_t r = 1, base = a, e = ORC_M31_P - 2;
    while (e) {
        if (e & 1) r = orc_m31_mul(r, base);
        base = orc_m31_mul(base, base);
        e >>= 1;
    }
    return r;
}

/* ------------------------------------------- Montgomery fields (R = 2^32) */

/* BabyBear and KoalaBear share one restatement of the reference's BabyBear
 * field (src/field/babybear.rs); only the constants differ.  KoalaBear is on
 * the reference's roadmap (README.md:20, "standard NTT") and has no code
 * there: its constants are p = 2^31 - 2^24 + 1, the two-adicity 24 of p - 1 =
 * 2^24 * 127, and the 2^24-th root 3^127 (3 generates F_p^*; 3^127 =
 * 0x6ac49f88 is also Plonky3's KoalaBear two-adic generator). */
typedef struct {
    uint32_t p;
    uint32_t monty;   /* -p^-1 mod 2^32 */
    uint32_t r2;      /* R^2 mod p */
    uint32_t one;     /* R mod p */
    uint32_t root;    /* primitive 2^max_log-th root, Montgomery form */
    uint32_t max_log; /* two-adicity of p - 1 */
} mfield;

static const mfield BB = {ORC_BB_P, 2013265919u /* babybear.rs:25 */, 1172168163u /* babybear.rs:21 */,
                          268435454u /* babybear.rs:18 */, 1476048622u /* to_monty(440564289), babybear.rs:33 */,
                          27};
static const mfield KB = {ORC_KB_P, 0x7EFFFFFFu, 402124772u, 33554430u, 0 /* set by kb_field() */, 24};
#define KB_ROOT24_STD 1791270792u /* 3^127 mod p */

/* src/field/babybear.rs:40-51 */
static uint32_t mf_reduce(const mfield *f, uint64_t t) {
    uint32_t m = (uint32_t)t * f->monty;
    uint64_t mp = (uint64_t)m * f->p;
    /* t + mp < 2^64 for every t = a*b with a, b < 2^32 used here */
    uint32_t r = (uint32_t)((t + mp) >> 32);
    return r >= f->p ? r - f->p : r;
}
static uint32_t mf_to_monty(const mfield *f, uint32_t v) { return mf_reduce(f, (uint64_t)v * f->r2); } /* :54-59 */
static uint32_t mf_add(const mfield *f, uint32_t a, uint32_t b) {                                     /* :113-122 */
    uint32_t s = a + b;
    return s >= f->p ? s - f->p : s;
}
static uint32_t mf_sub(const mfield *f, uint32_t a, uint32_t b) { return a >= b ? a - b : f->p - (b - a); } /* :124-131 */
static uint32_t mf_mul(const mfield *f, uint32_t a, uint32_t b) { return mf_reduce(f, (uint64_t)a * b); } /* :133-137 */
static uint32_t mf_inv(const mfield *f, uint32_t a) {                                                   /* :147-161 */
    uint32_t r = f->one, base = a, e = f->p - 2;
    while (e) {
        if (e & 1) r = mf_mul(f, r, base);
        base = mf_mul(f, base, base);
        e >>= 1;
    }
    return r;
}
/* src/field/babybear.rs:69-76: square the 2^max_log-th root (max_log - log_n) times. */
static uint32_t mf_two_adic_generator(const mfield *f, uint32_t log_n) {
    if (log_n > f->max_log) return 0;
    uint32_t g = f->root;
    for (uint32_t i = 0; i < f->max_log - log_n; ++i) g = mf_mul(f, g, g);
    return g;
}

static const mfield *kb_field(void) {
    static mfield kb;
    static int ready = 0;
    if (!ready) {
        kb = KB;
        kb.root = mf_to_monty(&kb, KB_ROOT24_STD);
        ready = 1;
    }
    return &kb;
}
static const mfield *mfield_of(int field) { return field == 2 ? kb_field() : &BB; }

uint32_t orc_bb_monty_reduce(uint64_t t) { return mf_reduce(&BB, t); }
uint32_t orc_bb_to_monty(uint32_t v) { return mf_to_monty(&BB, v); }
uint32_t orc_bb_from_monty(uint32_t m) { return mf_reduce(&BB, m); } /* src/field/babybear.rs:62-65 */
uint32_t orc_bb_add(uint32_t a, uint32_t b) { return mf_add(&BB, a, b); }
uint32_t orc_bb_sub(uint32_t a, uint32_t b) { return mf_sub(&BB, a, b); }
uint32_t orc_bb_mul(uint32_t a, uint32_t b) { return mf_mul(&BB, a, b); }
uint32_t orc_bb_reduce(uint64_t v) { return mf_to_monty(&BB, (uint32_t)(v % ORC_BB_P)); } /* :139-145 */
uint32_t orc_bb_inv(uint32_t a) { return mf_inv(&BB, a); }
uint32_t orc_bb_two_adic_generator(uint32_t log_n) { return mf_two_adic_generator(&BB, log_n); }

uint32_t orc_kb_monty_reduce(uint64_t t) { return mf_reduce(kb_field(), t); }
uint32_t orc_kb_to_monty(uint32_t v) { return mf_to_monty(kb_field(), v); }
uint32_t orc_kb_from_monty(uint32_t m) { return mf_reduce(kb_field(), m); }
uint32_t orc_kb_add(uint32_t a, uint32_t b) { return mf_add(kb_field(), a, b); }
uint32_t orc_kb_sub(uint32_t a, uint32_t b) { return mf_sub(kb_field(), a, b); }
uint32_t orc_kb_mul(uint32_t a, uint32_t b) { return mf_mul(kb_field(), a, b); }
uint32_t orc_kb_inv(uint32_t a) { return mf_inv(kb_field(), a); }
uint32_t orc_kb_two_adic_generator(uint32_t log_n) { return mf_two_adic_generator(kb_field(), log_n); }

/* ---------------------------------------------------------------- circle */

typedef struct { uint32_t x, y; } cpt;

/* src/field/circle.rs:34-40 */
static cpt cpt_add(cpt a, cpt b) {
    cpt r;
    r.x = orc_m31_sub(orc_m31_mul(a.x, b.x), orc_m31_mul(a.y, b.y));
    r.y = orc_m31_add(orc_m31_mul(a.x, b.y), orc_m31_mul(a.y, b.x));
    return r;
}

/* src/field/circle.rs:53-60 */
static cpt cpt_double(cpt a) {
    cpt r;
    r.x = orc_m31_sub(orc_m31_mul(2, orc_m31_mul(a.x, a.x)), 1);
    r.y = orc_m31_mul(2, orc_m31_mul(a.x, a.y));
    return r;
}

/* src/field/circle.rs:63-74 (double-and-add using the group law) */
static cpt cpt_mul_scalar(cpt p, uint64_t k) {
    cpt r = {1, 0}, base = p;
    while (k) {
        if (k & 1) r = cpt_add(r, base);
        base = cpt_add(base, base);
        k >>= 1;
    }
    return r;
}

/* src/field/circle.rs:78-85: CIRCLE_GEN = (2, 1268011823) doubled 31-k times. */
static cpt cpt_subgroup_gen(uint32_t log_size) {
    cpt p = {2u, 1268011823u};
    for (uint32_t i = 0; i < 31 - log_size; ++i) p = cpt_double(p);
    return p;
}

typedef struct { cpt initial, step; uint32_t log_size; } coset;

/* src/field/circle.rs:108-117 */
static coset coset_odds(uint32_t log_size) {
    coset c;
    c.step = cpt_subgroup_gen(log_size);
    c.initial = cpt_subgroup_gen(log_size + 1);
    c.log_size = log_size;
    return c;
}

/* src/field/circle.rs:136-138 */
static cpt coset_at(const coset *c, uint64_t i) {
    return cpt_add(c->initial, cpt_mul_scalar(c->step, i));
}

/* src/field/circle.rs:143-150 */
static coset coset_double(const coset *c) {
    coset d;
    d.initial = cpt_double(c->initial);
    d.step = cpt_double(c->step);
    d.log_size = c->log_size - 1;
    return d;
}

void orc_circle_odds_point(uint32_t log_size, uint64_t index, uint32_t *x, uint32_t *y) {
    coset c = coset_odds(log_size);
    cpt p = coset_at(&c, index);
    *x = p.x;
    *y = p.y;
}

/* src/ntt/m31/twiddles.rs:71-79 */
static uint64_t bitrev(uint64_t v, uint32_t bits) {
    uint64_t r = 0;
    for (uint32_t i = 0; i < bits; ++i) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

/* ------------------------------------------------------------- twiddles */

/* src/field/babybear.rs:83-97 (forward) and src/ntt/babybear/twiddles.rs:13-19
 * (inverse = element-wise inv of the forward table). */
static int mf_twiddles(const mfield *f, uint32_t log_n, int inverse, uint32_t *out) {
    if (log_n > f->max_log) return -1;
    for (uint32_t s = 0; s < log_n; ++s) {
        size_t half = (size_t)1 << s;
        uint32_t g = mf_two_adic_generator(f, s + 1);
        uint32_t w = f->one;
        uint32_t *layer = out + (half - 1);
        for (size_t j = 0; j < half; ++j) {
            layer[j] = inverse ? mf_inv(f, w) : w;
            w = mf_mul(f, w, g);
        }
    }
    return 0;
}
int orc_bb_twiddles(uint32_t log_n, int inverse, uint32_t *out) { return mf_twiddles(&BB, log_n, inverse, out); }
int orc_kb_twiddles(uint32_t log_n, int inverse, uint32_t *out) { return mf_twiddles(kb_field(), log_n, inverse, out); }

/* src/ntt/m31/twiddles.rs:19-43 (forward) and :47-68 (inverse): coset odds(log_n),
 * layer l takes x (or y on the last layer) of current.at(bitrev(i, log_size-1)),
 * then the coset doubles. */
int orc_m31_twiddles(uint32_t log_n, int inverse, uint32_t *out) {
    if (log_n > 30) return -1;
    size_t n = (size_t)1 << log_n;
    coset cur = coset_odds(log_n);
    for (uint32_t l = 0; l < log_n; ++l) {
        size_t half = ((size_t)1 << cur.log_size) / 2;
        int last = (l == log_n - 1);
        uint32_t *layer = out + (n - ((size_t)1 << (log_n - l)));
        for (size_t i = 0; i < half; ++i) {
            cpt p = coset_at(&cur, bitrev(i, cur.log_size - 1));
            uint32_t t = last ? p.y : p.x;
            layer[i] = inverse ? orc_m31_inv(t) : t;
        }
        cur = coset_double(&cur);
    }
    return 0;
}

static int log2_exact(size_t n, uint32_t *log_n) {
    if (n == 0 || (n & (n - 1))) return -1;
    uint32_t l = 0;
    while (((size_t)1 << l) < n) ++l;
    *log_n = l;
    return 0;
}

/* ------------------------------------------------------------ transforms */

/* src/ntt/m31/cpu_reference.rs:46-76 (forward: layers log_n-1..0, CT butterfly
 * with the block-indexed twiddle, :134-139). */
static void m31_forward_tw(uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    for (int layer = (int)log_n - 1; layer >= 0; --layer) {
        size_t half = (size_t)1 << layer, full = half << 1;
        const uint32_t *t = tw + (n - ((size_t)1 << (log_n - layer)));
        for (size_t b = 0; b < n / full; ++b) {
            uint32_t w = t[b];
            for (size_t j = 0; j < half; ++j) {
                size_t i0 = b * full + j, i1 = i0 + half;
                uint32_t tmp = orc_m31_mul(data[i1], w);
                uint32_t v0 = data[i0];
                data[i0] = orc_m31_add(v0, tmp);
                data[i1] = orc_m31_sub(v0, tmp);
            }
        }
    }
}

int orc_m31_forward(uint32_t *data, size_t n) {
    uint32_t log_n;
    if (log2_exact(n, &log_n)) return -1;
    if (n == 1) return 0;
    uint32_t *tw = (uint32_t *)malloc((n - 1) * sizeof(uint32_t));
    orc_m31_twiddles(log_n, 0, tw);
    m31_forward_tw(data, n, log_n, tw);
    free(tw);
    return 0;
}

/* src/ntt/m31/cpu_reference.rs:78-114 (inverse: layers 0..log_n-1, GS butterfly
 * :145-150, then multiply by inv(n)). */
static void m31_inverse_tw(uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    for (uint32_t layer = 0; layer < log_n; ++layer) {
        size_t half = (size_t)1 << layer, full = half << 1;
        const uint32_t *t = tw + (n - ((size_t)1 << (log_n - layer)));
        for (size_t b = 0; b < n / full; ++b) {
            uint32_t w = t[b];
            for (size_t j = 0; j < half; ++j) {
                size_t i0 = b * full + j, i1 = i0 + half;
                uint32_t v0 = data[i0], v1 = data[i1];
                data[i0] = orc_m31_add(v0, v1);
                data[i1] = orc_m31_mul(orc_m31_sub(v0, v1), w);
            }
        }
    }
    uint32_t inv_n = orc_m31_inv(orc_m31_reduce((uint64_t)n));
    for (size_t i = 0; i < n; ++i) data[i] = orc_m31_mul(data[i], inv_n);
}

int orc_m31_inverse(uint32_t *data, size_t n) {
    uint32_t log_n;
    if (log2_exact(n, &log_n)) return -1;
    if (n == 1) return 0;
    uint32_t *tw = (uint32_t *)malloc((n - 1) * sizeof(uint32_t));
    orc_m31_twiddles(log_n, 1, tw);
    m31_inverse_tw(data, n, log_n, tw);
    free(tw);
    return 0;
}

/* src/ntt/babybear/cpu_reference.rs:34-62 (forward_ntt_u32: GS-DIF, layers
 * log_n-1..0, position-indexed twiddle tw[layer][j]). */
static void mf_forward_tw(const mfield *f, uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    for (int layer = (int)log_n - 1; layer >= 0; --layer) {
        size_t half = (size_t)1 << layer, full = half << 1;
        const uint32_t *t = tw + (half - 1);
        for (size_t b = 0; b < n / full; ++b) {
            for (size_t j = 0; j < half; ++j) {
                size_t i0 = b * full + j, i1 = i0 + half;
                uint32_t v0 = data[i0], v1 = data[i1];
                data[i0] = mf_add(f, v0, v1);
                data[i1] = mf_mul(f, mf_sub(f, v0, v1), t[j]);
            }
        }
    }
}

/* src/ntt/babybear/cpu_reference.rs:69-103 (inverse_ntt_u32: CT-DIT, layers
 * 0..log_n-1, then multiply by inv(n)). */
static void mf_inverse_tw(const mfield *f, uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    for (uint32_t layer = 0; layer < log_n; ++layer) {
        size_t half = (size_t)1 << layer, full = half << 1;
        const uint32_t *t = tw + (half - 1);
        for (size_t b = 0; b < n / full; ++b) {
            for (size_t j = 0; j < half; ++j) {
                size_t i0 = b * full + j, i1 = i0 + half;
                uint32_t tmp = mf_mul(f, data[i1], t[j]);
                uint32_t v0 = data[i0];
                data[i0] = mf_add(f, v0, tmp);
                data[i1] = mf_sub(f, v0, tmp);
            }
        }
    }
    uint32_t inv_n = mf_inv(f, mf_to_monty(f, (uint32_t)(n % f->p)));
    for (size_t i = 0; i < n; ++i) data[i] = mf_mul(f, data[i], inv_n);
}

static void bb_forward_tw(uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    mf_forward_tw(&BB, data, n, log_n, tw);
}
static void bb_inverse_tw(uint32_t *data, size_t n, uint32_t log_n, const uint32_t *tw) {
    mf_inverse_tw(&BB, data, n, log_n, tw);
}

static int mf_single(const mfield *f, int inverse, uint32_t *data, size_t n) {
    uint32_t log_n;
    if (log2_exact(n, &log_n)) return -1;
    if (n == 1) return 0;
    if (log_n > f->max_log) return -1;
    uint32_t *tw = (uint32_t *)malloc((n - 1) * sizeof(uint32_t));
    mf_twiddles(f, log_n, inverse, tw);
    if (inverse) mf_inverse_tw(f, data, n, log_n, tw);
    else mf_forward_tw(f, data, n, log_n, tw);
    free(tw);
    return 0;
}
int orc_bb_forward(uint32_t *data, size_t n) { return mf_single(&BB, 0, data, n); }
int orc_bb_inverse(uint32_t *data, size_t n) { return mf_single(&BB, 1, data, n); }
int orc_kb_forward(uint32_t *data, size_t n) { return mf_single(kb_field(), 0, data, n); }
int orc_kb_inverse(uint32_t *data, size_t n) { return mf_single(kb_field(), 1, data, n); }

/* ------------------------------------------------------- batched helpers */

typedef struct {
    int field, dir;
    size_t n;
    uint32_t log_n;
    const uint32_t *tw; /* table built once per call, like the reference's TwiddleCache */
    uint32_t *data;     /* SoA */
    uint32_t c0, c1;
} batch_job;

static void *batch_worker(void *arg) {
    batch_job *j = (batch_job *)arg;
    for (uint32_t c = j->c0; c < j->c1; ++c) {
        uint32_t *col = j->data + (size_t)c * j->n;
        if (j->n == 1) continue;
        if (j->field == 0)
            j->dir == 0 ? m31_forward_tw(col, j->n, j->log_n, j->tw) : m31_inverse_tw(col, j->n, j->log_n, j->tw);
        else if (j->dir == 0)
            mf_forward_tw(mfield_of(j->field), col, j->n, j->log_n, j->tw);
        else
            mf_inverse_tw(mfield_of(j->field), col, j->n, j->log_n, j->tw);
    }
    return NULL;
}

static void transpose(const uint32_t *src, uint32_t *dst, size_t rows, size_t cols) {
    /* src is rows x cols (row-major), dst is cols x rows */
    for (size_t r = 0; r < rows; ++r)
        for (size_t c = 0; c < cols; ++c) dst[c * rows + r] = src[r * cols + c];
}

static void run_columns(int field, int dir, size_t n, uint32_t batch, uint32_t *soa, int threads) {
    uint32_t log_n = 0;
    while (((size_t)1 << log_n) < n) ++log_n;
    uint32_t *tw = (uint32_t *)malloc((n > 1 ? n - 1 : 1) * sizeof(uint32_t));
    if (n > 1) {
        if (field == 0) orc_m31_twiddles(log_n, dir, tw);
        else mf_twiddles(mfield_of(field), log_n, dir, tw);
    }
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > batch) threads = (int)batch;
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    batch_job *jobs = (batch_job *)calloc((size_t)threads, sizeof(batch_job));
    for (int t = 0; t < threads; ++t) {
        jobs[t].field = field;
        jobs[t].dir = dir;
        jobs[t].n = n;
        jobs[t].log_n = log_n;
        jobs[t].tw = tw;
        jobs[t].data = soa;
        jobs[t].c0 = (uint32_t)((uint64_t)batch * t / threads);
        jobs[t].c1 = (uint32_t)((uint64_t)batch * (t + 1) / threads);
    }
    if (threads == 1) {
        batch_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, batch_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    }
    free(tid);
    free(jobs);
    free(tw);
}

/* The reference's batch APIs run the single-column transform on every column
 * of an SoA buffer (src/ntt/babybear/batch.rs:76-180 is checked against the
 * single-column path at :438-480); the row-major layout is the transpose. */
int orc_ntt_batch(int field, int dir, uint32_t log_n, uint32_t batch, int layout,
                  uint32_t *data, int threads) {
    if (field < 0 || field > 2 || batch == 0) return -1;
    if ((field == 0 && log_n > 30) || (field != 0 && log_n > mfield_of(field)->max_log)) return -1;
    size_t n = (size_t)1 << log_n;
    if (layout == 0) {
        run_columns(field, dir, n, batch, data, threads);
        return 0;
    }
    uint32_t *tmp = (uint32_t *)malloc(n * batch * sizeof(uint32_t));
    if (!tmp) return -1;
    transpose(data, tmp, n, batch);
    run_columns(field, dir, n, batch, tmp, threads);
    transpose(tmp, data, batch, n);
    free(tmp);
    return 0;
}

typedef struct {
    const uint32_t *in; /* SoA, n per column */
    uint32_t *out;      /* SoA, n_ext per column */
    size_t n, n_ext;
    uint32_t log_n, log_ext;
    const uint32_t *itw, *ftw; /* inverse size-n and forward size-n_ext tables, built once */
    const uint32_t *shift_pow;
    uint32_t c0, c1;
} lde_job;

/* src/lde/mod.rs:127-176 (coset_lde_batch_cpu): inverse_ntt_u32 per column,
 * zero-pad to n_ext, multiply row i by shift^i, forward_ntt_u32 per column. */
static void *lde_worker(void *arg) {
    lde_job *j = (lde_job *)arg;
    for (uint32_t c = j->c0; c < j->c1; ++c) {
        uint32_t *dst = j->out + (size_t)c * j->n_ext;
        memcpy(dst, j->in + (size_t)c * j->n, j->n * sizeof(uint32_t));
        if (j->n > 1) bb_inverse_tw(dst, j->n, j->log_n, j->itw);
        memset(dst + j->n, 0, (j->n_ext - j->n) * sizeof(uint32_t));
        for (size_t i = 0; i < j->n_ext; ++i) dst[i] = orc_bb_mul(dst[i], j->shift_pow[i]);
        if (j->n_ext > 1) bb_forward_tw(dst, j->n_ext, j->log_ext, j->ftw);
    }
    return NULL;
}

int orc_coset_lde_batch(const uint32_t *in, uint32_t log_n, uint32_t batch,
                        uint32_t added_bits, uint32_t shift_raw, int layout,
                        uint32_t *out, int threads) {
    uint32_t log_ext = log_n + added_bits;
    if (batch == 0 || log_ext + 1 > 27) return -1;
    size_t n = (size_t)1 << log_n, n_ext = (size_t)1 << log_ext;
    /* src/lde/mod.rs:158 shift = two_adic_generator(log_n_ext + 1); powers as
     * in compute_shift_powers (src/lde/mod.rs:116-124). */
    uint32_t shift = shift_raw ? shift_raw : orc_bb_two_adic_generator(log_ext + 1);
    uint32_t *pw = (uint32_t *)malloc(n_ext * sizeof(uint32_t));
    uint32_t cur = BB.one;
    for (size_t i = 0; i < n_ext; ++i) {
        pw[i] = cur;
        cur = orc_bb_mul(cur, shift);
    }
    uint32_t *itw = (uint32_t *)malloc(n * sizeof(uint32_t));
    uint32_t *ftw = (uint32_t *)malloc(n_ext * sizeof(uint32_t));
    orc_bb_twiddles(log_n, 1, itw);
    orc_bb_twiddles(log_ext, 0, ftw);
    const uint32_t *src = in;
    uint32_t *soa_in = NULL, *soa_out = out;
    if (layout == 1) {
        soa_in = (uint32_t *)malloc(n * batch * sizeof(uint32_t));
        soa_out = (uint32_t *)malloc(n_ext * batch * sizeof(uint32_t));
        transpose(in, soa_in, n, batch);
        src = soa_in;
    }
    if (threads < 1) threads = 1;
    if ((uint32_t)threads > batch) threads = (int)batch;
    pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
    lde_job *jobs = (lde_job *)calloc((size_t)threads, sizeof(lde_job));
    for (int t = 0; t < threads; ++t) {
        jobs[t].in = src;
        jobs[t].out = soa_out;
        jobs[t].n = n;
        jobs[t].n_ext = n_ext;
        jobs[t].log_n = log_n;
        jobs[t].log_ext = log_ext;
        jobs[t].itw = itw;
        jobs[t].ftw = ftw;
        jobs[t].shift_pow = pw;
        jobs[t].c0 = (uint32_t)((uint64_t)batch * t / threads);
        jobs[t].c1 = (uint32_t)((uint64_t)batch * (t + 1) / threads);
    }
    if (threads == 1) {
        lde_worker(&jobs[0]);
    } else {
        for (int t = 0; t < threads; ++t) pthread_create(&tid[t], NULL, lde_worker, &jobs[t]);
        for (int t = 0; t < threads; ++t) pthread_join(tid[t], NULL);
    }
    if (layout == 1) {
        transpose(soa_out, out, batch, n_ext);
        free(soa_in);
        free(soa_out);
    }
    free(tid);
    free(jobs);
    free(pw);
    free(itw);
    free(ftw);
    return 0;
}

/* ------------------------------------------------------------ generators */

/* benches/ntt_benchmark.rs:53-61 and src/ntt/test_utils.rs:38-46 (M31);
 * benches/ntt_benchmark.rs:482-490, :653-661 (BabyBear, to_monty). */
uint64_t orc_lcg_fill(uint64_t seed, uint32_t modulus, int monty, uint32_t *out, size_t n) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; ++i) {
        s = s * 6364136223846793005ull + 1ull;
        uint32_t v = (uint32_t)(s >> 33) % modulus;
        out[i] = monty == 2 ? orc_kb_to_monty(v) : monty ? orc_bb_to_monty(v) : v;
    }
    return s;
}

/* SplitMix64 (Steele, Lea, Flood 2014), used to draw job shapes. */
uint64_t orc_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
