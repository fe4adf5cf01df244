/*
 * sfntt_oracle.h -- CPU restatement of the reference's CPU NTT / LDE paths.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library (libsfntt.so) never
 * links or calls it.
 *
 * Every function restates one reference function; the citation is on the
 * definition in sfntt_oracle.c.  All values are u32; BabyBear values are the
 * raw Montgomery form (R = 2^32) exactly as the reference stores them.
 *
 * Flat twiddle layouts (the reference keeps Vec<Vec<u32>>, one Vec per layer):
 *   BabyBear: layer s (0 <= s < log_n) has 2^s entries at offset 2^s - 1.
 *   M31:      layer l (0 <= l < log_n) has 2^(log_n-l-1) entries at offset
 *             n - 2^(log_n-l).
 * Both tables hold n - 1 entries in total.
 */
#ifndef SFNTT_ORACLE_H
#define SFNTT_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_M31_P 0x7FFFFFFFu
#define ORC_BB_P 0x78000001u
#define ORC_KB_P 0x7F000001u

/* ---- M31 field ---- */
uint32_t orc_m31_reduce(uint64_t v);
uint32_t orc_m31_add(uint32_t a, uint32_t b);
uint32_t orc_m31_sub(uint32_t a, uint32_t b);
uint32_t orc_m31_mul(uint32_t a, uint32_t b);
uint32_t orc_m31_inv(uint32_t a);

/* ---- BabyBear field (Montgomery raw form) ---- */
uint32_t orc_bb_monty_reduce(uint64_t t);
uint32_t orc_bb_to_monty(uint32_t v);
uint32_t orc_bb_from_monty(uint32_t m);
uint32_t orc_bb_add(uint32_t a, uint32_t b);
uint32_t orc_bb_sub(uint32_t a, uint32_t b);
uint32_t orc_bb_mul(uint32_t a, uint32_t b);
uint32_t orc_bb_reduce(uint64_t v);
uint32_t orc_bb_inv(uint32_t a);
uint32_t orc_bb_two_adic_generator(uint32_t log_n); /* returns 0 if log_n > 27 */

/* ---- KoalaBear field (p = 2^31 - 2^24 + 1, Montgomery raw form, R = 2^32) ----
 * The reference lists KoalaBear as planned ("standard NTT", README.md:20) with
 * no code; these restate its BabyBear field and NTT with KoalaBear's
 * constants (two-adicity 24, 2^24-th root 3^127). */
uint32_t orc_kb_monty_reduce(uint64_t t);
uint32_t orc_kb_to_monty(uint32_t v);
uint32_t orc_kb_from_monty(uint32_t m);
uint32_t orc_kb_add(uint32_t a, uint32_t b);
uint32_t orc_kb_sub(uint32_t a, uint32_t b);
uint32_t orc_kb_mul(uint32_t a, uint32_t b);
uint32_t orc_kb_inv(uint32_t a);
uint32_t orc_kb_two_adic_generator(uint32_t log_n); /* returns 0 if log_n > 24 */

/* ---- circle group over M31 ---- */
void orc_circle_odds_point(uint32_t log_size, uint64_t index, uint32_t *x, uint32_t *y);

/* ---- twiddle tables ---- */
int orc_bb_twiddles(uint32_t log_n, int inverse, uint32_t *out);
int orc_kb_twiddles(uint32_t log_n, int inverse, uint32_t *out); /* BabyBear layout */
int orc_m31_twiddles(uint32_t log_n, int inverse, uint32_t *out);

/* ---- single-column transforms, in place; return 0 or -1 (InvalidSize) ---- */
int orc_m31_forward(uint32_t *data, size_t n);
int orc_m31_inverse(uint32_t *data, size_t n);
int orc_bb_forward(uint32_t *data, size_t n);
int orc_bb_inverse(uint32_t *data, size_t n);
int orc_kb_forward(uint32_t *data, size_t n);
int orc_kb_inverse(uint32_t *data, size_t n);

/* ---- batched helpers (columns split over `threads` pthreads) ----
 * field: 0 = M31, 1 = BabyBear, 2 = KoalaBear.  dir: 0 = forward, 1 = inverse.
 * layout: 0 = SoA (data[col*n + row]), 1 = row-major (data[row*batch + col]). */
int orc_ntt_batch(int field, int dir, uint32_t log_n, uint32_t batch, int layout,
                  uint32_t *data, int threads);

/* Coset LDE.  shift_raw = 0 selects the reference shift
 * two_adic_generator(log_n + added_bits + 1); any other value is used as the
 * Montgomery raw shift.  out holds batch * 2^(log_n+added_bits) values. */
int orc_coset_lde_batch(const uint32_t *in, uint32_t log_n, uint32_t batch,
                        uint32_t added_bits, uint32_t shift_raw, int layout,
                        uint32_t *out, int threads);

/* ---- synthetic inputs ----
 * The reference LCG: s = s*6364136223846793005 + 1; v = (s >> 33) % p;
 * monty = 1 stores BabyBear to_monty(v), 2 KoalaBear to_monty(v) (the
 * modulus must match). Returns the final state. */
uint64_t orc_lcg_fill(uint64_t seed, uint32_t modulus, int monty, uint32_t *out, size_t n);
uint64_t orc_splitmix64(uint64_t *state);

#ifdef __cplusplus
}
#endif
#endif
