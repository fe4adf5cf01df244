"""Pins the CPU oracle against the reference's known-answer tests and against
independent closed forms (tests/closed_forms.py).  CPU only."""
import random

import numpy as np
import pytest

import closed_forms as cf

M31_P, BB_P = cf.M31_P, cf.BB_P


# ------------------------------------------------------------ field KATs
def test_m31_reference_kats(orc):
    # src/field/m31.rs:148-203
    assert orc.m31_add(5, 3) == 8
    assert orc.m31_add(M31_P - 1, 1) == 0
    assert orc.m31_add(M31_P - 1, M31_P - 1) == M31_P - 2
    assert orc.m31_sub(8, 3) == 5
    assert orc.m31_sub(0, 1) == M31_P - 1
    assert orc.m31_mul(3, 7) == 21
    assert orc.m31_mul(M31_P - 1, M31_P - 1) == 1
    assert orc.m31_reduce(0) == 0 and orc.m31_reduce(M31_P) == 0
    assert orc.m31_mul(42, orc.m31_inv(42)) == 1
    # tests/gpu_integration.rs:104-142 edge cases against the closed form
    for a, b in [(0, 42), (1, 42), (M31_P - 1, M31_P - 1), (M31_P - 1, 1),
                 (1000000000, 1000000000), (42, 0)]:
        assert orc.m31_mul(a, b) == a * b % M31_P


def test_m31_closed_form_random(orc):
    rng = random.Random(1)
    for _ in range(2000):
        a, b = rng.randrange(M31_P), rng.randrange(M31_P)
        assert orc.m31_mul(a, b) == a * b % M31_P
        assert orc.m31_add(a, b) == (a + b) % M31_P
        assert orc.m31_sub(a, b) == (a - b) % M31_P
    for a in (1, 2, 7, 123456789, M31_P - 1):
        assert orc.m31_inv(a) == pow(a, M31_P - 2, M31_P)


def test_bb_reference_kats(orc):
    # tests/field_vectors.rs:14-72, src/field/babybear.rs:178-364
    tm, fm = orc.bb_to_monty, orc.bb_from_monty
    assert tm(1) == (1 << 32) % BB_P == 268435454
    assert tm(0) == 0 and fm(tm(1)) == 1
    assert fm(orc.bb_mul(tm(123), tm(456))) == 56088
    assert fm(orc.bb_mul(tm(BB_P - 1), tm(BB_P - 1))) == 1
    assert fm(orc.bb_add(tm(BB_P - 1), tm(1))) == 0
    assert fm(orc.bb_sub(0, tm(1))) == BB_P - 1
    assert fm(orc.bb_add(tm(3), tm(7))) == 10
    assert fm(orc.bb_sub(tm(10), tm(3))) == 7
    assert fm(orc.bb_mul(tm(3), tm(7))) == 21
    for v in (0, 1, 2, 42, 1000, BB_P - 1, BB_P - 2):
        assert fm(tm(v)) == v
    for v in (1, 2, 7, 100, 999999, BB_P - 1):
        assert fm(orc.bb_mul(tm(v), orc.bb_inv(tm(v)))) == 1
    seven_inv = fm(orc.bb_inv(tm(7)))
    assert seven_inv == pow(7, BB_P - 2, BB_P)
    assert fm(orc.bb_reduce(21)) == 21 and fm(orc.bb_reduce(BB_P + 5)) == 5
    assert orc.bb_monty_reduce(268435454 * 268435454) == 268435454
    assert (BB_P * 2013265919 + 1) % (1 << 32) == 0


def test_bb_closed_form_random(orc):
    rng = random.Random(2)
    rinv = pow(1 << 32, -1, BB_P)
    for _ in range(2000):
        a, b = rng.randrange(BB_P), rng.randrange(BB_P)
        assert orc.bb_mul(a, b) == a * b * rinv % BB_P
        assert orc.bb_add(a, b) == (a + b) % BB_P
        assert orc.bb_sub(a, b) == (a - b) % BB_P
        assert orc.bb_to_monty(a) == a * (1 << 32) % BB_P


def test_bb_two_adic_generators(orc):
    # tests/field_vectors.rs:74-116: (p-1) = 2^27 * 15, root orders exact
    assert BB_P - 1 == (1 << 27) * 15
    for k in range(1, 28):
        g = orc.bb_from_monty(orc.bb_two_adic_generator(k))
        assert g == cf.bb_root(k)
        assert pow(g, 1 << k, BB_P) == 1 and pow(g, 1 << (k - 1), BB_P) != 1
    assert orc.bb_two_adic_generator(28) == 0


def test_bb_twiddle_tables(orc):
    for log_n in (1, 3, 6, 10):
        fw, iv = orc.bb_twiddles(log_n), orc.bb_twiddles(log_n, inverse=True)
        for s in range(log_n):
            g = cf.bb_root(s + 1)
            for j in range(1 << s):
                assert orc.bb_from_monty(int(fw[(1 << s) - 1 + j])) == pow(g, j, BB_P)
                assert orc.bb_from_monty(int(iv[(1 << s) - 1 + j])) == pow(g, -j, BB_P)


# -------------------------------------------------------------- KoalaBear
# The reference has no KoalaBear code (README.md:20 lists it as planned), so
# the oracle's KoalaBear functions are pinned by closed forms alone:
# big-integer field arithmetic, root orders, Plonky3's published 2^24-th root,
# and the brute-force DFT.
KB_P = cf.KB_P


def test_kb_field_closed_forms(orc):
    assert KB_P == 0x7F000001 and (KB_P - 1) == (1 << 24) * 127
    assert (KB_P * 0x7EFFFFFF + 1) % (1 << 32) == 0  # -p^-1 mod 2^32
    rng = random.Random(3)
    rinv = pow(1 << 32, -1, KB_P)
    vals = [0, 1, 2, KB_P - 1, KB_P - 2, (1 << 31) - 1 - (1 << 24)] + [rng.randrange(KB_P) for _ in range(2000)]
    for a, b in zip(vals, reversed(vals)):
        assert orc.kb_mul(a, b) == a * b * rinv % KB_P
        assert orc.kb_add(a, b) == (a + b) % KB_P
        assert orc.kb_sub(a, b) == (a - b) % KB_P
        assert orc.kb_to_monty(a) == a * (1 << 32) % KB_P
        assert orc.kb_from_monty(orc.kb_to_monty(a)) == a
        if a:
            assert orc.kb_mul(orc.kb_inv(a), a) == orc.kb_to_monty(1)
    assert orc.kb_monty_reduce((KB_P - 1) * (KB_P - 1)) == (KB_P - 1) ** 2 * rinv % KB_P


def test_kb_two_adic_generators(orc):
    for k in range(1, 25):
        g = orc.kb_from_monty(orc.kb_two_adic_generator(k))
        assert g == cf.kb_root(k)
        assert pow(g, 1 << k, KB_P) == 1 and pow(g, 1 << (k - 1), KB_P) != 1
    # Plonky3's KoalaBear TWO_ADIC_GENERATOR (koala-bear/src/koala_bear.rs) is 0x6ac49f88
    assert orc.kb_from_monty(orc.kb_two_adic_generator(24)) == 0x6AC49F88
    assert orc.kb_two_adic_generator(25) == 0


def test_kb_twiddle_tables(orc):
    for log_n in (1, 4, 9):
        fw, iv = orc.kb_twiddles(log_n), orc.kb_twiddles(log_n, inverse=True)
        for s in range(log_n):
            g = cf.kb_root(s + 1)
            for j in range(1 << s):
                assert orc.kb_from_monty(int(fw[(1 << s) - 1 + j])) == pow(g, j, KB_P)
                assert orc.kb_from_monty(int(iv[(1 << s) - 1 + j])) == pow(g, -j, KB_P)


@pytest.mark.parametrize("log_n", range(0, 8))
def test_kb_forward_and_inverse_closed_forms(orc, log_n):
    n = 1 << log_n
    x = orc.lcg_fill(2000 + log_n, n, orc.FIELD_KB)
    assert int(x.max(initial=0)) < KB_P
    assert orc.kb_forward(x).tolist() == cf.kb_forward_closed(x.tolist())
    assert orc.kb_inverse(x).tolist() == cf.kb_inverse_closed(x.tolist())


def test_kb_size_limits(orc):
    with pytest.raises(orc.InvalidSize):
        orc.kb_forward(np.zeros(3, dtype=np.uint32))
    with pytest.raises(orc.InvalidSize):
        orc.ntt_batch(orc.FIELD_KB, False, 25, 1, np.zeros(1 << 25, dtype=np.uint32))


# ----------------------------------------------------------- circle group
def test_circle_kats(orc):
    # src/field/circle.rs:158-230
    x, y = cf.CIRCLE_GEN
    assert (x * x + y * y) % M31_P == 1
    assert cf.c_mul(cf.CIRCLE_GEN, 1 << 31) == (1, 0)
    assert cf.c_mul(cf.CIRCLE_GEN, 1 << 30) != (1, 0)
    for L in (1, 2, 4, 9):
        for i in (0, 1, 3, (1 << L) - 1):
            px, py = orc.circle_odds_point(L, i)
            assert (px * px + py * py) % M31_P == 1
            assert (px, py) == cf.c_add(cf.c_gen(L + 1), cf.c_mul(cf.c_gen(L), i))


# ---------------------------------------------------- transform closed forms
@pytest.mark.parametrize("log_n", range(0, 8))
def test_bb_forward_is_bitreversed_dft(orc, log_n):
    n = 1 << log_n
    x = orc.lcg_fill(1000 + log_n, n, orc.FIELD_BB)
    assert orc.bb_forward(x).tolist() == cf.bb_forward_closed(x.tolist())


@pytest.mark.parametrize("log_n", range(0, 9))
def test_m31_forward_circle_closed_form(orc, log_n):
    n = 1 << log_n
    a = orc.lcg_fill(n * 11111 + 98765, n, orc.FIELD_M31)
    assert orc.m31_forward(a).tolist() == cf.m31_forward_closed(a.tolist())


def test_m31_twiddle_layers_match_closed_form(orc):
    # tw[l][b] = phi_l at the block's first point: x(2^l P) (y on the last layer)
    L = 6
    n = 1 << L
    tw = orc.m31_twiddles(L)
    g2n, gn = cf.c_gen(L + 1), cf.c_gen(L)
    for l in range(L):
        off = n - (1 << (L - l))
        for b in range(1 << (L - l - 1)):
            i = b << (l + 1)  # first output index of the block
            P = cf.c_add(g2n, cf.c_mul(gn, cf.bitrev(i, L)))
            Q = cf.c_mul(P, 1 << l)
            assert int(tw[off + b]) == (Q[1] if l == L - 1 else Q[0])
    itw = orc.m31_twiddles(L, inverse=True)
    assert all(int(a) * int(b) % M31_P == 1 for a, b in zip(tw, itw))


@pytest.mark.parametrize("field", [0, 1, 2])
@pytest.mark.parametrize("log_n", [1, 2, 5, 10, 12])
def test_roundtrip_pins_inverse(orc, field, log_n):
    # forward is pinned by the closed forms above; inverse(forward(x)) == x then
    # fixes inverse uniquely (forward is a bijection).
    n = 1 << log_n
    x = orc.lcg_fill(n * 11111 + 98765, n, field)
    fwd = [orc.m31_forward, orc.bb_forward, orc.kb_forward][field]
    inv = [orc.m31_inverse, orc.bb_inverse, orc.kb_inverse][field]
    y = fwd(x)
    assert not np.array_equal(y, x)
    assert np.array_equal(inv(y), x)
    assert np.array_equal(fwd(inv(x)), x)


def test_reference_transform_kats(orc):
    # M31 n=2: tw[0][0] = y(G_4) = p-1, so fwd([a, b]) = [a - b, a + b]
    assert orc.m31_forward([1, 2]).tolist() == [M31_P - 1, 3]
    # tests/ntt_cross_validation.rs:367-376: [1,0,0,0] -> all ones; zeros -> zeros
    assert orc.m31_forward([1, 0, 0, 0]).tolist() == [1, 1, 1, 1]
    for L in range(1, 9):
        assert not orc.m31_forward(np.zeros(1 << L, np.uint32)).any()
        assert not orc.bb_forward(np.zeros(1 << L, np.uint32)).any()
    # size-1 identity, invalid sizes (src/ntt/test_utils.rs:110-152)
    assert orc.m31_forward([42]).tolist() == [42]
    assert orc.bb_forward([orc.bb_to_monty(42)]).tolist() == [orc.bb_to_monty(42)]
    for bad in ([], [1, 1, 1]):
        for fn in (orc.m31_forward, orc.m31_inverse, orc.bb_forward, orc.bb_inverse):
            with pytest.raises(orc.InvalidSize):
                fn(bad)
    # (1 + 2x)(3 + 4x) = 3 + 10x + 8x^2 (tests/ntt_cross_validation.rs:112-148)
    tm = orc.bb_to_monty
    a = orc.bb_forward([tm(1), tm(2), 0, 0])
    b = orc.bb_forward([tm(3), tm(4), 0, 0])
    c = [orc.bb_mul(int(u), int(v)) for u, v in zip(a, b)]
    assert [orc.bb_from_monty(int(v)) for v in orc.bb_inverse(c)] == [3, 10, 8, 0]
    # linearity (src/ntt/*/cpu_reference.rs test_linearity)
    for field, fwd, add in ((0, orc.m31_forward, orc.m31_add), (1, orc.bb_forward, orc.bb_add)):
        a = [(i * 1 + 1) for i in range(16)]
        b = [(i * 3 + 2) for i in range(16)]
        if field == 1:
            a, b = [tm(v) for v in a], [tm(v) for v in b]
        s = [add(u, v) for u, v in zip(a, b)]
        assert fwd(s).tolist() == [add(int(u), int(v)) for u, v in zip(fwd(a), fwd(b))]


def test_lcg_matches_reference_generator(orc):
    s = 12345 + 65536
    vals = []
    for _ in range(8):
        s = (s * 6364136223846793005 + 1) % (1 << 64)
        vals.append(cf.to_monty((s >> 33) % BB_P))
    assert orc.lcg_fill(12345 + 65536, 8, orc.FIELD_BB).tolist() == vals


# ------------------------------------------------------------- coset LDE
@pytest.mark.parametrize("log_n,added_bits", [(1, 1), (3, 1), (4, 2), (5, 1), (2, 3)])
@pytest.mark.parametrize("shift", [None, 31])
def test_coset_lde_closed_form(orc, log_n, added_bits, shift):
    batch = 3
    n = 1 << log_n
    data = orc.lcg_fill(42, n * batch, orc.FIELD_BB)
    shift_raw = 0 if shift is None else orc.bb_to_monty(shift)
    out = orc.coset_lde_batch(data, log_n, batch, added_bits, shift_raw)
    N = n << added_bits
    for c in range(batch):
        exp = cf.coset_lde_closed(data[c * n:(c + 1) * n].tolist(), added_bits, shift)
        assert out[c * N:(c + 1) * N].tolist() == exp


@pytest.mark.parametrize("log_n,added_bits", [(1, 1), (3, 1), (4, 2)])
@pytest.mark.parametrize("shift", [None, 31])
def test_natural_order_lde_is_rows_bitrev_around_the_reference(orc, log_n, added_bits, shift):
    # The natural-order LDE (sfn_coset_lde_natural) is pinned as: bit-reverse
    # the input rows, run the reference LDE, bit-reverse the output rows; and
    # that composition is checked here against direct evaluation P(s * w_N^k).
    batch = 2
    n = 1 << log_n
    data = orc.lcg_fill(4242, n * batch, orc.FIELD_BB)
    shift_raw = 0 if shift is None else orc.bb_to_monty(shift)
    N = n << added_bits
    for row_major in (False, True):
        x = data if not row_major else data.reshape(batch, n).T.copy().ravel()
        lay = orc.LAYOUT_ROW_MAJOR if row_major else orc.LAYOUT_SOA
        y = orc.coset_lde_batch(cf.rows_bitrev(x, log_n, batch, row_major), log_n, batch, added_bits, shift_raw,
                                layout=lay)
        y = cf.rows_bitrev(y, log_n + added_bits, batch, row_major)
        if row_major:
            y = y.reshape(N, batch).T.ravel()
        for c in range(batch):
            exp = cf.coset_lde_natural_closed(data[c * n:(c + 1) * n].tolist(), added_bits, shift)
            assert y[c * N:(c + 1) * N].tolist() == exp


def test_layouts_are_transposes(orc):
    log_n, batch, a = 6, 5, 1
    n, N = 1 << log_n, 1 << (log_n + a)
    soa = orc.lcg_fill(7, n * batch, orc.FIELD_BB)
    rm = soa.reshape(batch, n).T.copy().ravel()
    out_soa = orc.coset_lde_batch(soa, log_n, batch, a)
    out_rm = orc.coset_lde_batch(rm, log_n, batch, a, layout=orc.LAYOUT_ROW_MAJOR)
    assert np.array_equal(out_rm.reshape(N, batch).T.ravel(), out_soa)
    for field in (0, 1, 2):
        x = orc.lcg_fill(9, n * batch, field)
        y_soa = orc.ntt_batch(field, False, log_n, batch, x)
        y_rm = orc.ntt_batch(field, False, log_n, batch, x.reshape(batch, n).T.copy().ravel(),
                             layout=orc.LAYOUT_ROW_MAJOR)
        assert np.array_equal(y_rm.reshape(n, batch).T.ravel(), y_soa)
        single = np.concatenate([[orc.m31_forward, orc.bb_forward, orc.kb_forward][field](
            x[c * n:(c + 1) * n]) for c in range(batch)])
        assert np.array_equal(single, y_soa)
