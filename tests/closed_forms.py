"""Independent closed forms used to pin the CPU oracle (pure Python ints).

Nothing here restates the reference's butterfly loops: each function states
what the transform *is* mathematically and evaluates it by brute force, so a
dropped term, a wrong twiddle index, a transposed layer order or a wrong
bit-reversal in the oracle (or the GPU path) disagrees with it.
"""
from __future__ import annotations

M31_P = (1 << 31) - 1
BB_P = (1 << 31) - (1 << 27) + 1
BB_ROOT27 = 440564289  # primitive 2^27-th root of unity, standard form
R = 1 << 32


def bitrev(i: int, bits: int) -> int:
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


# ------------------------------------------------------------- BabyBear
def to_monty(v: int) -> int:
    return (v * R) % BB_P


def from_monty(m: int) -> int:
    return (m * pow(R, -1, BB_P)) % BB_P


def bb_root(log_n: int) -> int:
    """Primitive 2^log_n-th root of unity (standard form)."""
    return pow(BB_ROOT27, 1 << (27 - log_n), BB_P)


def bb_forward_closed(raw: list[int]) -> list[int]:
    """out[i] = X[bitrev(i)], X_k = sum_j x_j w^(jk), w = primitive n-th root.
    Values in and out are Montgomery raw (the map is linear, so the closed
    form is applied to the standard values)."""
    n = len(raw)
    L = n.bit_length() - 1
    x = [from_monty(v) for v in raw]
    w = bb_root(L)
    X = [sum(x[j] * pow(w, j * k, BB_P) for j in range(n)) % BB_P for k in range(n)]
    return [to_monty(X[bitrev(i, L)]) for i in range(n)]


# ------------------------------------------------------------ KoalaBear
# p = 2^31 - 2^24 + 1, p - 1 = 2^24 * 127; 3 generates F_p^*, so 3^127 is a
# primitive 2^24-th root of unity.
KB_P = (1 << 31) - (1 << 24) + 1
KB_GEN = 3


def kb_root(log_n: int) -> int:
    """Primitive 2^log_n-th root of unity (standard form), derived from the
    multiplicative generator rather than from any table."""
    return pow(KB_GEN, (KB_P - 1) >> log_n, KB_P)


def kb_to_monty(v: int) -> int:
    return (v * R) % KB_P


def kb_from_monty(m: int) -> int:
    return (m * pow(R, -1, KB_P)) % KB_P


def kb_forward_closed(raw: list[int]) -> list[int]:
    """KoalaBear forward NTT in the BabyBear convention: out[i] = X[bitrev(i)],
    X_k = sum_j x_j w^(jk), Montgomery raw in and out."""
    n = len(raw)
    L = n.bit_length() - 1
    x = [kb_from_monty(v) for v in raw]
    w = kb_root(L)
    X = [sum(x[j] * pow(w, j * k, KB_P) for j in range(n)) % KB_P for k in range(n)]
    return [kb_to_monty(X[bitrev(i, L)]) for i in range(n)]


def kb_inverse_closed(raw: list[int]) -> list[int]:
    """Inverse of kb_forward_closed: bit-reversed evaluations in, coefficients
    out, x_j = n^-1 sum_k y[bitrev(k)] w^(-jk)."""
    n = len(raw)
    L = n.bit_length() - 1
    y = [kb_from_monty(v) for v in raw]
    winv = pow(kb_root(L), -1, KB_P)
    ninv = pow(n, -1, KB_P)
    x = [ninv * sum(y[bitrev(k, L)] * pow(winv, j * k, KB_P) for k in range(n)) % KB_P for j in range(n)]
    return [kb_to_monty(v) for v in x]


def coset_lde_closed(col_raw: list[int], added_bits: int, shift_std: int | None) -> list[int]:
    """Reference LDE semantics for one column: the input holds evaluations of a
    degree < n polynomial P over H_n in bit-reversed order; the output holds P
    evaluated over s*H_N (N = n*2^a) in bit-reversed order."""
    n = len(col_raw)
    L = n.bit_length() - 1
    N = n << added_bits
    LN = L + added_bits
    if shift_std is None:
        shift_std = bb_root(LN + 1)
    ev = [0] * n
    for i, v in enumerate(col_raw):
        ev[bitrev(i, L)] = from_monty(v)  # ev[k] = P(w_n^k)
    wn = bb_root(L)
    inv_n = pow(n, -1, BB_P)
    coef = [inv_n * sum(ev[k] * pow(wn, (-j * k) % n, BB_P) for k in range(n)) % BB_P
            for j in range(n)]
    wN = bb_root(LN)
    out = [0] * N
    for k in range(N):
        pt = shift_std * pow(wN, k, BB_P) % BB_P
        val = 0
        for c in reversed(coef):
            val = (val * pt + c) % BB_P
        out[bitrev(k, LN)] = to_monty(val)
    return out


def coset_lde_natural_closed(col_raw: list[int], added_bits: int, shift_std: int | None) -> list[int]:
    """Natural row order on both sides: input row k = P(w_n^k), output row k =
    P(s * w_N^k) (Plonky3's coset_lde_batch order)."""
    n = len(col_raw)
    L = n.bit_length() - 1
    rev_in = [0] * n
    for k, v in enumerate(col_raw):
        rev_in[bitrev(k, L)] = v
    out_rev = coset_lde_closed(rev_in, added_bits, shift_std)
    LN = L + added_bits
    return [out_rev[bitrev(k, LN)] for k in range(len(out_rev))]


def rows_bitrev(a, log_rows: int, batch: int, row_major: bool):
    """Row r of the result is row bitrev(r) of `a` (numpy, SoA or row-major)."""
    import numpy as np
    idx = np.array([bitrev(r, log_rows) for r in range(1 << log_rows)], dtype=np.int64)
    if row_major:
        return a.reshape(1 << log_rows, batch)[idx].ravel()
    return a.reshape(batch, 1 << log_rows)[:, idx].ravel()


# ------------------------------------------------------------------ M31
CIRCLE_GEN = (2, 1268011823)


def c_add(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % M31_P, (a[0] * b[1] + a[1] * b[0]) % M31_P)


def c_mul(p, k: int):
    r, base = (1, 0), p
    while k:
        if k & 1:
            r = c_add(r, base)
        base = c_add(base, base)
        k >>= 1
    return r


def c_gen(log_size: int):
    """Generator of the circle subgroup of order 2^log_size."""
    return c_mul(CIRCLE_GEN, 1 << (31 - log_size))


def m31_forward_closed(a: list[int]) -> list[int]:
    """Closed form of the reference's circle transform:
        out[i] = sum_k a[k] * prod_{l : bit l of k set} phi_l(P_i)
    with P_i = G_{2n} + bitrev(i, L) * G_n (the i-th point of odds(L) taken in
    bit-reversed order), phi_l(P) = x(2^l P) for l < L-1 and
    phi_{L-1}(P) = y(2^(L-1) P)."""
    n = len(a)
    L = n.bit_length() - 1
    if n == 1:
        return list(a)
    g2n, gn = c_gen(L + 1), c_gen(L)
    out = []
    for i in range(n):
        P = c_add(g2n, c_mul(gn, bitrev(i, L)))
        phi = []
        Q = P
        for l in range(L):
            phi.append(Q[1] if l == L - 1 else Q[0])
            Q = c_add(Q, Q)
        basis = [1] * n
        for k in range(1, n):
            top = k.bit_length() - 1
            basis[k] = basis[k ^ (1 << top)] * phi[top] % M31_P
        out.append(sum(ak * bk for ak, bk in zip(a, basis)) % M31_P)
    return out
