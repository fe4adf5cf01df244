"""Device known-answer tests: the reference's device-arithmetic KATs and
hand-derived 2- and 4-point transforms, checked on the GPU through the C ABI.

The reference tests its Metal field kernels directly (m31_add on 1024 pairs
a = 7i + 3, b = 13i + 11; m31_mul on 1024 pairs a = (17i + 5) mod p,
b = (23i + 7) mod p; m31_mul edge cases 0, 1, p - 1, (p - 1)^2, 10^9 * 10^9,
x * 0; tests/gpu_integration.rs:22-142).  This library exposes field
multiplication as sfn_pointwise_mul (NttBackend::pointwise_mul,
src/ntt/mod.rs:46-48) and addition / subtraction through the 2-point
transforms, whose outputs are x + y and x - y (BabyBear / KoalaBear: twiddle 1;
M31: the one twiddle y(G_4) = p - 1, so [a, b] -> [a - b, a + b]).  Expected
values are Python big-integer arithmetic, not the oracle.
"""
import ctypes

import numpy as np
import pytest

import closed_forms as cf
import sfntt
from sfntt import _lib as L

pytestmark = pytest.mark.gpu

M31_P, BB_P, KB_P = 0x7FFFFFFF, 0x78000001, 0x7F000001
R = 1 << 32


@pytest.fixture(scope="module")
def ctx():
    return sfntt.HipContext(0)


def mont(v, p):
    return v * R % p


def edge_pairs(p):
    # tests/gpu_integration.rs:110-111, and the largest canonical values
    a = [0, 1, p - 1, p - 1, 1000000000, 42, p - 2, 2]
    b = [42, 42, p - 1, 1, 1000000000, 0, p - 2, (p + 1) // 2]
    return a, b


def test_m31_device_mul_reference_kats(ctx):
    # tests/gpu_integration.rs:63-101 (sequences) and :104-142 (edge cases)
    n = 1024
    a = [(i * 17 + 5) % M31_P for i in range(n)]
    b = [(i * 23 + 7) % M31_P for i in range(n)]
    ea, eb = edge_pairs(M31_P)
    a, b = a + ea, b + eb
    got = ctx.pointwise_mul(L.FIELD_M31, np.array(a, np.uint32), np.array(b, np.uint32))
    exp = [x * y % M31_P for x, y in zip(a, b)]
    assert got.tolist() == exp
    assert exp[n + 2] == 1  # (p - 1)^2 = 1
    assert exp[n + 4] == 10**18 % M31_P


@pytest.mark.parametrize("field,p", [(L.FIELD_BABYBEAR, BB_P), (L.FIELD_KOALABEAR, KB_P)], ids=["bb", "kb"])
def test_monty_device_mul_kats(ctx, field, p):
    # Montgomery fields: raw values in, raw value of the product out
    n = 1024
    a = [(i * 17 + 5) % p for i in range(n)]
    b = [(i * 23 + 7) % p for i in range(n)]
    ea, eb = edge_pairs(p)
    a, b = a + ea, b + eb
    got = ctx.pointwise_mul(field, np.array([mont(x, p) for x in a], np.uint32),
                            np.array([mont(y, p) for y in b], np.uint32))
    assert got.tolist() == [mont(x * y % p, p) for x, y in zip(a, b)]


@pytest.mark.parametrize("field,p", [(L.FIELD_M31, M31_P), (L.FIELD_BABYBEAR, BB_P), (L.FIELD_KOALABEAR, KB_P)],
                         ids=["m31", "bb", "kb"])
def test_device_add_sub_through_two_point_transforms(ctx, field, p):
    # tests/gpu_integration.rs:22-61: a = 7i + 3, b = 13i + 11, plus wrap-around pairs;
    # 1024 + 6 columns of [a, b] (SoA) in one batched 2-point forward transform
    n = 1024
    a = [i * 7 + 3 for i in range(n)] + [p - 1, p - 1, 0, 0, p - 5, 1]
    b = [i * 13 + 11 for i in range(n)] + [p - 1, 1, p - 1, 0, 7, p - 1]
    conv = (lambda v: v) if field == L.FIELD_M31 else (lambda v: mont(v, p))
    x = np.array([conv(v) for pair in zip(a, b) for v in pair], np.uint32)
    y, _ = ctx.ntt(field, L.VARIANT_AUTO, L.FORWARD, 1, len(a), x)
    y = y.reshape(-1, 2)
    if field == L.FIELD_M31:
        exp = [((u - v) % p, (u + v) % p) for u, v in zip(a, b)]
    else:
        exp = [(mont((u + v) % p, p), mont((u - v) % p, p)) for u, v in zip(a, b)]
    assert [tuple(int(t) for t in row) for row in y] == exp


@pytest.mark.parametrize("field,p", [(L.FIELD_BABYBEAR, BB_P), (L.FIELD_KOALABEAR, KB_P)], ids=["bb", "kb"])
def test_monty_four_point_transform_hand_derived(ctx, field, p):
    # forward: natural input, bit-reversed output [X0, X2, X1, X3] with
    # X_k = sum_j x_j w^(jk), w = the field's 4th root of unity (w^2 = -1)
    w = cf.bb_root(2) if field == L.FIELD_BABYBEAR else cf.kb_root(2)
    assert w * w % p == p - 1
    cols = [[1, 2, 3, 4], [0, 0, 0, 1], [p - 1, 5, 0, 123456789], [7, 7, 7, 7]]
    x = np.array([mont(v, p) for c in cols for v in c], np.uint32)
    y, _ = ctx.ntt(field, L.VARIANT_AUTO, L.FORWARD, 2, len(cols), x)
    for k, (x0, x1, x2, x3) in enumerate(cols):
        X0 = (x0 + x1 + x2 + x3) % p
        X2 = (x0 - x1 + x2 - x3) % p
        X1 = (x0 + w * x1 - x2 - w * x3) % p
        X3 = (x0 - w * x1 - x2 + w * x3) % p
        assert y[4 * k:4 * k + 4].tolist() == [mont(v, p) for v in (X0, X2, X1, X3)], cols[k]
    z, _ = ctx.ntt(field, L.VARIANT_AUTO, L.INVERSE, 2, len(cols), y)
    assert np.array_equal(z, x)


def test_m31_two_and_four_point_transforms_hand_derived(ctx):
    # n = 2: [1, 2] -> [p - 1, 3] (the one twiddle is y(G_4) = p - 1)
    y, _ = ctx.ntt(L.FIELD_M31, L.VARIANT_AUTO, L.FORWARD, 1, 1, np.array([1, 2], np.uint32))
    assert y.tolist() == [M31_P - 1, 3]
    # n = 4 over odds(2) = {G_8 + i G_4}: out[i] = a0 + a1 x(P) + a2 y(2P) + a3 x(P) y(2P)
    # with P the point at bit-reversed index i (layer 0 twiddles x(P), layer 1 y(2P))
    cols = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1], [3, 1, 4, 1]]
    x = np.array([v for c in cols for v in c], np.uint32)
    y, _ = ctx.ntt(L.FIELD_M31, L.VARIANT_AUTO, L.FORWARD, 2, len(cols), x)
    g8, g4 = cf.c_gen(3), cf.c_gen(2)
    for k, (a0, a1, a2, a3) in enumerate(cols):
        exp = []
        for i in range(4):
            P = cf.c_add(g8, cf.c_mul(g4, cf.bitrev(i, 2)))
            x1, y2 = P[0], cf.c_add(P, P)[1]
            exp.append((a0 + a1 * x1 + a2 * y2 + a3 * x1 * y2) % M31_P)
        assert y[4 * k:4 * k + 4].tolist() == exp, cols[k]
    z, _ = ctx.ntt(L.FIELD_M31, L.VARIANT_AUTO, L.INVERSE, 2, len(cols), y)
    assert np.array_equal(z, x)


def test_device_lde_rejects_overlapping_buffers(ctx):
    # the LDE writes column b at b * N while later input columns are unread:
    # any overlap of [in, in + batch*n) and [out, out + batch*N) is refused
    log_n, batch, a = 10, 4, 1
    n_in, n_out = batch << log_n, batch << (log_n + a)
    buf = sfntt.DeviceBuffer(ctx, n_in + n_out)
    base = buf.ptr.value
    lib = ctx._lib
    for fn in (lib.sfn_coset_lde_batch_device, lib.sfn_coset_lde_natural_device):
        for in_off, out_off in [(0, 0), (0, 4 * (n_in - 1)), (4 * n_out - 4, 0), (4 * 1024, 0)]:
            rc = fn(ctx.ptr, log_n, batch, a, 0, L.LAYOUT_SOA, ctypes.c_void_p(base + in_off),
                    ctypes.c_void_p(base + out_off), None)
            assert rc == L.ERR_INVALID_ARGUMENT, (in_off, out_off)
        # adjacent ranges do not overlap
        ctx.check(fn(ctx.ptr, log_n, batch, a, 0, L.LAYOUT_SOA, ctypes.c_void_p(base),
                     ctypes.c_void_p(base + 4 * n_in), None))
    ctx.check(lib.sfn_synchronize(ctx.ptr, None))
    cfg = L.ServerConfig(0, 1, 0, 0, 0)
    srv = ctypes.c_void_p()
    ctx.check(lib.sfn_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(srv)))
    t = ctypes.c_uint64()
    bad = L.Request(L.JOB_LDE, L.FIELD_BABYBEAR, 0, log_n, batch, a, 0, 1, ctypes.c_void_p(base),
                    ctypes.c_void_p(base + 4 * (n_in // 2)))
    assert lib.sfn_server_submit(srv, ctypes.byref(bad), ctypes.byref(t)) == L.ERR_INVALID_ARGUMENT
    good = L.Request(L.JOB_LDE, L.FIELD_BABYBEAR, 0, log_n, batch, a, 0, 1, ctypes.c_void_p(base),
                     ctypes.c_void_p(base + 4 * n_in))
    ctx.check(lib.sfn_server_submit(srv, ctypes.byref(good), ctypes.byref(t)))
    ctx.check(lib.sfn_server_wait(srv, t.value, None))
    lib.sfn_server_destroy(srv)
    buf.free()
