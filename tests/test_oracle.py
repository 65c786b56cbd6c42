"""Pins the CPU oracle (oracle/) against the reference's known answers and
against closed forms computed independently in tests/gfpy.py.

Every check here is chosen so that a plausible slip in the oracle (wrong
twiddle index, swapped butterfly operands, wrong chunk shift, dropped term)
fails at least one of them.
"""
import itertools

import numpy as np
import pytest

import gfpy
import oracle as O

T_ = O.tables()
BASIS = [int(v) for v in T_["basis"]]
SUB = [int(v) for v in T_["subspace"]]


# ---------------------------------------------------------------- field KATs
def test_field_kats():
    # gf2p8/src/poly_11d.rs:62-82
    assert O.mul(2, 0x80) == 0x1D
    assert O.mul(0x57, 0) == 0 and O.mul(0x57, 1) == 0x57
    assert 0x57 ^ 0x83 == 0xD4


def test_mul_exhaustive():
    # rs/src/poly_11d_lut.rs:58-66 (mul_lut == mul), here over all pairs and
    # against the independent textbook multiply.
    for a in range(256):
        for b in range(0, 256, 1):
            m = int(gfpy.MUL[a, b])
            assert O.mul(a, b) == m
            assert O.mul_lut(a, b) == m


def test_exp_log_inv_tables():
    exp, log, inv = T_["exp"], T_["log"], T_["inv"]
    for i in range(255):
        assert exp[i] == gfpy.gpow(2, i)
        assert exp[255 + i] == exp[i]
        assert log[exp[i]] == i
    assert inv[0] == 0
    for a in range(1, 256):
        assert gfpy.gmul(a, int(inv[a])) == 1


def test_gfni_matrices_exhaustive():
    # rs/src/poly_11d_lut.rs:68-81: matrix applied as vgf2p8affineqb does
    # (result bit i = parity(A.byte[7-i] & x)) equals the field product.
    gf = T_["gfni"]
    for t in range(256):
        A = int(gf[t])
        rows = [(A >> (8 * (7 - i))) & 0xFF for i in range(8)]
        for x in range(256):
            r = 0
            for i in range(8):
                r |= (bin(rows[i] & x).count("1") & 1) << i
            assert r == gfpy.MUL[t, x]


# ---------------------------------------------------------- basis and s_j
def test_cantor_basis_kat_and_properties():
    # KAT rs/src/poly_11d_lut.rs:419-429
    assert BASIS == [1, 214, 152, 146, 86, 200, 88, 231]
    assert [SUB[1 << i] for i in range(8)] == BASIS
    # chain v_i^2 + v_i = v_{i-1} and trace 0 (gf2p8/src/poly_11d.rs:99-129)
    for i in range(1, 8):
        assert gfpy.gmul(BASIS[i], BASIS[i]) ^ BASIS[i] == BASIS[i - 1]
    for i in range(7):
        assert gfpy.trace(BASIS[i]) == 0
    # GF(2)-independence: the span has 256 distinct points
    assert sorted(SUB) == list(range(256))
    # omega_i = xor of basis elements over set bits (gf2p8/src/field.rs:218-226)
    for i in range(256):
        assert SUB[i] == gfpy.span(BASIS, 8)[i]


def test_subspace_polys_bruteforce():
    # gf2p8/src/poly_11d.rs:187-202
    spv = T_["spv"]
    for j in range(9):
        for x in range(256):
            assert spv[j, x] == gfpy.subspace_poly(BASIS, j, x), (j, x)


def test_normalisation_and_derivative_constants():
    spv = T_["spv"]
    # rs/src/poly_11d_lut.rs:123-129: s_j(v_j) = 1
    for j in range(8):
        assert spv[j, BASIS[j]] == 1
    # gf2p8/src/poly_11d.rs:204-213: product of nonzero elements of V_i = 1
    for i in range(9):
        p = 1
        for a in gfpy.span(BASIS, i):
            if a:
                p = gfpy.gmul(p, a)
        assert p == 1
    # s_j is GF(2)-linear
    for j in range(9):
        for x, y in [(3, 77), (200, 5), (255, 128)]:
            assert spv[j, x ^ y] == spv[j, x] ^ spv[j, y]


def test_subspace_poly_coeffs():
    # gf2p8/src/field.rs:381-395: s_{j+1} = s_j^2 + s_j on coefficient masks
    m = 1
    masks = []
    for _ in range(9):
        masks.append(m >> 1)
        m = (m << 1) ^ m
    assert list(T_["coeffs"]) == [x & 0xFF for x in masks]


# -------------------------------------------------------------- transforms
def fft_closed_form(d, k, beta):
    n = 1 << k
    out = []
    for i in range(n):
        x = beta ^ SUB[i]
        acc = 0
        for j in range(n):
            if d[j]:
                acc ^= gfpy.gmul(d[j], gfpy.lch_basis(BASIS, j, x))
        out.append(acc)
    return out


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_fft_matches_closed_form(k):
    rng = np.random.default_rng(100 + k)
    n = 1 << k
    for beta_idx in [0, n, 3 * n % 256, 255 & ~(n - 1)]:
        beta = SUB[beta_idx]
        d = rng.integers(0, 256, n, dtype=np.uint8)
        got = d.copy()
        O.fft_sharded(got, 1, k, beta)
        assert list(got) == fft_closed_form([int(v) for v in d], k, beta)
        back = got.copy()
        O.ifft_sharded(back, 1, k, beta)
        assert np.array_equal(back, d)


def test_sharded_equals_columnwise():
    rng = np.random.default_rng(7)
    k, L = 5, 37
    n = 1 << k
    sh = rng.integers(0, 256, n * L, dtype=np.uint8)
    beta = SUB[64]
    a = sh.copy()
    O.fft_sharded(a, L, k, beta)
    b = sh.reshape(n, L).copy()
    for c in range(L):
        col = np.ascontiguousarray(b[:, c])
        O.fft_sharded(col, 1, k, beta)
        b[:, c] = col
    assert np.array_equal(a.reshape(n, L), b)


@pytest.mark.skipif(not O.gfni_supported(), reason="host lacks AVX-512BW+GFNI")
@pytest.mark.parametrize("unrolled", [True, False])
@pytest.mark.parametrize("L", [1, 63, 64, 65, 128, 200])
def test_gfni_transforms_equal_lut(L, unrolled):
    # rs/src/kernel/gfni_kernel.rs:324-387; beta = 0 through the generated
    # fixed sequences (the reference's dispatch) and through the recursion
    O.gfni_set_unrolled(unrolled)
    try:
        for k in range(1, 9):
            n = 1 << k
            sh = np.array([(i * 37 + j * 13 + 1) & 0xFF for i in range(n) for j in range(L)], np.uint8)
            for beta in (0, SUB[n % 256]):
                a, b = sh.copy(), sh.copy()
                O.fft_sharded(a, L, k, beta)
                O.fft_sharded(b, L, k, beta, gfni=True)
                assert np.array_equal(a, b), (k, beta)
                O.ifft_sharded(a, L, k, beta)
                O.ifft_sharded(b, L, k, beta, gfni=True)
                assert np.array_equal(a, b) and np.array_equal(a, sh), (k, beta)
    finally:
        O.gfni_set_unrolled(True)


# ------------------------------------------------------------------ encode
def lagrange_parity(n, k, msg, L):
    T = n - k
    xs = [SUB[T + j] for j in range(k)]
    M = gfpy.lagrange_eval_matrix(xs, [SUB[i] for i in range(T)])
    return gfpy.apply_matrix(M, msg.reshape(k, L))


POW2_CODES = [(n, T) for n in (2, 4, 8, 16, 32) for T in (1, 2, 4, 8, 16) if T < n]


@pytest.mark.parametrize("n,T", POW2_CODES)
def test_encode_equals_lagrange(n, T):
    rng = np.random.default_rng(n * 1000 + T)
    L = 3
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = O.encode(n, T, msg, L)
    assert np.array_equal(par.reshape(T, L), lagrange_parity(n, k, msg, L))
    if O.gfni_supported():
        assert np.array_equal(O.encode(n, T, msg, L, gfni=True), par)


def test_encode_large_equals_lagrange():
    rng = np.random.default_rng(99)
    for n, T in [(64, 32), (256, 128), (256, 32)]:
        L = 2
        k = n - T
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        par = O.encode(n, T, msg, L)
        assert np.array_equal(par.reshape(T, L), lagrange_parity(n, k, msg, L)), (n, T)


def test_encode_accumulates_parity():
    # rs/src/codec.rs:259-264: parity is XORed into, then FFT_T(., 0) applied;
    # by linearity the result is encode(msg) ^ FFT_T(parity_in, 0).
    rng = np.random.default_rng(5)
    n, T, L = 32, 8, 11
    msg = rng.integers(0, 256, (n - T) * L, dtype=np.uint8)
    pin = rng.integers(0, 256, T * L, dtype=np.uint8)
    got = O.encode(n, T, msg, L, parity=pin.copy())
    f = pin.copy()
    O.fft_sharded(f, L, 3, 0)
    assert np.array_equal(got, O.encode(n, T, msg, L) ^ f)


def test_encode_sharded_generalizes_scalar_pattern():
    # rs/src/poly_11d_lut.rs:453-469 with the message pattern of :100-121
    for n in (4, 8, 16, 32, 64, 128, 256):
        T = n // 2
        k = n - T
        msg = np.array([(T + i) & 0xFF for i in range(k)], np.uint8)
        par = O.encode(n, T, msg, 1)
        assert np.array_equal(par.reshape(T, 1), lagrange_parity(n, k, msg, 1))


def test_ext_encode_kpow2():
    rng = np.random.default_rng(3)
    for n, k in [(8, 2), (16, 4), (32, 8), (256, 16)]:
        L = 2
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        par = O.ext_encode_kpow2(n, k, msg, L)
        assert np.array_equal(par.reshape(n - k, L), lagrange_parity(n, k, msg, L)), (n, k)
    # agrees with the reference algorithm where both apply (k == T)
    for n in (4, 16, 64, 256):
        k = n // 2
        msg = rng.integers(0, 256, k * 5, dtype=np.uint8)
        assert np.array_equal(O.ext_encode_kpow2(n, k, msg, 5), O.encode(n, n - k, msg, 5))


# ----------------------------------------------------------------- recover
def make_codeword(n, T, L, rng):
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = O.encode_any(n, k, msg, L)
    return np.concatenate([par, msg])


REF_RECOVER_CODES = [(2, 1), (4, 1), (4, 2), (8, 1), (8, 2), (8, 4), (16, 1), (16, 2), (16, 4),
                     (16, 8), (32, 1), (32, 2), (32, 4), (32, 8), (32, 16), (64, 1), (64, 2),
                     (64, 4), (64, 8), (64, 16), (64, 32), (128, 1), (128, 2), (128, 4),
                     (128, 8), (128, 16), (128, 32), (128, 64), (256, 1), (256, 2), (256, 4),
                     (256, 8), (256, 16), (256, 32), (256, 64), (256, 128)]


@pytest.mark.parametrize("n,T", REF_RECOVER_CODES)
def test_recover_reference_pattern(n, T):
    # rs/src/poly_11d_lut.rs:496-566: L = 1, erasures at 1, 3, ..., 2T-1,
    # message bytes T + i (the reference's generate_sharded_lch_codeword).
    k = n - T
    msg = np.array([(T + i) & 0xFF for i in range(k)], np.uint8)
    cw = np.concatenate([O.encode(n, T, msg, 1), msg])
    pos = list(range(1, 2 * T, 2))
    rx = cw.copy()
    rx[pos] = 0
    assert O.recover(n, T, rx, 1, pos) == 1
    assert np.array_equal(rx, cw)


@pytest.mark.parametrize("n,T", [(8, 4), (16, 8), (32, 16), (64, 32), (64, 8), (256, 128),
                                 (256, 32)])
def test_recover_random_patterns(n, T):
    rng = np.random.default_rng(n + 17 * T)
    L = 9
    cw = make_codeword(n, T, L, rng)
    for e in sorted({0, 1, T // 2, T}):
        pos = rng.permutation(n)[:e]
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        rx = rx.reshape(-1)
        assert O.recover(n, T, rx, L, pos) == 1
        assert np.array_equal(rx, cw), e
        if O.gfni_supported():
            rx2 = cw.copy().reshape(n, L)
            rx2[pos] = 0
            rx2 = rx2.reshape(-1)
            O.recover(n, T, rx2, L, pos, gfni=True)
            assert np.array_equal(rx2, cw)


def test_recover_garbage_in_erased_shards_xors():
    # the reference writes the computed error value, so erased content g
    # comes back as c ^ g (linearity of rs/src/codec.rs:479-569).
    rng = np.random.default_rng(11)
    n, T, L = 32, 16, 13
    cw = make_codeword(n, T, L, rng)
    pos = rng.permutation(n)[:10]
    g = rng.integers(0, 256, (10, L), dtype=np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = g
    rx = rx.reshape(-1)
    assert O.recover(n, T, rx, L, pos) == 1
    exp = cw.copy().reshape(n, L)
    exp[pos] ^= g
    assert np.array_equal(rx.reshape(n, L), exp)


def test_recover_too_many_erasures_returns_false_untouched():
    rng = np.random.default_rng(12)
    n, T, L = 16, 4, 5
    cw = make_codeword(n, T, L, rng)
    rx = cw.copy()
    assert O.recover(n, T, rx, L, [0, 1, 2, 3, 4]) == 0
    assert np.array_equal(rx, cw)


def test_recompute_data_from_parity():
    # rs/src/poly_11d_lut.rs:471-494
    rng = np.random.default_rng(13)
    for n in (4, 8, 16, 32, 64, 128, 256):
        T, L = n // 2, 16
        cw = make_codeword(n, T, L, rng)
        rx = cw.copy()
        rx[T * L:] = 0
        O.recompute_data_from_parity(n, T, rx, L)
        assert np.array_equal(rx, cw)


def test_ext_recover_lagrange_roundtrip():
    rng = np.random.default_rng(14)
    for n, k in [(16, 4), (32, 8), (256, 16)]:
        L = 3
        T = n - k
        cw = make_codeword(n, T, L, rng)
        pos = rng.permutation(n)[:T]
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        rx = rx.reshape(-1)
        assert O.ext_recover_lagrange(n, k, rx, L, pos) == 1
        assert np.array_equal(rx, cw)
