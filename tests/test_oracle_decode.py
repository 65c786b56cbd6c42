"""Pins the oracle's restatement of the reference's scalar error decoder
(rs/src/codec.rs:28-301, 341-420) and of the novel-basis polynomial
arithmetic it uses (rs/src/poly_arith.rs), against:
  * the reference's own tests (rs/src/poly_11d_lut.rs:157-233, 260-404);
  * closed forms evaluated with the textbook field in gfpy (products and
    quotients evaluated at points, the product rule at known roots, the
    locator's roots at the error positions, unique decoding within T/2);
  * the reference's claim that its EEA and HGCD key-equation solvers are
    functionally equivalent (rs/src/codec.rs:212-214)."""
import numpy as np
import pytest

import gfpy
import oracle as O

TAB = O.tables()
SUB = [int(v) for v in TAB["subspace"]]
BASIS = [int(v) for v in TAB["basis"]]
REF_CODES = [(4, 2), (8, 4), (16, 8), (32, 16), (64, 32), (128, 64), (256, 128)]


def xeval_table(points, count):
    """X_j(x) for j < count and x in points (textbook, gfpy)."""
    return {x: [gfpy.lch_basis(BASIS, j, x) for j in range(count)] for x in points}


def peval(c, x, tab):
    acc = 0
    for j, v in enumerate(c[:len(tab[x])]):
        if v:
            acc ^= gfpy.gmul(int(v), tab[x][j])
    return acc


def poly(coeffs, n=256):
    p = np.zeros(n, np.uint8)
    for i, v in coeffs.items():
        p[i] = v
    return p


# --- rs/src/poly_11d_lut.rs:260-388, the reference's poly_mul_lnh tests -----

def test_poly_mul_lnh_reference_tests():
    zero = np.zeros(256, np.uint8)
    a = poly({0: 0x53, 1: 0xca, 3: 1})
    assert np.array_equal(O.poly_mul_lnh(a, zero), zero)
    assert np.array_equal(O.poly_mul_lnh(zero, a), zero)
    one = poly({0: 1})
    a = poly({0: 0xab, 2: 0x3f, 7: 0x11})
    assert np.array_equal(O.poly_mul_lnh(a, one), a)
    assert np.array_equal(O.poly_mul_lnh(one, a), a)
    r = O.poly_mul_lnh(poly({0: 0x53}), poly({0: 0xca}))
    assert r[0] == O.mul_lut(0x53, 0xca) and not r[1:].any()
    a, b = poly({0: 1, 1: 0xff}), poly({0: 0x57, 2: 0x83, 5: 0x2a})
    assert np.array_equal(O.poly_mul_lnh(a, b), O.poly_mul_lnh(b, a))
    a, b, c = poly({0: 3, 1: 9}), poly({0: 0xe5, 3: 0x7c}), poly({0: 0xaa, 1: 0x55})
    assert np.array_equal(O.poly_mul_lnh(O.poly_mul_lnh(a, b), c), O.poly_mul_lnh(a, O.poly_mul_lnh(b, c)))
    a, b, c = poly({0: 0x11, 2: 0x22}), poly({1: 0x33, 3: 0x44}), poly({0: 0x55, 2: 0x66})
    assert np.array_equal(O.poly_mul_lnh(a, b ^ c), O.poly_mul_lnh(a, b) ^ O.poly_mul_lnh(a, c))
    r = O.poly_mul_lnh(poly({3: 1}), poly({5: 1}))
    assert max(np.nonzero(r)[0]) == 8


def test_poly_mul_lnh_is_the_polynomial_product():
    rng = np.random.default_rng(11)
    pts = [int(x) for x in rng.integers(1, 256, 5)]
    tab = xeval_table(pts, 64)
    for _ in range(12):
        da, db = int(rng.integers(0, 32)), int(rng.integers(0, 32))
        a, b = np.zeros(256, np.uint8), np.zeros(256, np.uint8)
        a[:da + 1] = rng.integers(0, 256, da + 1)
        b[:db + 1] = rng.integers(0, 256, db + 1)
        p = O.poly_mul_lnh(a, b)
        assert not p[64:].any()
        for x in pts:
            assert peval(p, x, tab) == gfpy.gmul(peval(a, x, tab), peval(b, x, tab))


def test_poly_mul_lnh_panics_where_the_reference_does():
    # deg a + deg b + 1 > N: the reference slices va[..n] beyond its [G; N]
    assert O.poly_mul_lnh(poly({20: 1}, 32), poly({20: 1}, 32), 32) is None


# --- rs/src/poly_11d_lut.rs:390-404 and the division identity --------------

def test_poly_div_lnh_by_constant_scales_coefficients():
    a = poly({0: 0xef, 3: 0xcd, 7: 0xab})
    q, r = O.poly_div_lnh(a, poly({0: 0x53}))
    assert not r.any()
    cinv = int(TAB["inv"][0x53])
    assert list(q) == [O.mul_lut(int(v), cinv) for v in a]
    assert O.poly_div_lnh(a, np.zeros(256, np.uint8)) is None


def test_poly_div_lnh_is_division_with_remainder():
    rng = np.random.default_rng(12)
    pts = [int(x) for x in rng.integers(1, 256, 4)]
    tab = xeval_table(pts, 64)
    for _ in range(10):
        da, db = int(rng.integers(0, 64)), int(rng.integers(0, 40))
        a, b = np.zeros(256, np.uint8), np.zeros(256, np.uint8)
        a[:da + 1] = rng.integers(0, 256, da + 1)
        b[:db] = rng.integers(0, 256, db)
        b[db] = int(rng.integers(1, 256))
        q, r = O.poly_div_lnh(a, b)
        rd = max(list(np.nonzero(r)[0]) + [-1])
        assert rd < db
        for x in pts:
            assert peval(a, x, tab) == gfpy.gmul(peval(q, x, tab), peval(b, x, tab)) ^ peval(r, x, tab)


# --- derivatives ---------------------------------------------------------------

def test_deriv_lnh_closed_form():
    """deriv_lnh is the superset sum res[i] = sum_{s >= i} c[s], which
    evaluates to sum_s c_s prod_{b in s} (s_b(x) + 1)."""
    rng = np.random.default_rng(13)
    n = 32
    pts = [int(x) for x in rng.integers(0, 256, 6)]
    tab = xeval_table(pts, n)
    for _ in range(4):
        c = rng.integers(0, 256, n).astype(np.uint8)
        d = O.deriv_lnh(c, n)
        for x in pts:
            want = 0
            for s in range(n):
                t = int(c[s])
                for b in range(5):
                    if (s >> b) & 1:
                        t = gfpy.gmul(t, gfpy.subspace_poly(BASIS, b, x) ^ 1)
                want ^= t
            assert peval(d, x, tab) == want


@pytest.mark.parametrize("n", [8, 32, 256])
def test_formal_deriv_product_rule_at_roots(n):
    """p = prod_{i in S} (x + w_i): p'(w_i) = prod_{j in S, j != i} (w_i + w_j)."""
    rng = np.random.default_rng(14 + n)
    k = int(np.log2(n))
    S = sorted(int(v) for v in rng.permutation(n)[:max(1, n // 3)])
    ev = np.zeros(n, np.uint8)
    for j in range(n):
        v = 1
        for i in S:
            v = gfpy.gmul(v, SUB[j] ^ SUB[i])
        ev[j] = v
    coef = ev.copy()
    O.ifft_sharded(coef, 1, k, 0)
    d = O.formal_deriv(coef, n)
    O.fft_sharded(d, 1, k, 0)
    for i in S:
        want = 1
        for j in S:
            if j != i:
                want = gfpy.gmul(want, SUB[i] ^ SUB[j])
        assert int(d[i]) == want


# --- key equation -------------------------------------------------------------

@pytest.mark.parametrize("n,T", [(8, 4), (32, 16), (64, 32), (256, 128), (256, 32), (256, 2), (128, 4)])
def test_key_equation_hgcd_equals_eea(n, T):
    rng = np.random.default_rng(n + 7 * T)
    for _ in range(60):
        syn = np.zeros(n, np.uint8)
        d = int(rng.integers(0, T))
        syn[:d + 1] = rng.integers(0, 256, d + 1)
        if not syn.any():
            continue
        a, b = O.key_equation(syn, n, T), O.key_equation(syn, n, T, eea=True)
        assert isinstance(a, tuple) and isinstance(b, tuple)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def codeword(n, T, rng):
    msg = rng.integers(0, 256, n - T, dtype=np.uint8)
    return np.concatenate([O.encode(n, T, msg, 1), msg])


@pytest.mark.parametrize("n,T", [(16, 8), (64, 32), (256, 128), (256, 16)])
def test_locator_roots_are_the_error_positions(n, T):
    rng = np.random.default_rng(15 * n + T)
    k = int(np.log2(T))
    for _ in range(10):
        c = codeword(n, T, rng)
        ne = int(rng.integers(1, T // 2 + 1))
        pos = set(int(v) for v in rng.permutation(n)[:ne])
        r = c.copy()
        r[sorted(pos)] ^= rng.integers(1, 256, ne, dtype=np.uint8)
        v1, lam = O.key_equation(O.syndrome(n, T, r), n, T)
        roots = set()
        for ch in range(n // T):
            ev = lam.copy()
            O.fft_sharded(ev, 1, k, SUB[ch * T])
            roots |= {ch * T + o for o in range(T) if ev[o] == 0}
        assert roots == pos
        assert max(np.nonzero(lam)[0]) == ne


# --- decode: the reference's tests (rs/src/poly_11d_lut.rs:157-233) ------------

def lch_codeword(n, T):
    msg = np.array([(T + i) & 0xFF for i in range(n - T)], np.uint8)
    return np.concatenate([O.encode(n, T, msg, 1), msg])


@pytest.mark.parametrize("n,T", REF_CODES)
def test_decode_reference_tests(n, T):
    c = lch_codeword(n, T)
    rc, kind, w = O.decode_scalar(n, T, c)
    assert rc == 1 and kind == 0 and np.array_equal(w, c)
    for base in (0, T):  # corrupt parity, then corrupt data
        r = c.copy()
        for i in range(T // 2):
            r[base + i] = O.mul(int(r[base + i]), i + 2)
        rc, kind, w = O.decode_scalar(n, T, r)
        assert rc == 1 and np.array_equal(w, c), (n, T, base)


@pytest.mark.parametrize("n,T", [(4, 2), (8, 4), (16, 8), (32, 16), (64, 32), (256, 128), (256, 32), (128, 8)])
def test_decode_formal_is_unique_decoding(n, T):
    """With the formal derivative every pattern of <= T/2 errors decodes to
    the transmitted codeword (bounded-distance uniqueness)."""
    rng = np.random.default_rng(16 * n + T)
    for _ in range(25):
        c = codeword(n, T, rng)
        ne = int(rng.integers(1, T // 2 + 1))
        pos = rng.permutation(n)[:ne]
        r = c.copy()
        r[pos] ^= rng.integers(1, 256, ne, dtype=np.uint8)
        rc, kind, w = O.decode_scalar(n, T, r, formal=True)
        assert rc == 1 and kind == 1 and np.array_equal(w, c)


@pytest.mark.parametrize("n,T", [(16, 8), (32, 16)])
def test_decode_reference_error_values_closed_form(n, T):
    """Reference mode (deriv_lnh in step 4): every corrected value v_p obeys
    v_p * D(w_p) = e_p * lambda'(w_p), where e_p is the true error, D is the
    closed form of deriv_lnh and lambda' the formal derivative; it changes
    nothing outside the error positions."""
    rng = np.random.default_rng(17 * n + T)
    pts = SUB[:n]
    tab = xeval_table(pts, n)
    for _ in range(6):
        c = codeword(n, T, rng)
        ne = int(rng.integers(1, T // 2 + 1))
        pos = rng.permutation(n)[:ne]
        r = c.copy()
        r[pos] ^= rng.integers(1, 256, ne, dtype=np.uint8)
        rc, _, w = O.decode_scalar(n, T, r)
        assert rc == 1
        assert set(np.nonzero(w != r)[0]) <= set(int(p) for p in pos)
        _, lam = O.key_equation(O.syndrome(n, T, r), n, T)
        dref, dfor = O.deriv_lnh(lam, n), O.formal_deriv(lam, n)
        for p in pos:
            x = SUB[int(p)]
            v, e = int(w[p] ^ r[p]), int(c[p] ^ r[p])
            a, b = peval(dref, x, tab), peval(dfor, x, tab)
            if a == 0:
                assert v == 0  # the reference skips a zero denominator
            else:
                assert gfpy.gmul(v, a) == gfpy.gmul(e, b)


def test_decode_beyond_capacity_false_leaves_word_untouched():
    rng = np.random.default_rng(18)
    n, T = 64, 8
    seen_false = 0
    for _ in range(60):
        c = codeword(n, T, rng)
        pos = rng.permutation(n)[:T + 3]
        r = c.copy()
        r[pos] ^= rng.integers(1, 256, len(pos), dtype=np.uint8)
        rc, kind, w = O.decode_scalar(n, T, r)
        if rc == 0:
            seen_false += 1
            assert kind == 2 and np.array_equal(w, r)
    assert seen_false > 0


def test_decode_sharded_equals_columnwise():
    rng = np.random.default_rng(19)
    n, T, L = 32, 8, 23
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    cw = np.concatenate([O.encode(n, T, msg, L), msg]).reshape(n, L)
    rx = cw.copy()
    for j in range(L):
        ne = int(rng.integers(0, T + 2))
        rx[rng.permutation(n)[:ne], j] ^= rng.integers(1, 256, ne, dtype=np.uint8)
    for formal in (False, True):
        got = rx.copy().reshape(-1)
        rc, st = O.decode_sharded(n, T, got, L, formal=formal)
        got = got.reshape(n, L)
        for j in range(L):
            rcj, kind, w = O.decode_scalar(n, T, rx[:, j], formal=formal)
            assert st[j] == kind
            assert np.array_equal(got[:, j], w if rcj == 1 else rx[:, j])
