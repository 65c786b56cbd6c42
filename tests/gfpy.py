"""Independent textbook GF(2^8) arithmetic for pinning the oracle.

Written from the mathematics (GF(2)[x]/(x^8+x^4+x^3+x^2+1), generator 2),
not from the oracle's C code, so that a mistake in one is caught by the other.
"""
import numpy as np

POLY = 0x11D


def gmul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        b >>= 1
        a <<= 1
        if a & 0x100:
            a ^= POLY
    return r


def gpow(a, e):
    r = 1
    for _ in range(e):
        r = gmul(r, a)
    return r


# full multiplication table (65536 entries) for vectorised use
MUL = np.array([[gmul(a, b) for b in range(256)] for a in range(256)], dtype=np.uint8)
INV = np.zeros(256, np.uint8)
for _a in range(1, 256):
    INV[_a] = int(np.nonzero(MUL[_a] == 1)[0][0])


def vmul(scalar, arr):
    """scalar * arr elementwise (arr uint8 array)."""
    return MUL[scalar][arr]


def trace(x):
    t, s = x, x
    for _ in range(7):
        t = gmul(t, t)
        s ^= t
    return s & 1


def span(basis, k):
    out = []
    for i in range(1 << k):
        v = 0
        for b in range(k):
            if (i >> b) & 1:
                v ^= basis[b]
        out.append(v)
    return out


def subspace_poly(basis, j, x):
    """s_j(x) = prod over a in V_j of (x + a), by brute force."""
    r = 1
    for a in span(basis, j):
        r = gmul(r, x ^ a)
    return r


def lch_basis(basis, i, x):
    """X_i(x) = prod over set bits b of i of s_b(x) / s_b(v_b) (LCH basis)."""
    r = 1
    for b in range(8):
        if (i >> b) & 1:
            num = subspace_poly(basis, b, x)
            den = subspace_poly(basis, b, basis[b])
            r = gmul(r, gmul(num, int(INV[den])))
    return r


def lagrange_eval_matrix(xs, targets):
    """Matrix M (len(targets) x len(xs)) with f(t) = sum_j M[t, j] f(x_j) for
    every polynomial of degree < len(xs)."""
    k = len(xs)
    M = np.zeros((len(targets), k), np.uint8)
    for ti, t in enumerate(targets):
        for j in range(k):
            num, den = 1, 1
            for l in range(k):
                if l != j:
                    num = gmul(num, t ^ xs[l])
                    den = gmul(den, xs[j] ^ xs[l])
            M[ti, j] = gmul(num, int(INV[den]))
    return M


def apply_matrix(M, rows):
    """rows: (k, L) uint8; returns (m, L) = M @ rows over GF(2^8)."""
    out = np.zeros((M.shape[0], rows.shape[1]), np.uint8)
    for i in range(M.shape[0]):
        acc = np.zeros(rows.shape[1], np.uint8)
        for j in range(M.shape[1]):
            if M[i, j]:
                acc ^= MUL[M[i, j]][rows[j]]
        out[i] = acc
    return out
