"""ctypes front-end for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The C sources
restate the reference crate's sharded codec (see the file headers of
ars_oracle.c / ars_oracle_gfni.c for the file:line citations).

Shard layout is the reference's: shard i of a codeword occupies bytes
[i*L, (i+1)*L); parity shards 0..T-1 come first, message shards T..n-1 follow
(rs/src/poly_11d_lut.rs:100-121).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.aro_mul.restype = ctypes.c_uint8
        L.aro_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
        L.aro_mul_lut.restype = ctypes.c_uint8
        L.aro_mul_lut.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
        L.aro_tables.restype = None
        L.aro_tables.argtypes = [_u8p] * 7 + [ctypes.POINTER(ctypes.c_uint64)]
        for name in ("aro_fft_sharded", "aro_ifft_sharded", "aro_gfni_fft_sharded",
                     "aro_gfni_ifft_sharded"):
            f = getattr(L, name)
            f.restype = None
            f.argtypes = [_u8p, _sz, ctypes.c_int, ctypes.c_uint8]
        for name in ("aro_encode_systematic_sharded", "aro_gfni_encode_systematic_sharded"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_sz, _sz, _u8p, _u8p, _u8p, _sz]
        for name in ("aro_recover_erasure_shards", "aro_gfni_recover_erasure_shards"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_sz, _sz, _u8p, _u8p, _sz, _u8p, _sz]
        L.aro_recompute_data_from_parity_sharded.restype = ctypes.c_int
        L.aro_recompute_data_from_parity_sharded.argtypes = [_sz, _sz, _u8p, _u8p, _sz]
        L.aro_ext_encode_kpow2.restype = ctypes.c_int
        L.aro_ext_encode_kpow2.argtypes = [_sz, _sz, _u8p, _u8p, _sz]
        L.aro_ext_recover_lagrange.restype = ctypes.c_int
        L.aro_ext_recover_lagrange.argtypes = [_sz, _sz, _u8p, _sz, _u8p, _sz]
        L.aro_gfni_supported.restype = ctypes.c_int
        L.aro_gfni_supported.argtypes = []
        L.aro_gfni_time_codec.restype = ctypes.c_int
        L.aro_gfni_time_codec.argtypes = [_sz, _sz, _sz, _u8p, _sz, ctypes.c_int, _sz,
                                          ctypes.POINTER(ctypes.c_double)]
        L.aro_gfni_set_unrolled.restype = None
        L.aro_gfni_set_unrolled.argtypes = [ctypes.c_int]
        L.aro_decode_systematic_scalar.restype = ctypes.c_int
        L.aro_decode_systematic_scalar.argtypes = [_sz, _sz, _u8p, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.aro_decode_sharded.restype = ctypes.c_int
        L.aro_decode_sharded.argtypes = [_sz, _sz, _u8p, _sz, _u8p, ctypes.c_int]
        L.aro_formal_deriv.restype = None
        L.aro_formal_deriv.argtypes = [_u8p, _sz, _u8p]
        L.aro_poly_mul_lnh.restype = ctypes.c_int
        L.aro_poly_mul_lnh.argtypes = [_u8p, _u8p, _sz, _u8p]
        L.aro_poly_div_lnh.restype = ctypes.c_int
        L.aro_poly_div_lnh.argtypes = [_u8p, _u8p, _sz, _u8p, _u8p]
        L.aro_deriv_lnh.restype = None
        L.aro_deriv_lnh.argtypes = [_u8p, _sz, _u8p]
        L.aro_syndrome_scalar.restype = None
        L.aro_syndrome_scalar.argtypes = [_u8p, _sz, _sz, _u8p]
        L.aro_key_equation.restype = ctypes.c_int
        L.aro_key_equation.argtypes = [_u8p, _sz, _sz, ctypes.c_int, _u8p, _u8p]
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u8p)


def tables():
    """Return the generated tables as numpy arrays (rs/build.rs:549-601)."""
    exp = np.zeros(510, np.uint8)
    log = np.zeros(256, np.uint8)
    inv = np.zeros(256, np.uint8)
    basis = np.zeros(8, np.uint8)
    sub = np.zeros(256, np.uint8)
    spv = np.zeros((9, 256), np.uint8)
    coeff = np.zeros(9, np.uint8)
    gfni = np.zeros(256, np.uint64)
    lib().aro_tables(_p(exp), _p(log), _p(inv), _p(basis), _p(sub), _p(spv), _p(coeff),
                     gfni.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return dict(exp=exp, log=log, inv=inv, basis=basis, subspace=sub, spv=spv,
                coeffs=coeff, gfni=gfni)


def mul(a, b):
    return lib().aro_mul(a, b)


def mul_lut(a, b):
    return lib().aro_mul_lut(a, b)


def gfni_supported():
    return bool(lib().aro_gfni_supported())


def fft_sharded(shards, L, k, beta, gfni=False):
    f = lib().aro_gfni_fft_sharded if gfni else lib().aro_fft_sharded
    f(_p(shards), L, k, beta)


def ifft_sharded(shards, L, k, beta, gfni=False):
    f = lib().aro_gfni_ifft_sharded if gfni else lib().aro_ifft_sharded
    f(_p(shards), L, k, beta)


def encode(n, T, message, L, parity=None, gfni=False, workspace=None):
    """Codec::encode_systematic_sharded (rs/src/codec.rs:241-265).

    `parity` is XOR-accumulated into (the reference's behaviour); a fresh
    zeroed buffer is used when it is None.  `workspace` (T*L bytes, contents
    irrelevant) may be passed to reuse one, as the reference's callers do.
    Returns the parity array."""
    if parity is None:
        parity = np.zeros(T * L, np.uint8)
    ws = np.zeros(T * L, np.uint8) if workspace is None else workspace
    assert ws.size >= T * L
    f = lib().aro_gfni_encode_systematic_sharded if gfni else lib().aro_encode_systematic_sharded
    rc = f(n, T, _p(message), _p(parity), _p(ws), L)
    if rc != 0:
        raise ValueError(f"oracle encode rejected (n={n}, T={T})")
    return parity


def recover(n, T, received, L, positions, gfni=False, workspace=None):
    """Codec::recover_erasure_shards (rs/src/codec.rs:479-569); in place.
    Returns the reference's bool as 1/0 (negative for rejected input).
    `workspace` (n*L bytes) may be passed to reuse one."""
    ws = np.zeros(n * L, np.uint8) if workspace is None else workspace
    assert ws.size >= n * L
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint8))
    f = lib().aro_gfni_recover_erasure_shards if gfni else lib().aro_recover_erasure_shards
    return f(n, T, _p(received), _p(ws), L, _p(pos) if len(pos) else None, len(pos))


def recompute_data_from_parity(n, T, received, L):
    """Codec::recompute_data_from_parity_sharded (rs/src/codec.rs:321-340)."""
    ws = np.zeros(T * L, np.uint8)
    rc = lib().aro_recompute_data_from_parity_sharded(n, T, _p(received), _p(ws), L)
    if rc != 0:
        raise ValueError("oracle recompute rejected")


def ext_encode_kpow2(n, k, message, L):
    parity = np.zeros((n - k) * L, np.uint8)
    rc = lib().aro_ext_encode_kpow2(n, k, _p(message), _p(parity), L)
    if rc != 0:
        raise ValueError("oracle ext encode rejected")
    return parity


def ext_recover_lagrange(n, k, received, L, positions):
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint8))
    return lib().aro_ext_recover_lagrange(n, k, _p(received), L, _p(pos) if len(pos) else None,
                                          len(pos))


def encode_any(n, k, message, L):
    """Parity for any supported (n, k): the reference algorithm when T is a
    power of two, the k-block extension otherwise."""
    T = n - k
    if T & (T - 1) == 0:
        return encode(n, T, message, L)
    return ext_encode_kpow2(n, k, message, L)


def recover_any(n, k, received, L, positions):
    T = n - k
    if T & (T - 1) == 0:
        return recover(n, T, received, L, positions)
    return ext_recover_lagrange(n, k, received, L, positions)


# --- error decoding (rs/src/codec.rs:28-301, 341-420; rs/src/poly_arith.rs) ---

PANIC = -3  # the reference would panic at this input


def _poly(a, n):
    out = np.zeros(n, np.uint8)
    a = np.asarray(a, dtype=np.uint8)
    out[:len(a)] = a
    return out


def poly_mul_lnh(a, b, n=256):
    """CantorBasisPolyArith::poly_mul_lnh (rs/src/poly_arith.rs:198-231) on
    [G; n] arrays.  Returns the product, or None where the reference panics."""
    a, b, out = _poly(a, n), _poly(b, n), np.zeros(n, np.uint8)
    rc = lib().aro_poly_mul_lnh(_p(a), _p(b), n, _p(out))
    return None if rc else out


def poly_div_lnh(a, b, n=256):
    """poly_div_lnh (rs/src/poly_arith.rs:233-267): (q, r), or None for b = 0."""
    a, b = _poly(a, n), _poly(b, n)
    q, r = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    rc = lib().aro_poly_div_lnh(_p(a), _p(b), n, _p(q), _p(r))
    if rc == 1:
        return q, r
    if rc == 0:
        return None
    raise ValueError("reference would panic")


def deriv_lnh(c, n=256):
    """poly::deriv_lnh (rs/src/poly_arith.rs:91-116) over n coefficients."""
    c, out = _poly(c, n), np.zeros(n, np.uint8)
    lib().aro_deriv_lnh(_p(c), n, _p(out))
    return out


def syndrome(n, T, received):
    """compute_syndrome_scalar (rs/src/codec.rs:266-301): n entries, first T set."""
    r = np.ascontiguousarray(np.asarray(received, dtype=np.uint8))
    out = np.zeros(n, np.uint8)
    lib().aro_syndrome_scalar(_p(r), n, T, _p(out))
    return out


def key_equation(syn, n, T, eea=False):
    """solve_key_equation_hgcd (rs/src/codec.rs:212-239) or, with eea=True,
    solve_key_equation_eea (rs/src/codec.rs:28-80): (v1, lambda) or None;
    PANIC where the reference would panic."""
    s = _poly(syn, n)
    v1, lam = np.zeros(n, np.uint8), np.zeros(n, np.uint8)
    rc = lib().aro_key_equation(_p(s), n, T, 1 if eea else 0, _p(v1), _p(lam))
    if rc == 1:
        return v1, lam
    if rc == 0:
        return None
    return PANIC


def formal_deriv(c, n=256):
    """The formal derivative in the basis X (extension; not the reference's
    deriv_lnh): D'[i] = xor over bits b not in i of D[i | 2^b]."""
    c, out = _poly(c, n), np.zeros(n, np.uint8)
    lib().aro_formal_deriv(_p(c), n, _p(out))
    return out


def decode_scalar(n, T, received, formal=False):
    """Codec::decode_systematic_scalar (rs/src/codec.rs:341-420) on a copy.
    Returns (rc, kind, word): rc 1/0 as the reference's bool (PANIC where it
    would panic); kind 0 zero syndrome, 1 true after the key equation,
    2 false, 3 panic.  formal=True evaluates errors with the formal
    derivative instead of the reference's deriv_lnh (extension)."""
    w = np.array(received, dtype=np.uint8)
    kind = ctypes.c_int(0)
    rc = lib().aro_decode_systematic_scalar(n, T, _p(w), ctypes.byref(kind), 1 if formal else 0)
    return rc, kind.value, w


def decode_sharded(n, T, received, L, formal=False):
    """Each byte column of the n x L codeword decoded as decode_systematic_scalar
    decodes one received word; in place.  Returns (rc, status per column)."""
    status = np.zeros(L, np.uint8)
    rc = lib().aro_decode_sharded(n, T, _p(received), L, _p(status), 1 if formal else 0)
    return rc, status


def gfni_time_codec(n, T, L, positions, reps, offset=0):
    """Median ns of one GFNI encode and one recover, timed inside C on
    buffers allocated once, starting `offset` bytes past a 64-byte boundary
    (CPU baseline only).  None without GFNI."""
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.uint8))
    out = (ctypes.c_double * 2)()
    rc = lib().aro_gfni_time_codec(n, T, L, _p(pos), len(pos), reps, offset, out)
    return None if rc else (out[0], out[1])


def gfni_set_unrolled(on):
    """beta = 0 GFNI transforms as the reference's generated fixed sequences
    (True, its default dispatch) or through the generic recursion (False)."""
    lib().aro_gfni_set_unrolled(1 if on else 0)
