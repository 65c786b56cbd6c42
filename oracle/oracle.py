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


def encode(n, T, message, L, parity=None, gfni=False):
    """Codec::encode_systematic_sharded (rs/src/codec.rs:241-265).

    `parity` is XOR-accumulated into (the reference's behaviour); a fresh
    zeroed buffer is used when it is None.  Returns the parity array."""
    if parity is None:
        parity = np.zeros(T * L, np.uint8)
    ws = np.zeros(T * L, np.uint8)
    f = lib().aro_gfni_encode_systematic_sharded if gfni else lib().aro_encode_systematic_sharded
    rc = f(n, T, _p(message), _p(parity), _p(ws), L)
    if rc != 0:
        raise ValueError(f"oracle encode rejected (n={n}, T={T})")
    return parity


def recover(n, T, received, L, positions, gfni=False):
    """Codec::recover_erasure_shards (rs/src/codec.rs:479-569); in place.
    Returns the reference's bool as 1/0 (negative for rejected input)."""
    ws = np.zeros(n * L, np.uint8)
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
