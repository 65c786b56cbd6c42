"""The phase-B probes of libars_probe.so (bench infrastructure: the bit-sliced
prototype and the product's v_perm form of the same map) compute phase B of
the 256-point recover: on 16 rows it is the 16-point transform pair at
beta = 0 with the derivative between, since s_{l}(omega_{16h}) =
s_{l-4}(omega_h) for the Cantor basis.  Both must equal the oracle's
IFFT_16 / derivative / FFT_16 applied the probe's number of times."""
import ctypes
import os

import numpy as np
import pytest

import _pkg
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def probe():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _pkg.load().lib()
    path = os.path.join(_pkg.PKG_DIR, "libars_probe.so")
    assert os.path.exists(path), "libars_probe.so is not built"
    lib = ctypes.CDLL(path)
    lib.ars_probe_phaseb.restype = ctypes.c_int
    lib.ars_probe_phaseb.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double)]
    return lib


def derivative16(rows):
    # novel-basis formal derivative on 16 points: d[h] = xor over bits b of h
    # that are clear of x[h | 2^b] (every s_b has x-coefficient 1)
    out = np.zeros_like(rows)
    for h in range(16):
        for b in range(4):
            if not (h >> b) & 1:
                out[h] ^= rows[h | (1 << b)]
    return out


@pytest.mark.parametrize("L", [32, 4096, 65536 + 96])
def test_phaseb_forms_match_oracle(probe, L):
    maps = probe.ars_probe_phaseb_maps()
    rng = np.random.default_rng(L)
    src = rng.integers(0, 256, 16 * L, dtype=np.uint8)
    exp = src.copy()
    for _ in range(maps):
        O.ifft_sharded(exp, L, 4, 0)
        exp = np.ascontiguousarray(derivative16(exp.reshape(16, L)).reshape(-1))
        O.fft_sharded(exp, L, 4, 0)
    d_src = torch.from_numpy(src).cuda()
    for variant in (0, 1):
        out = torch.zeros_like(d_src)
        us = ctypes.c_double()
        print(f"phase-B probe variant {variant}, L={L}", flush=True)
        assert probe.ars_probe_phaseb(variant, d_src.data_ptr(), out.data_ptr(), L, 1, ctypes.byref(us)) == 0
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), exp), f"variant {variant}"


def test_phaseb_rejects_bad_sizes(probe):
    d = torch.zeros(16 * 64, dtype=torch.uint8, device="cuda")
    us = ctypes.c_double()
    assert probe.ars_probe_phaseb(0, d.data_ptr(), d.data_ptr(), 48, 1, ctypes.byref(us)) == -1
    assert probe.ars_probe_phaseb(2, d.data_ptr(), d.data_ptr(), 64, 1, ctypes.byref(us)) == -1
