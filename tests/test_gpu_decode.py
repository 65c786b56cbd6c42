"""GPU parity of sharded error decoding (ars_decode, through the C ABI)
against the oracle's restatement of Codec::decode_systematic_scalar
(rs/src/codec.rs:341-420) applied to every byte column: corrected bytes,
per-column status and return code, byte for byte, in the reference's mode and
with the formal derivative."""
import numpy as np
import pytest

import _pkg
import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_rx(n, T, L, rng, max_err=None):
    """Codeword columns with, per column, 0..max_err random errors."""
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    cw = np.concatenate([O.encode(n, T, msg, L), msg]).reshape(n, L)
    rx = cw.copy()
    hi = T + 2 if max_err is None else max_err
    for j in range(L):
        ne = int(rng.integers(0, hi + 1))
        pos = rng.permutation(n)[:min(ne, n)]
        rx[pos, j] ^= rng.integers(1, 256, len(pos), dtype=np.uint8)
    return cw, rx


def gpu_decode(ars, n, k, rx, L, formal, host_keq=False):
    c = ars.Codec(n, k)
    d = dev(rx.reshape(-1))
    st = torch.full((L,), 7, dtype=torch.uint8, device="cuda")
    ok, counts = c.decode_systematic_sharded(d, L, formal_derivative=formal, status=st,
                                             host_key_equation=host_keq)
    return ok, counts, host(d).reshape(n, L), host(st)


def check_against_oracle(ars, n, T, rx, L, formal, host_keq=False):
    got_ok, counts, got, st = gpu_decode(ars, n, n - T, rx, L, formal, host_keq)
    exp = rx.reshape(-1).copy()
    rc, est = O.decode_sharded(n, T, exp, L, formal=formal)
    assert rc in (0, 1), "the reference would panic on this input"
    assert np.array_equal(st, est), (n, T, L, formal)
    assert np.array_equal(got, exp.reshape(n, L)), (n, T, L, formal)
    assert got_ok == (rc == 1)
    assert counts == (int((est == 0).sum()), int((est == 1).sum()), int((est == 2).sum()))


CODES = [(n, T) for n in (4, 8, 16, 32, 64, 128, 256) for T in (2, 4, 8, 16, 32, 64, 128) if T < n]


@pytest.mark.parametrize("n,T", CODES)
@pytest.mark.parametrize("L", [1, 37, 64])
def test_decode_matches_oracle(ars, n, T, L):
    rng = np.random.default_rng(n * 1009 + T * 17 + L)
    _, rx = make_rx(n, T, L, rng)
    for formal in (False, True):
        for host_keq in (False, True):  # key equations on the GPU, then on the host
            check_against_oracle(ars, n, T, rx, L, formal, host_keq)


@pytest.mark.parametrize("n,T", [(16, 8), (64, 32), (256, 128), (256, 16), (128, 4)])
def test_decode_formal_corrects_up_to_half_T(ars, n, T):
    rng = np.random.default_rng(3 * n + T)
    L = 256
    cw, rx = make_rx(n, T, L, rng, max_err=T // 2)
    ok, counts, got, st = gpu_decode(ars, n, n - T, rx, L, True)
    assert ok and counts[2] == 0
    assert np.array_equal(got, cw)


@pytest.mark.parametrize("n,T", [(4, 2), (32, 16), (256, 128)])
def test_decode_reference_test_patterns_sharded(ars, n, T):
    """rs/src/poly_11d_lut.rs:179-233 on shards: column 0 is the reference's
    test codeword (message[i] = T + i) and comes back exactly in the
    reference's mode; the other columns are other codewords with the same
    corruption, where the reference's error values can be wrong, so they are
    compared with the oracle.  The formal derivative restores every column."""
    L = 64
    k = n - T
    msg = np.array([(T + i + j) & 0xFF for i in range(k) for j in range(L)], np.uint8)
    cw = np.concatenate([O.encode(n, T, msg, L), msg]).reshape(n, L)
    for base in (0, T):
        rx = cw.copy()
        for i in range(T // 2):
            rx[base + i] = [O.mul(int(v), i + 2) for v in rx[base + i]]
        ok, counts, got, _ = gpu_decode(ars, n, k, rx, L, False)
        assert ok and np.array_equal(got[:, 0], cw[:, 0])
        check_against_oracle(ars, n, T, rx, L, False)
        ok, counts, got, _ = gpu_decode(ars, n, k, rx, L, True)
        assert ok and np.array_equal(got, cw)


def test_decode_whole_shard_corruption(ars):
    """Corrupted shards: every column has errors at the same positions."""
    rng = np.random.default_rng(5)
    n, T, L = 64, 16, 4096
    cw, _ = make_rx(n, T, L, rng, max_err=0)
    rx = cw.copy()
    bad = rng.permutation(n)[:5]
    rx[bad] ^= rng.integers(1, 256, (5, L), dtype=np.uint8)
    for host_keq in (False, True):
        ok, counts, got, st = gpu_decode(ars, n, n - T, rx, L, True, host_keq)
        assert ok and counts == (0, L, 0) and np.array_equal(got, cw)
        check_against_oracle(ars, n, T, rx[:, :512].copy(), 512, False, host_keq)


def test_decode_unaligned_and_clean(ars):
    rng = np.random.default_rng(6)
    n, T, L = 32, 8, 101
    cw, rx = make_rx(n, T, L, rng)
    for formal in (False, True):
        check_against_oracle(ars, n, T, rx, L, formal)
    # an unaligned base pointer
    buf = torch.zeros(n * L + 1, dtype=torch.uint8, device="cuda")
    buf[1:] = dev(rx.reshape(-1))
    view = buf[1:]
    ok, counts = ars.Codec(n, n - T).decode_systematic_sharded(view, L, formal_derivative=True)
    exp = rx.reshape(-1).copy()
    O.decode_sharded(n, T, exp, L, formal=True)
    assert np.array_equal(host(view), exp)
    # a clean codeword is left exactly as it is
    ok, counts, got, st = gpu_decode(ars, n, n - T, cw, L, False)
    assert ok and counts == (L, 0, 0) and not st.any() and np.array_equal(got, cw)


def test_decode_rejects_unsupported_codes(ars):
    d = torch.zeros(8 * 16, dtype=torch.uint8, device="cuda")
    ctx = ars.Codec(8, 4).ctx
    for n, k in [(8, 7), (8, 5), (256, 16)]:  # T = 1, 3, 240
        with pytest.raises(ars.ArsError) as ei:
            ctx.decode(n, k, d.data_ptr(), 16)
        assert ei.value.code == ars.ARS_E_UNSUPPORTED
    with pytest.raises(ars.ArsError) as ei:
        ctx.decode(8, 4, d.data_ptr(), 16, flags=8)
    assert ei.value.code == ars.ARS_E_INVAL


def test_decode_full_size_scrub(ars):
    """(256,128) with 1 MiB shards: a clean scrub leaves every byte as it
    was; then 2000 corrupted columns (<= 64 errors each) come back exactly
    with the formal derivative, and in the reference's mode equal the oracle
    on those columns."""
    n, k, L = 256, 128, 1 << 20
    T = n - k
    g = torch.Generator(device="cuda")
    g.manual_seed(43)
    msg = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda", generator=g)
    par = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
    c = ars.Codec(n, k)
    c.encode_systematic_sharded(msg, par, None, L, accumulate=False)
    cw = torch.cat([par, msg])
    rx = cw.clone()
    ok, counts = c.decode_systematic_sharded(rx, L)
    assert ok and counts == (L, 0, 0) and torch.equal(rx, cw)
    rng = np.random.default_rng(7)
    cols = np.sort(rng.permutation(L)[:2000])
    host_rx = rx.view(n, L)[:, torch.from_numpy(cols).cuda()].cpu().numpy()
    for j in range(len(cols)):
        ne = int(rng.integers(1, T // 2 + 1))
        host_rx[rng.permutation(n)[:ne], j] ^= rng.integers(1, 256, ne, dtype=np.uint8)
    idx = torch.from_numpy(cols).cuda()
    bad = dev(host_rx)
    for formal in (True, False):
        rx = cw.clone().view(n, L)
        rx[:, idx] = bad
        ok, counts = c.decode_systematic_sharded(rx.view(-1), L, formal_derivative=formal)
        assert counts[0] == L - len(cols)
        sub = rx[:, idx].cpu().numpy()
        if formal:
            assert ok and counts == (L - len(cols), len(cols), 0)
            assert torch.equal(rx.view(-1), cw)
        else:
            exp = np.ascontiguousarray(host_rx).reshape(-1).copy()
            O.decode_sharded(n, T, exp, len(cols))
            assert np.array_equal(sub, exp.reshape(n, len(cols)))
