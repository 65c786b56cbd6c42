"""GPU parity: the HIP path (through the C ABI) against the CPU oracle, byte
for byte, on seeded inputs; plus size-independent properties at full size."""
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


def gpu_encode(ars, n, k, msg, L, parity_in=None, accumulate=False):
    c = ars.Codec(n, k)
    dm = dev(msg)
    dp = dev(parity_in) if parity_in is not None else torch.zeros((n - k) * L, dtype=torch.uint8, device="cuda")
    c.encode_systematic_sharded(dm, dp, None, L, accumulate=accumulate)
    return host(dp)


def gpu_recover(ars, n, k, rx, L, pos):
    c = ars.Codec(n, k)
    d = dev(rx)
    ok = c.recover_erasure_shards(d, None, L, pos)
    return ok, host(d)


POW2 = [(n, T) for n in (2, 4, 8, 16, 32, 64, 128, 256) for T in (1, 2, 4, 8, 16, 32, 64, 128) if T < n]


@pytest.mark.parametrize("n,T", POW2)
@pytest.mark.parametrize("L", [1, 64, 100, 1030])
def test_encode_matches_oracle(ars, n, T, L):
    rng = np.random.default_rng(n * 7919 + T * 31 + L)
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    exp = O.encode(n, T, msg, L)
    got = gpu_encode(ars, n, k, msg, L)
    assert np.array_equal(got, exp), (n, T, L)


@pytest.mark.parametrize("n,T", [(4, 2), (32, 8), (64, 32), (256, 128), (256, 32)])
def test_encode_accumulate_matches_reference_semantics(ars, n, T):
    rng = np.random.default_rng(17 + n + T)
    L = 256
    k = n - T
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    pin = rng.integers(0, 256, T * L, dtype=np.uint8)
    exp = O.encode(n, T, msg, L, parity=pin.copy())
    got = gpu_encode(ars, n, k, msg, L, parity_in=pin, accumulate=True)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("n,k", [(8, 2), (16, 4), (32, 8), (64, 16), (256, 16), (256, 32), (256, 64)])
def test_encode_kpow2_extension(ars, n, k):
    rng = np.random.default_rng(n * 3 + k)
    L = 128
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    exp = O.ext_encode_kpow2(n, k, msg, L)
    assert np.array_equal(gpu_encode(ars, n, k, msg, L), exp)


def make_cw(n, k, L, rng):
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    return np.concatenate([O.encode_any(n, k, msg, L), msg])


@pytest.mark.parametrize("n,T", POW2)
@pytest.mark.parametrize("L", [1, 64, 132])
def test_recover_matches_oracle(ars, n, T, L):
    rng = np.random.default_rng(n * 131 + T * 17 + L)
    k = n - T
    cw = make_cw(n, k, L, rng)
    for e in sorted({1, max(1, T // 2), T}):
        pos = rng.permutation(n)[:e].astype(np.uint8)
        rx = cw.copy().reshape(n, L)
        rx[pos] = rng.integers(0, 256, (e, L), dtype=np.uint8) if e % 2 else 0
        rx = rx.reshape(-1)
        exp = rx.copy()
        assert O.recover(n, T, exp, L, pos) == 1
        ok, got = gpu_recover(ars, n, k, rx, L, pos)
        assert ok
        assert np.array_equal(got, exp), (n, T, L, e)


@pytest.mark.parametrize("n,T", POW2)
def test_recover_reference_odd_pattern(ars, n, T):
    # rs/src/poly_11d_lut.rs:536-566 at L = 1 and L = 64
    k = n - T
    for L in (1, 64):
        msg = np.array([(T + i + j) & 0xFF for i in range(k) for j in range(L)], np.uint8)
        cw = np.concatenate([O.encode(n, T, msg, L), msg])
        pos = list(range(1, 2 * T, 2))
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        ok, got = gpu_recover(ars, n, k, rx.reshape(-1), L, pos)
        assert ok and np.array_equal(got, cw)


@pytest.mark.parametrize("n,k", [(16, 4), (32, 8), (256, 16)])
def test_recover_kpow2_extension(ars, n, k):
    rng = np.random.default_rng(5 * n + k)
    L = 64
    T = n - k
    cw = make_cw(n, k, L, rng)
    for e in (1, T // 3, T):
        pos = rng.permutation(n)[:e].astype(np.uint8)
        rx = cw.copy().reshape(n, L)
        rx[pos] = 0
        ok, got = gpu_recover(ars, n, k, rx.reshape(-1), L, pos)
        assert ok and np.array_equal(got, cw), (n, k, e)


def test_recover_edge_cases(ars):
    rng = np.random.default_rng(1)
    n, k, L = 64, 32, 64
    cw = make_cw(n, k, L, rng)
    # e > T -> false, untouched (rs/src/codec.rs:491-493)
    ok, got = gpu_recover(ars, n, k, cw.copy(), L, list(range(33)))
    assert not ok and np.array_equal(got, cw)
    # e == 0 -> true, untouched
    ok, got = gpu_recover(ars, n, k, cw.copy(), L, [])
    assert ok and np.array_equal(got, cw)
    c = ars.Codec(n, k)
    d = dev(cw)
    with pytest.raises(ars.ArsError) as ei:
        c.recover_erasure_shards(d, None, L, [3, 64])
    assert ei.value.code == ars.ARS_E_POSITION
    with pytest.raises(ars.ArsError) as ei:
        c.recover_erasure_shards(d, None, L, [3, 5, 3])
    assert ei.value.code == ars.ARS_E_DUPLICATE
    # bad (n, k) is rejected at the C ABI (the reference's const generics
    # would not compile; rs/src/lib.rs:16-19)
    for bad_n, bad_k in [(48, 16), (512, 256), (64, 64), (64, 0), (256, 100)]:
        with pytest.raises(ars.ArsError) as ei:
            c.ctx.encode(bad_n, bad_k, d.data_ptr(), d.data_ptr(), L)
        assert ei.value.code in (ars.ARS_E_INVAL, ars.ARS_E_UNSUPPORTED)


def guarded(data, shift, pad=256):
    """`data` on the device starting `shift` bytes past a 4-byte boundary,
    with `pad` guard bytes of 0xA5 on both sides; returns (buffer, ptr)."""
    buf = torch.full((len(data) + shift + pad + 4,), 0xA5, dtype=torch.uint8, device="cuda")
    base = pad + shift
    assert (buf.data_ptr() + pad) % 4 == 0
    buf[base:base + len(data)] = dev(data)
    return buf, buf.data_ptr() + base


def unguard(buf, shift, size, pad=256):
    h = host(buf)
    base = pad + shift
    assert np.all(h[:base] == 0xA5) and np.all(h[base + size:] == 0xA5), "write outside the buffer"
    return h[base:base + size]


@pytest.mark.parametrize("n,T", [(16, 8), (64, 32), (256, 128), (256, 16), (128, 64)])
@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_guarded_buffers_match_oracle(ars, n, T, shift):
    """Shard length 68 (17 column quads: the last tile is partial) on buffers
    that start `shift` bytes past a 4-byte boundary.  Shift 0 takes the dword
    I/O path; recover at shifts 1-3 the shifted dword path (one buffer, one
    shift); encode at shifts 1 and 3 the byte-wise path (message and parity
    at different shifts) and at shift 2 the shifted dword path; results and
    the guard bytes on both sides of every buffer are exact."""
    k, L = n - T, 68
    rng = np.random.default_rng(n * 7 + T * 3 + shift)
    ctx = ars.Context(0)
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    exp = O.encode_any(n, k, msg, L)
    bm, pm = guarded(msg, shift)
    ps = (4 - shift) % 4
    bp, pp = guarded(np.zeros(T * L, np.uint8), ps)
    ctx.encode(n, k, pm, pp, L)
    assert np.array_equal(unguard(bp, ps, T * L), exp)
    assert np.array_equal(unguard(bm, shift, k * L), msg)
    cw = np.concatenate([exp, msg])
    pos = rng.permutation(n)[:T].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = 0
    rx = rx.reshape(-1)
    b, p = guarded(rx, shift)
    assert ctx.recover(n, k, p, L, pos) == 1
    assert np.array_equal(unguard(b, shift, n * L), cw)
    plan = torch.empty(ars.lib().ars_recover_plan_size(), dtype=torch.uint8, device="cuda")
    assert ctx.recover_plan(n, k, pos, plan.data_ptr()) == 1
    b, p = guarded(rx, shift)
    assert ctx.recover_with_plan(n, k, p, L, plan.data_ptr()) == 1
    assert np.array_equal(unguard(b, shift, n * L), cw)
    if n == 2 * T:
        lost = cw.copy()
        lost[T * L:] = 0
        b, p = guarded(lost, shift)
        ctx.recompute_data(n, k, p, L)
        assert np.array_equal(unguard(b, shift, n * L), cw)
    ctx.close()


@pytest.mark.parametrize("shift", [1, 2, 3])
@pytest.mark.parametrize("n,T", [(256, 128), (64, 32)])
def test_shifted_dword_path_equals_aligned(ars, n, T, shift):
    """Message and parity (and the codeword) both `shift` bytes past a 4-byte
    boundary with shard_len % 4 == 0: the shifted dword path (each lane's
    dword spans the boundary between columns) on launches of many tiles,
    against the same code on aligned buffers, and no byte outside the
    buffers written."""
    k, L = n - T, 65536 + 68
    rng = np.random.default_rng(n + T + shift)
    ctx = ars.Context(0)
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    bm0, pm0 = guarded(msg, 0)
    bp0, pp0 = guarded(np.zeros(T * L, np.uint8), 0)
    ctx.encode(n, k, pm0, pp0, L)
    exp = unguard(bp0, 0, T * L)
    # the aligned result, pinned on the first and last 8 columns by the oracle
    for cols in (slice(0, 8), slice(L - 8, L)):
        assert np.array_equal(exp.reshape(T, L)[:, cols],
                              O.encode(n, T, msg.reshape(k, L)[:, cols].copy().reshape(-1), 8).reshape(T, 8))
    bm, pm = guarded(msg, shift)
    bp, pp = guarded(np.zeros(T * L, np.uint8), shift)
    ctx.encode(n, k, pm, pp, L)
    assert np.array_equal(unguard(bp, shift, T * L), exp)
    assert np.array_equal(unguard(bm, shift, k * L), msg)
    cw = np.concatenate([exp, msg])
    pos = rng.permutation(n)[:T].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = rng.integers(0, 256, (T, L), dtype=np.uint8)  # garbage: the result is c ^ g
    want = cw.copy().reshape(n, L)
    want[pos] ^= rx[pos]
    b, p = guarded(rx.reshape(-1), shift)
    assert ctx.recover(n, k, p, L, pos) == 1
    assert np.array_equal(unguard(b, shift, n * L), want.reshape(-1))
    ctx.close()


def test_recompute_data_matches_oracle(ars):
    rng = np.random.default_rng(2)
    for n in (2, 4, 8, 16, 32, 64, 128, 256):
        T, L = n // 2, 96
        cw = make_cw(n, n - T, L, rng)
        rx = cw.copy()
        rx[T * L:] = 0
        d = dev(rx)
        ars.Codec(n, n - T).recompute_data_from_parity_sharded(d, None, L)
        assert np.array_equal(host(d), cw), n


@pytest.mark.parametrize("n,k,L,e", [(256, 128, 200 * 1024 + 64, 128), (64, 32, 70001, 32), (256, 16, 150000, 240),
                                     (256, 128, 131072 + 1001, 3), (128, 64, 262144 + 4097, 1)])
def test_host_slice_streaming_chunks(ars, n, k, L, e):
    """Host slices larger than one pipeline chunk, ragged last chunks, few or
    many erasures."""
    rng = np.random.default_rng(L)
    T = n - k
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = np.zeros(T * L, np.uint8)
    c = ars.Codec(n, k)
    c.encode_systematic_sharded(msg, par, None, L, accumulate=False)
    exp = O.encode_any(n, k, msg, L)
    assert np.array_equal(par, exp)
    if T & (T - 1) == 0:  # accumulate into non-zero parity, chunk by chunk
        par0 = rng.integers(0, 256, T * L, dtype=np.uint8)
        acc = par0.copy()
        c.encode_systematic_sharded(msg, acc, None, L)
        assert np.array_equal(acc, O.encode(n, T, msg, L, parity=par0.copy(), gfni=O.gfni_supported()))
    cw = np.concatenate([par, msg])
    pos = rng.permutation(n)[:e].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = 0
    rx = rx.reshape(-1).copy()
    assert c.recover_erasure_shards(rx, None, L, pos)
    assert np.array_equal(rx, cw)


def test_host_slice_variants(ars):
    rng = np.random.default_rng(3)
    n, k, L = 128, 96, 256
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = np.zeros((n - k) * L, np.uint8)
    ars.Codec(n, k).encode_systematic_sharded(msg, par, None, L)
    assert np.array_equal(par, O.encode(n, n - k, msg, L))
    cw = np.concatenate([par, msg])
    pos = rng.permutation(n)[:32].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = 0
    rx = rx.reshape(-1).copy()
    assert ars.Codec(n, k).recover_erasure_shards(rx, None, L, pos)
    assert np.array_equal(rx, cw)


def test_full_size_256_128_1mib(ars):
    """The headline shape: (256,128), 1 MiB shards.  Encode is compared with
    the oracle byte for byte; recover must return the original codeword
    (uniqueness) and equal the oracle's recover."""
    n, k, L = 256, 128, 1 << 20
    T = n - k
    rng = np.random.default_rng(43)
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    use_gfni = O.gfni_supported()
    exp_par = O.encode(n, T, msg, L, gfni=use_gfni)
    dm = dev(msg)
    dp = torch.zeros(T * L, dtype=torch.uint8, device="cuda")
    ars.Codec(n, k).encode_systematic_sharded(dm, dp, None, L, accumulate=False)
    got_par = host(dp)
    assert np.array_equal(got_par, exp_par)
    cw = torch.cat([dp, dm])
    pos = rng.permutation(n)[:T].astype(np.uint8)
    cwv = cw.view(n, L)
    cwv[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
    assert ars.Codec(n, k).recover_erasure_shards(cw, None, L, pos)
    got = host(cw)
    assert np.array_equal(got[:T * L], exp_par)
    assert np.array_equal(got[T * L:], msg)


@pytest.mark.parametrize("L", [(1 << 24) - 64, (1 << 24) - 1])
def test_largest_codeword_under_4gib(ars, L):
    """(256,128) with n * shard_len just under the 4 GiB limit, on the dword
    path (L % 4 == 0) and on the byte path.  Columns are independent, so the
    encode is compared with the oracle on column slices at both ends of the
    shards (the largest row and column offsets); recover of 128 erasures must
    return the whole 4 GiB codeword."""
    n, k = 256, 128
    T = n - k
    assert n * L < 1 << 32
    ctx = ars.Context(0)
    g = torch.Generator(device="cuda")
    g.manual_seed(L)
    cw = torch.empty(n * L, dtype=torch.uint8, device="cuda")
    cw[T * L:] = torch.randint(0, 256, (k * L,), dtype=torch.uint8, device="cuda", generator=g)
    assert ctx.encode(n, k, cw.data_ptr() + T * L, cw.data_ptr(), L) == 0
    torch.cuda.synchronize()
    rows = cw.view(n, L)
    w = 4099
    for c0 in (0, L - w):
        sl = rows[:, c0:c0 + w].cpu().numpy()
        exp = O.encode(n, T, np.ascontiguousarray(sl[T:]).reshape(-1), w)
        assert np.array_equal(np.ascontiguousarray(sl[:T]).reshape(-1), exp), c0
    ref = cw.clone()
    pos = np.random.default_rng(L).permutation(n)[:T].astype(np.uint8)
    rows[torch.from_numpy(pos.astype(np.int64)).cuda()] = 0
    assert ctx.recover(n, k, cw.data_ptr(), L, pos) == 1
    torch.cuda.synchronize()
    assert torch.equal(cw, ref)
    del cw, ref, rows
    torch.cuda.empty_cache()
    ctx.close()


@pytest.mark.parametrize("n,T", [(2, 1), (16, 8), (32, 16), (64, 32), (128, 64), (256, 128), (256, 16)])
def test_fused_and_split_recover_agree(ars, n, T):
    """ars_recover computes the erasure plan inside the recover kernel when
    its launch has at most one workgroup per compute unit, and launches the
    plan kernel first otherwise.  On both sides of that threshold (and on the
    byte path) the result equals the oracle's, garbage in the erased rows
    included."""
    k = n - T
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    cols = 1024 if n <= 16 else (128 if n == 256 else 256)  # byte columns per workgroup
    rng = np.random.default_rng(n * 5 + T)
    ctx = ars.Context(0)
    for L in (cols * cus, cols * cus + 4, cols * cus + 3):
        cw = make_cw(n, k, L, rng)
        pos = rng.permutation(n)[:T].astype(np.uint8)
        rx = cw.copy().reshape(n, L)
        rx[pos] = rng.integers(0, 256, (T, L), dtype=np.uint8)
        rx = rx.reshape(-1)
        exp = rx.copy()
        assert O.recover(n, T, exp, L, pos) == 1
        d = dev(rx)
        assert ctx.recover(n, k, d.data_ptr(), L, pos) == 1
        assert np.array_equal(host(d), exp), (n, T, L)
    ctx.close()
