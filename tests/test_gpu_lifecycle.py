"""Context lifecycle and concurrent use: contexts created and destroyed in a
loop keep producing oracle-exact results (each destroy frees the context's
host-slice pipeline and decode scratch), and two threads with their own
contexts can run host-slice encode / recover and error decoding at once."""
import threading

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


def work(ars, ctx, n, k, L, seed):
    """Host-slice encode + recover and a device decode on `ctx`; returns
    a list of failure descriptions (empty when all match the oracle)."""
    T = n - k
    rng = np.random.default_rng(seed)
    bad = []
    msg = rng.integers(0, 256, k * L, dtype=np.uint8)
    par = np.zeros(T * L, np.uint8)
    rc = ars.lib().ars_encode_host(ctx.handle, n, k, msg.ctypes.data, par.ctypes.data, L, 0)
    exp = O.encode(n, T, msg, L)
    if rc != 0 or not np.array_equal(par, exp):
        bad.append(f"encode seed {seed}")
    cw = np.concatenate([exp, msg])
    pos = rng.permutation(n)[:T].astype(np.uint8)
    rx = cw.copy().reshape(n, L)
    rx[pos] = 0
    rx = rx.reshape(-1).copy()
    rc = ars.lib().ars_recover_host(ctx.handle, n, k, rx.ctypes.data, L, pos.ctypes.data, len(pos))
    if rc != 1 or not np.array_equal(rx, cw):
        bad.append(f"recover seed {seed}")
    er = cw.copy().reshape(n, L)
    for j in range(L):
        ne = int(rng.integers(0, T // 2 + 1))
        er[rng.permutation(n)[:ne], j] ^= rng.integers(1, 256, ne, dtype=np.uint8)
    d = torch.from_numpy(er.reshape(-1).copy()).cuda()
    rc, counts = ctx.decode(n, k, d.data_ptr(), L, ars.ARS_DEC_FORMAL_DERIVATIVE)
    torch.cuda.synchronize()
    if rc != 1 or counts[2] != 0 or not np.array_equal(d.cpu().numpy(), cw):
        bad.append(f"decode seed {seed}")
    return bad


def test_contexts_come_and_go(ars):
    for it in range(6):
        ctx = ars.Context(0)
        bad = work(ars, ctx, 64, 32, 3001 + 977 * it, it)
        ctx.close()
        assert not bad, bad


def test_two_threads_two_contexts(ars):
    results = {}

    def run(i):
        ctx = ars.Context(0)
        try:
            out = []
            for r in range(3):
                out += work(ars, ctx, 128, 64, 20000 + 4096 * r + i, 10 * i + r)
            results[i] = out
        finally:
            ctx.close()

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=300)
    assert set(results) == {0, 1}
    assert not results[0] and not results[1], results


def test_host_slices_after_a_rejected_call(ars):
    """A host-slice call that fails (duplicate positions) leaves nothing in
    flight that a later call on the same context could pick up."""
    ctx = ars.Context(0)
    try:
        n, k, L = 64, 32, 300000  # several pipeline chunks
        T = n - k
        rng = np.random.default_rng(77)
        msg = rng.integers(0, 256, k * L, dtype=np.uint8)
        cw = np.concatenate([O.encode(n, T, msg, L), msg])
        rx = cw.copy()
        dup = np.array([3, 5, 3], np.uint8)
        rc = ars.lib().ars_recover_host(ctx.handle, n, k, rx.ctypes.data, L, dup.ctypes.data, len(dup))
        assert rc == ars.ARS_E_DUPLICATE
        for seed in range(2):
            bad = work(ars, ctx, n, k, L // 3 + seed, 200 + seed)
            assert not bad, bad
    finally:
        ctx.close()
