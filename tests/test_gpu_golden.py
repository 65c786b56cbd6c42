"""The HIP path through the C ABI against the committed fixtures
(tests/golden/fixtures.json): encode, recover and error decoding (both
modes), byte for byte, with the decode statuses."""
import json
import os

import numpy as np
import pytest

import _pkg

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fixtures.json")))


@pytest.fixture(scope="module")
def ars():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _pkg.load()


def b(h):
    return np.frombuffer(bytes.fromhex(h), np.uint8).copy()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_encode_fixtures(ars):
    for r in FIX["encode"]:
        n, k, L = r["n"], r["k"], r["shard_len"]
        dp = torch.zeros((n - k) * L, dtype=torch.uint8, device="cuda")
        ars.Codec(n, k).encode_systematic_sharded(dev(b(r["message"])), dp, None, L, accumulate=False)
        assert dp.cpu().numpy().tobytes().hex() == r["parity"], (n, k, L)


def test_recover_fixtures(ars):
    for r in FIX["recover"]:
        n, k, L = r["n"], r["k"], r["shard_len"]
        d = dev(b(r["received"]))
        assert ars.Codec(n, k).recover_erasure_shards(d, None, L, r["positions"])
        assert d.cpu().numpy().tobytes().hex() == r["recovered"], (n, k, L, len(r["positions"]))


def test_decode_fixtures(ars):
    for r in FIX["decode"]:
        n, k, L = r["n"], r["k"], r["shard_len"]
        for mode in ("reference", "formal"):
            d = dev(b(r["received"]))
            st = torch.zeros(L, dtype=torch.uint8, device="cuda")
            ok, _ = ars.Codec(n, k).decode_systematic_sharded(d, L, formal_derivative=mode == "formal", status=st)
            assert int(ok) == r[mode]["ok"], (n, k, mode)
            assert list(st.cpu().numpy()) == r[mode]["status"], (n, k, mode)
            assert d.cpu().numpy().tobytes().hex() == r[mode]["decoded"], (n, k, mode)
