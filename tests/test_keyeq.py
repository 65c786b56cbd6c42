"""CPU unit test of the product's host key-equation solver
(csrc/ars_keyeq.hpp), built with g++ and compared with the oracle's
restatement of the reference's HGCD solver (rs/src/codec.rs:212-239) and its
derivative (rs/src/poly_arith.rs:91-116): lambda, v and the derivative in
the basis X, and deg lambda, for random syndromes of every degree and for the
syndromes of received words with errors within and beyond T/2."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import _pkg
import oracle as O

CSRC = os.path.join(_pkg.PKG_DIR, "csrc")

DRIVER = r'''
#include "ars_keyeq.hpp"
#include <cstdio>
#include <vector>
int main(int argc, char **argv) {
    FILE *f = std::fopen(argv[1], "rb");
    unsigned hdr[4];
    if (std::fread(hdr, 4, 4, f) != 4) return 2;
    const unsigned n = hdr[0], T = hdr[1], formal = hdr[2], count = hdr[3];
    std::vector<uint8_t> syn((size_t)count * T);
    if (std::fread(syn.data(), 1, syn.size(), f) != syn.size()) return 3;
    std::fclose(f);
    std::vector<ars::DecPlan> plans(count);
    std::vector<uint8_t> ok(count);
    for (unsigned i = 0; i < count; ++i)
        ok[i] = ars::solve_key_equation(&syn[(size_t)i * T], n, T, formal != 0, plans[i]);
    FILE *o = std::fopen(argv[2], "wb");
    std::fwrite(ok.data(), 1, count, o);
    std::fwrite(plans.data(), sizeof(ars::DecPlan), count, o);
    std::fclose(o);
    return 0;
}
'''

PLAN = np.dtype([("lam", np.uint8, 128), ("v", np.uint8, 128), ("lp", np.uint8, 128),
                 ("deg", np.int16), ("pad", np.uint8, 2)])


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ missing")
    d = tmp_path_factory.mktemp("keyeq")
    src = d / "keyeq.cpp"
    src.write_text(DRIVER)
    exe = d / "keyeq"
    # under AddressSanitizer and UndefinedBehaviorSanitizer: every finding aborts
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", CSRC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return d, exe


def solve(driver, n, T, syns, formal):
    d, exe = driver
    inp, out = d / "in.bin", d / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, T, int(formal), len(syns)], np.uint32).tobytes())
        f.write(np.ascontiguousarray(np.array(syns, np.uint8)).tobytes())
    r = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    ok = raw[:len(syns)]
    plans = np.frombuffer(raw[len(syns):].tobytes(), PLAN)
    return ok, plans


def syndromes(n, T, rng, count):
    out = []
    k = n - T
    for i in range(count):
        if i % 3 == 0:  # a random syndrome of random degree
            s = np.zeros(T, np.uint8)
            d = int(rng.integers(0, T))
            s[:d + 1] = rng.integers(0, 256, d + 1)
            if not s.any():
                s[0] = 1
        else:  # the syndrome of a codeword with errors
            msg = rng.integers(0, 256, k, dtype=np.uint8)
            c = np.concatenate([O.encode(n, T, msg, 1), msg])
            ne = int(rng.integers(1, T // 2 + 1)) if i % 3 == 1 else int(rng.integers(T // 2 + 1, T + 3))
            pos = rng.permutation(n)[:min(ne, n)]
            c[pos] ^= rng.integers(1, 256, len(pos), dtype=np.uint8)
            s = O.syndrome(n, T, c)[:T]
            if not s.any():
                continue
        out.append(s)
    return out


@pytest.mark.parametrize("n,T", [(4, 2), (8, 4), (16, 8), (32, 16), (64, 32), (128, 64), (256, 128),
                                 (256, 32), (256, 2), (64, 4), (128, 8)])
def test_keyeq_matches_oracle_hgcd(driver, n, T):
    rng = np.random.default_rng(n * 31 + T)
    syns = syndromes(n, T, rng, 45)
    for formal in (False, True):
        ok, plans = solve(driver, n, T, syns, formal)
        for s, good, p in zip(syns, ok, plans):
            assert good
            v1, lam = O.key_equation(np.concatenate([s, np.zeros(n - T, np.uint8)]), n, T)
            dl = max(list(np.nonzero(lam)[0]) + [-1])
            assert p["deg"] == dl
            assert np.array_equal(p["lam"][:T], lam[:T])
            assert np.array_equal(p["v"][:T], v1[:T])
            d = O.formal_deriv(lam, n) if formal else O.deriv_lnh(lam, n)
            assert np.array_equal(p["lp"][:T], d[:T])


def test_keyeq_zero_syndrome_is_rejected(driver):
    ok, _ = solve(driver, 16, 8, [np.zeros(8, np.uint8)], False)
    assert not ok[0]
