"""The CPU oracle (oracle/ars_oracle.c, oracle/ars_oracle_gfni.c) under
AddressSanitizer and UndefinedBehaviorSanitizer.

A C driver, linked with both restatements built with
-fsanitize=address,undefined (every finding aborts), runs each exported
codec function on buffers of exactly the sizes the reference's callers pass
(so an overrun is an ASan error, not a silent read into slack), over every
pow2 (n, T) with n <= 64 plus (256,128), (256,16), (256,4) and ragged shard
lengths, and checks the round trips: LUT encode == GFNI encode, recover of
zeroed and garbage-filled erasures (c ^ g), recompute, error decoding of
<= T/2 errors per column in the formal-derivative mode, the k = 2^KL
extension encode and the Lagrange recover, and the key-equation solvers and
novel-basis polynomial arithmetic on random input."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

DRIVER = r'''
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int aro_encode_systematic_sharded(size_t n, size_t T, const uint8_t *m, uint8_t *p, uint8_t *ws, size_t L);
int aro_recover_erasure_shards(size_t n, size_t T, uint8_t *r, uint8_t *ws, size_t L, const uint8_t *pos, size_t e);
int aro_recompute_data_from_parity_sharded(size_t n, size_t T, uint8_t *r, uint8_t *ws, size_t L);
int aro_gfni_supported(void);
void aro_gfni_set_unrolled(int on);
int aro_gfni_encode_systematic_sharded(size_t n, size_t T, const uint8_t *m, uint8_t *p, uint8_t *ws, size_t L);
int aro_gfni_recover_erasure_shards(size_t n, size_t T, uint8_t *r, uint8_t *ws, size_t L, const uint8_t *pos,
                                    size_t e);
int aro_decode_sharded(size_t n, size_t T, uint8_t *r, size_t L, uint8_t *status, int formal);
int aro_ext_encode_kpow2(size_t n, size_t k, const uint8_t *m, uint8_t *p, size_t L);
int aro_ext_recover_lagrange(size_t n, size_t k, uint8_t *r, size_t L, const uint8_t *pos, size_t e);
int aro_key_equation(const uint8_t *syn, size_t n, size_t T, int use_eea, uint8_t *v1, uint8_t *lambda);
int aro_poly_mul_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out);
int aro_poly_div_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *q, uint8_t *r);
void aro_deriv_lnh(const uint8_t *c, size_t n, uint8_t *out);
void aro_formal_deriv(const uint8_t *c, size_t n, uint8_t *out);
void aro_syndrome_scalar(const uint8_t *r, size_t n, size_t T, uint8_t *syn);

static uint64_t s = 88172645463325252ull;
static uint8_t rnd(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (uint8_t)(s >> 24); }
static uint8_t *buf(size_t len) { uint8_t *b = malloc(len ? len : 1); for (size_t i = 0; i < len; i++) b[i] = rnd(); return b; }
static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { fails++; printf(__VA_ARGS__); printf("\n"); } } while (0)

static void positions(size_t n, size_t e, uint8_t *pos) {
    uint8_t perm[256];
    for (size_t i = 0; i < n; i++) perm[i] = (uint8_t)i;
    for (size_t i = n - 1; i > 0; i--) { size_t j = rnd() % (i + 1); uint8_t t = perm[i]; perm[i] = perm[j]; perm[j] = t; }
    memcpy(pos, perm, e);
}

static void codec(size_t n, size_t T, size_t L, int gfni) {
    size_t k = n - T;
    uint8_t *msg = buf(k * L), *par = calloc(T * L, 1), *ws = buf(T * L);
    CHECK(aro_encode_systematic_sharded(n, T, msg, par, ws, L) == 0, "encode %zu %zu", n, T);
    if (gfni) {
        uint8_t *par2 = calloc(T * L, 1), *ws2 = buf(T * L);
        CHECK(aro_gfni_encode_systematic_sharded(n, T, msg, par2, ws2, L) == 0, "gfni encode");
        CHECK(memcmp(par, par2, T * L) == 0, "gfni encode differs %zu %zu %zu", n, T, L);
        free(par2); free(ws2);
    }
    uint8_t *cw = malloc(n * L);
    memcpy(cw, par, T * L);
    memcpy(cw + T * L, msg, k * L);
    size_t es[3] = {1, T / 2 ? T / 2 : 1, T};
    for (int t = 0; t < 3; t++) {
        uint8_t pos[256];
        size_t e = es[t];
        positions(n, e, pos);
        for (int g = 0; g <= gfni; g++) {
            uint8_t *rx = malloc(n * L), *w = buf(n * L), *gar = buf(e * L);
            memcpy(rx, cw, n * L);
            for (size_t i = 0; i < e; i++) memcpy(rx + (size_t)pos[i] * L, gar + i * L, L);
            int rc = g ? aro_gfni_recover_erasure_shards(n, T, rx, w, L, pos, e)
                       : aro_recover_erasure_shards(n, T, rx, w, L, pos, e);
            CHECK(rc == 1, "recover rc %d", rc);
            for (size_t i = 0; i < e; i++)
                for (size_t j = 0; j < L; j++) rx[(size_t)pos[i] * L + j] ^= gar[i * L + j];
            CHECK(memcmp(rx, cw, n * L) == 0, "recover %s %zu %zu %zu e=%zu", g ? "gfni" : "lut", n, T, L, e);
            free(rx); free(w); free(gar);
        }
    }
    if (n == 2 * T) {
        uint8_t *rx = malloc(n * L), *w = buf(T * L);
        memcpy(rx, cw, n * L);
        memset(rx + T * L, 0, k * L);
        CHECK(aro_recompute_data_from_parity_sharded(n, T, rx, w, L) == 0, "recompute rc");
        CHECK(memcmp(rx, cw, n * L) == 0, "recompute %zu", n);
        free(rx); free(w);
    }
    if (T >= 2 && n <= 64) {
        /* <= T/2 errors in every column, formal-derivative mode */
        uint8_t *rx = malloc(n * L), *st = malloc(L);
        memcpy(rx, cw, n * L);
        for (size_t j = 0; j < L; j++) {
            uint8_t pos[256];
            size_t ne = rnd() % (T / 2 + 1);
            positions(n, ne, pos);
            for (size_t i = 0; i < ne; i++) rx[(size_t)pos[i] * L + j] ^= (uint8_t)(1 + rnd() % 255);
        }
        CHECK(aro_decode_sharded(n, T, rx, L, st, 1) == 1, "decode rc");
        CHECK(memcmp(rx, cw, n * L) == 0, "decode %zu %zu %zu", n, T, L);
        /* the reference's mode on the same kind of input: must not fault */
        memcpy(rx, cw, n * L);
        rx[rnd() % (n * L)] ^= 1;
        (void)aro_decode_sharded(n, T, rx, L, st, 0);
        free(rx); free(st);
    }
    free(msg); free(par); free(ws); free(cw);
}

static void extension(size_t n, size_t k, size_t L) {
    size_t T = n - k;
    uint8_t *msg = buf(k * L), *par = malloc(T * L);
    CHECK(aro_ext_encode_kpow2(n, k, msg, par, L) == 0, "ext encode");
    uint8_t *cw = malloc(n * L);
    memcpy(cw, par, T * L);
    memcpy(cw + T * L, msg, k * L);
    uint8_t pos[256];
    positions(n, T, pos);
    uint8_t *rx = malloc(n * L);
    memcpy(rx, cw, n * L);
    for (size_t i = 0; i < T; i++) memset(rx + (size_t)pos[i] * L, 0, L);
    CHECK(aro_ext_recover_lagrange(n, k, rx, L, pos, T) == 1, "lagrange rc");
    CHECK(memcmp(rx, cw, n * L) == 0, "lagrange %zu %zu", n, k);
    free(msg); free(par); free(cw); free(rx);
}

static void algebra(size_t n, size_t T) {
    uint8_t *a = buf(n), *b = buf(n), *o = malloc(n), *q = malloc(n), *r = malloc(n);
    for (size_t i = n / 2; i < n; i++) a[i] = b[i] = 0; /* products stay below n */
    (void)aro_poly_mul_lnh(a, b, n, o);
    (void)aro_poly_div_lnh(a, b, n, q, r);
    aro_deriv_lnh(a, n, o);
    aro_formal_deriv(a, n, o);
    uint8_t *w = buf(n), *syn = malloc(n), *v1 = malloc(n), *lam = malloc(n);
    aro_syndrome_scalar(w, n, T, syn);
    for (int eea = 0; eea < 2; eea++) (void)aro_key_equation(syn, n, T, eea, v1, lam);
    free(a); free(b); free(o); free(q); free(r); free(w); free(syn); free(v1); free(lam);
}

int main(void) {
    int gfni = aro_gfni_supported();
    size_t Ls[4] = {1, 3, 64, 67};
    for (size_t n = 2; n <= 64; n *= 2)
        for (size_t T = 1; T < n; T *= 2)
            for (int l = 0; l < 4; l++) codec(n, T, Ls[l], gfni);
    size_t big[3][2] = {{256, 128}, {256, 16}, {256, 4}};
    for (int i = 0; i < 3; i++) codec(big[i][0], big[i][1], 67, gfni);
    if (gfni) {
        aro_gfni_set_unrolled(0);
        codec(64, 32, 67, 1);
        codec(256, 128, 5, 1);
        aro_gfni_set_unrolled(1);
    }
    extension(16, 4, 5);
    extension(32, 8, 3);
    extension(256, 16, 2);
    for (size_t n = 4; n <= 256; n *= 2) algebra(n, n / 2);
    printf("fails=%d gfni=%d\n", fails, gfni);
    return fails != 0;
}
'''


def test_oracle_under_asan_ubsan(tmp_path):
    if not shutil.which("gcc"):
        pytest.skip("gcc missing")
    src = tmp_path / "driver.c"
    src.write_text(DRIVER)
    exe = tmp_path / "driver"
    r = subprocess.run(["gcc", "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", str(src), os.path.join(ORACLE, "ars_oracle.c"),
                        os.path.join(ORACLE, "ars_oracle_gfni.c"), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "fails=0" in r.stdout
