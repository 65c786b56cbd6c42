/*
 * ars_oracle.c -- CPU ORACLE for the sharded additive-FFT Reed-Solomon path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the algorithm of
 * the reference crate (vkomenda/additive-fft-reed-solomon) for the sharded
 * encoder and the sharded erasure recoverer, so that the GPU path can be
 * checked against it byte for byte.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product library never does.
 *
 * Every function cites the reference file:line it follows.  Parity of this
 * restatement is pinned (see tests/test_oracle.py) by:
 *   - the reference's own known-answer tests (Cantor basis KAT
 *     rs/src/poly_11d_lut.rs:419-429, 2*0x80 = 0x1d gf2p8/src/poly_11d.rs:77-82,
 *     normalisation factors = 1 rs/src/poly_11d_lut.rs:123-129, derivative
 *     constants = 1 gf2p8/src/poly_11d.rs:204-213, subspace polynomials equal
 *     the brute-force product gf2p8/src/poly_11d.rs:187-202, exhaustive GFNI
 *     matrix == mul rs/src/poly_11d_lut.rs:68-81);
 *   - closed forms checked independently in Python: FFT == direct evaluation
 *     of sum_j d_j X_j(beta + w_i) (rs/src/gf2p8lut.rs:91-114), encode ==
 *     Lagrange interpolation through the message points, recover == original
 *     codeword for any erasure set of size <= T (uniqueness), and the
 *     reference's round-trip tests (rs/src/poly_11d_lut.rs:431-566).
 *
 * Layout (the reference's): a codeword of n shards, each shard_len bytes,
 * shard i at bytes [i*L, (i+1)*L); parity shards 0..T-1, message T..n-1
 * (rs/src/poly_11d_lut.rs:100-121).
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>

#define FIELD 256
#define EXP_SIZE 510

static uint8_t EXP[EXP_SIZE], LOG[FIELD], INV[FIELD];
static uint8_t BASIS[8], SUBSPACE[FIELD], SPV[9][FIELD], SPCOEFF[9];
static uint64_t GFNI[FIELD];
static int tables_ready = 0;

/* gf2p8/src/field.rs:22-37 (Gf2p8::mul, shift-and-add modulo POLY=0x11d,
 * gf2p8/src/poly_11d.rs:18-21). */
uint8_t aro_mul(uint8_t x, uint8_t y) {
    uint16_t a = x, b = y, r = 0;
    for (int i = 0; i < 8; i++) {
        if (b & 1) r ^= a;
        a <<= 1;
        if (a & 0x100) a ^= 0x11d;
        b >>= 1;
    }
    return (uint8_t)r;
}

/* gf2p8/src/field.rs:122-130 (trace by repeated squaring). */
static int gf_trace(uint8_t x) {
    uint8_t t = x, s = x;
    for (int i = 0; i < 7; i++) { t = aro_mul(t, t); s ^= t; }
    return s & 1;
}

/* gf2p8/src/field.rs:134-147 (first root of x^2+x = c in ascending order). */
static int gf_solve_quadratic(uint8_t c, uint8_t *out) {
    if (gf_trace(c)) return 0;
    for (int i = 0; i < 256; i++) {
        uint8_t x = (uint8_t)i;
        if ((uint8_t)(aro_mul(x, x) ^ x) == c) { *out = x; return 1; }
    }
    return 0;
}

static void build_tables(void) {
    if (tables_ready) return;
    /* gf2p8/src/field.rs:68-88: EXP[i] = 2^i for i < 255, duplicated. */
    uint8_t x = 1;
    for (int i = 0; i < FIELD - 1; i++) {
        EXP[i] = x;
        LOG[x] = (uint8_t)i;
        x = aro_mul(x, 2);
    }
    for (int i = 0; i < FIELD - 1; i++) EXP[FIELD - 1 + i] = EXP[i];
    /* gf2p8/src/field.rs:90-102: INV[a] = EXP[255 - LOG[a]], INV[0] = 0. */
    INV[0] = 0;
    for (int i = 1; i < FIELD; i++) INV[i] = EXP[FIELD - 1 - LOG[i]];
    /* gf2p8/src/field.rs:174-198: Cantor basis v0 = 1, v_i^2 + v_i = v_{i-1},
     * choosing the trace-0 root. */
    uint8_t cur = 1;
    int len = 0;
    BASIS[len++] = cur;
    uint8_t nxt;
    while (len < 8 && gf_solve_quadratic(cur, &nxt)) {
        cur = gf_trace(nxt) ? (uint8_t)(nxt ^ 1) : nxt;
        BASIS[len++] = cur;
    }
    /* gf2p8/src/field.rs:218-235: omega_i = xor of v_b over set bits b of i. */
    for (int i = 0; i < FIELD; i++) {
        uint8_t p = 0;
        for (int b = 0; b < 8; b++) if ((i >> b) & 1) p ^= BASIS[b];
        SUBSPACE[i] = p;
    }
    /* gf2p8/src/field.rs:312-336: s_0(x) = x, s_{j+1} = s_j (s_j + s_j(v_j)). */
    for (int i = 0; i < FIELD; i++) SPV[0][i] = (uint8_t)i;
    for (int j = 0; j < 8; j++) {
        uint8_t bj = SPV[j][BASIS[j]];
        for (int i = 0; i < FIELD; i++) {
            uint8_t s = SPV[j][i];
            SPV[j + 1][i] = aro_mul(s, (uint8_t)(s ^ bj));
        }
    }
    /* gf2p8/src/field.rs:381-395: coefficient masks, s_{j+1} = s_j^2 + s_j. */
    uint16_t m[9];
    m[0] = 1;
    for (int j = 0; j < 8; j++) m[j + 1] = (uint16_t)((m[j] << 1) ^ m[j]);
    for (int j = 0; j < 9; j++) SPCOEFF[j] = (uint8_t)(m[j] >> 1);
    /* gf2p8/src/field.rs:150-161 + gf2p8/src/bit_matrix.rs:19-31, 80-82:
     * columns tau*2^i, transposed to rows, packed big-endian. */
    for (int t = 0; t < FIELD; t++) {
        uint8_t cols[8], rows[8];
        for (int i = 0; i < 8; i++) cols[i] = aro_mul((uint8_t)t, (uint8_t)(1u << i));
        for (int r = 0; r < 8; r++) {
            uint8_t v = 0;
            for (int c = 0; c < 8; c++) v |= (uint8_t)(((cols[c] >> r) & 1) << c);
            rows[r] = v;
        }
        uint64_t w = 0;
        for (int r = 0; r < 8; r++) w = (w << 8) | rows[r];
        GFNI[t] = w;
    }
    tables_ready = 1;
}

/* Export the generated tables (rs/build.rs:549-601 writes the same set). */
void aro_tables(uint8_t *exp, uint8_t *log, uint8_t *inv, uint8_t *basis,
                uint8_t *subspace, uint8_t *spv, uint8_t *spcoeff, uint64_t *gfni) {
    build_tables();
    if (exp) memcpy(exp, EXP, sizeof EXP);
    if (log) memcpy(log, LOG, sizeof LOG);
    if (inv) memcpy(inv, INV, sizeof INV);
    if (basis) memcpy(basis, BASIS, sizeof BASIS);
    if (subspace) memcpy(subspace, SUBSPACE, sizeof SUBSPACE);
    if (spv) memcpy(spv, SPV, sizeof SPV);
    if (spcoeff) memcpy(spcoeff, SPCOEFF, sizeof SPCOEFF);
    if (gfni) memcpy(gfni, GFNI, sizeof GFNI);
}

/* rs/src/poly_11d_lut.rs:10-17 (log/exp multiply with zero check). */
uint8_t aro_mul_lut(uint8_t a, uint8_t b) {
    build_tables();
    if (a == 0 || b == 0) return 0;
    return EXP[LOG[a] + LOG[b]];
}

/* rs/src/gf2p8lut.rs:28-30 (256-entry table for a fixed multiplier). */
static void make_mul_lut(uint8_t c, uint8_t lut[FIELD]) {
    for (int i = 0; i < FIELD; i++) lut[i] = aro_mul_lut(c, (uint8_t)i);
}

/* rs/src/kernel/lut_kernel.rs:11-18 */
static void butterfly_fwd(uint8_t *a, uint8_t *b, size_t L, const uint8_t *lut) {
    for (size_t j = 0; j < L; j++) {
        a[j] ^= lut[b[j]];
        b[j] ^= a[j];
    }
}

/* rs/src/kernel/lut_kernel.rs:20-26 */
static void butterfly_inv(uint8_t *a, uint8_t *b, size_t L, const uint8_t *lut) {
    for (size_t j = 0; j < L; j++) {
        b[j] ^= a[j];
        a[j] ^= lut[b[j]];
    }
}

/* rs/src/kernel/lut_kernel.rs:28-54 (LNH Algorithm 1, sharded, recursive). */
void aro_fft_sharded(uint8_t *shards, size_t L, int k, uint8_t beta) {
    build_tables();
    if (k == 0) return;
    size_t half = (size_t)1 << (k - 1);
    uint8_t lut[FIELD];
    make_mul_lut(SPV[k - 1][beta], lut);
    for (size_t i = 0; i < half; i++)
        butterfly_fwd(shards + i * L, shards + (i + half) * L, L, lut);
    uint8_t next_beta = beta ^ BASIS[k - 1];
    aro_fft_sharded(shards, L, k - 1, beta);
    aro_fft_sharded(shards + half * L, L, k - 1, next_beta);
}

/* rs/src/kernel/lut_kernel.rs:56-82 (LNH Algorithm 2, sharded, recursive). */
void aro_ifft_sharded(uint8_t *shards, size_t L, int k, uint8_t beta) {
    build_tables();
    if (k == 0) return;
    size_t half = (size_t)1 << (k - 1);
    uint8_t next_beta = beta ^ BASIS[k - 1];
    aro_ifft_sharded(shards, L, k - 1, beta);
    aro_ifft_sharded(shards + half * L, L, k - 1, next_beta);
    uint8_t lut[FIELD];
    make_mul_lut(SPV[k - 1][beta], lut);
    for (size_t i = 0; i < half; i++)
        butterfly_inv(shards + i * L, shards + (i + half) * L, L, lut);
}

/* rs/src/kernel/lut_kernel.rs:84-89 */
static void scale(const uint8_t *src, uint8_t *dst, size_t L, uint8_t c) {
    uint8_t lut[FIELD];
    make_mul_lut(c, lut);
    for (size_t j = 0; j < L; j++) dst[j] = lut[src[j]];
}

/* rs/src/kernel/lut_kernel.rs:91-96 */
static void scale_in_place(uint8_t *dst, size_t L, uint8_t c) {
    uint8_t lut[FIELD];
    make_mul_lut(c, lut);
    for (size_t j = 0; j < L; j++) dst[j] = lut[dst[j]];
}

static int is_pow2(size_t v) { return v && !(v & (v - 1)); }
static int ilog2(size_t v) { int r = 0; while ((size_t)1 << r < v) r++; return r; }

/* The reference's contract: N <= 256 power of two, 1 <= T < N power of two
 * (rs/src/lib.rs:16-19). */
static int valid_ref_code(size_t n, size_t T) {
    return n >= 2 && n <= 256 && is_pow2(n) && T >= 1 && T < n && is_pow2(T);
}

/* rs/src/codec.rs:241-265.  Parity is XOR-accumulated into the caller's
 * buffer, exactly as the reference does; callers pass zeroed parity. */
int aro_encode_systematic_sharded(size_t n, size_t T, const uint8_t *message,
                                  uint8_t *parity, uint8_t *workspace, size_t L) {
    build_tables();
    if (!valid_ref_code(n, T)) return -1;
    int t_log = ilog2(T);
    size_t k = n - T;
    for (size_t i = 0; i < k / T; i++) {
        memcpy(workspace, message + i * T * L, T * L);
        uint8_t omega = SUBSPACE[(i + 1) * T];
        aro_ifft_sharded(workspace, L, t_log, omega);
        for (size_t j = 0; j < T * L; j++) parity[j] ^= workspace[j];
    }
    aro_fft_sharded(parity, L, t_log, 0);
    return 0;
}

/* rs/src/codec.rs:436-465 (monomial-basis erasure locator and the products
 * of differences, O(e^2)). */
static void erasure_locator_and_denominators(const uint8_t *pos, size_t e,
                                             uint8_t *lambda, uint8_t *denoms) {
    uint8_t pts[256];
    for (size_t i = 0; i < e; i++) pts[i] = SUBSPACE[pos[i]];
    memset(lambda, 0, 257);
    lambda[0] = 1;
    for (size_t k = 0; k < e; k++) {
        uint8_t p = pts[k];
        for (size_t j = k + 1; j >= 1; j--) lambda[j] = lambda[j - 1] ^ aro_mul_lut(p, lambda[j]);
        lambda[0] = aro_mul_lut(p, lambda[0]);
    }
    for (size_t j = 0; j < e; j++) {
        uint8_t acc = 1;
        for (size_t i = 0; i < e; i++)
            if (i != j) acc = aro_mul_lut(acc, pts[j] ^ pts[i]);
        denoms[j] = acc;
    }
}

/* rs/src/codec.rs:479-569.  Returns 1 (true), 0 (false: e > T), or -1 for a
 * code the reference's contract excludes.  Out-of-range positions are
 * rejected with -2 (the reference panics on slicing, codec.rs:562-563). */
int aro_recover_erasure_shards(size_t n, size_t T, uint8_t *received, uint8_t *workspace,
                               size_t L, const uint8_t *pos, size_t e) {
    build_tables();
    if (!valid_ref_code(n, T)) return -1;
    if (e > T) return 0;
    if (e == 0) return 1;
    for (size_t i = 0; i < e; i++) if (pos[i] >= n) return -2;
    int t_log = ilog2(T), n_log = ilog2(n);
    uint8_t lambda[257], denoms[256];
    erasure_locator_and_denominators(pos, e, lambda, denoms);
    /* chunk 0 at omega_0 (codec.rs:503-511) */
    memcpy(workspace, received, T * L);
    aro_ifft_sharded(workspace, L, t_log, 0);
    /* chunks c >= 1 at omega_{cT}, XOR-accumulated (codec.rs:513-523) */
    for (size_t c = 1; c < n / T; c++) {
        uint8_t *tmp = workspace + T * L;
        memcpy(tmp, received + c * T * L, T * L);
        aro_ifft_sharded(tmp, L, t_log, SUBSPACE[c * T]);
        for (size_t j = 0; j < T * L; j++) workspace[j] ^= tmp[j];
    }
    memset(workspace + T * L, 0, (n - T) * L);
    /* Horner evaluation of lambda at all omega_i (codec.rs:527-536) */
    uint8_t lev[256];
    for (size_t i = 0; i < n; i++) {
        uint8_t p = SUBSPACE[i];
        uint8_t v = lambda[e];
        for (size_t j = e; j-- > 0;) v = aro_mul_lut(v, p) ^ lambda[j];
        lev[i] = v;
    }
    aro_fft_sharded(workspace, L, n_log, 0);                     /* codec.rs:539 */
    for (size_t i = 0; i < n; i++) scale_in_place(workspace + i * L, L, lev[i]); /* :542-547 */
    aro_ifft_sharded(workspace, L, n_log, 0);                    /* codec.rs:550 */
    memmove(workspace, workspace + T * L, e * L);                /* codec.rs:553 */
    memset(workspace + e * L, 0, (n - e) * L);                   /* codec.rs:554 */
    aro_fft_sharded(workspace, L, n_log, 0);                     /* codec.rs:557 */
    for (size_t i = 0; i < e; i++)                               /* codec.rs:560-566 */
        scale(workspace + (size_t)pos[i] * L, received + (size_t)pos[i] * L, L, INV[denoms[i]]);
    return 1;
}

/* rs/src/codec.rs:321-340 (n = 2T only). */
int aro_recompute_data_from_parity_sharded(size_t n, size_t T, uint8_t *received,
                                           uint8_t *workspace, size_t L) {
    build_tables();
    if (!valid_ref_code(n, T) || n != 2 * T) return -1;
    int t_log = ilog2(T);
    memcpy(workspace, received, T * L);
    aro_ifft_sharded(workspace, L, t_log, 0);
    aro_fft_sharded(workspace, L, t_log, SUBSPACE[T]);
    memcpy(received + T * L, workspace, T * L);
    return 0;
}

/* rs/src/poly_arith.rs:129-189: the scalar transforms are the sharded ones
 * at shard_len = 1 (the reference asserts this, rs/src/poly_11d_lut.rs:453-469). */
void aro_fft_scalar(uint8_t *c, int k, uint8_t beta) { aro_fft_sharded(c, 1, k, beta); }
void aro_ifft_scalar(uint8_t *c, int k, uint8_t beta) { aro_ifft_sharded(c, 1, k, beta); }

/* ---------------------------------------------------------------------------
 * Extension beyond the reference's contract (T not a power of two, e.g.
 * (n=256, k=16) where T = 240).  The reference cannot encode or recover such
 * codes (rs/src/codec.rs:252 takes trailing_zeros(T)).  The code itself is
 * still defined: the unique codeword (f(w_0), ..., f(w_{n-1})) with deg f < k
 * whose positions T..n-1 hold the message.  These two functions compute it by
 * a different route than the product path and are pinned only by the
 * Lagrange brute force in tests/test_oracle.py.
 * ------------------------------------------------------------------------- */

/* Encode when k is a power of two: f's novel-basis coefficients are
 * IFFT_k(message, w_{n-k}); parity block j is FFT_k(coeffs, w_{jk}). */
int aro_ext_encode_kpow2(size_t n, size_t k, const uint8_t *message, uint8_t *parity, size_t L) {
    build_tables();
    if (!(n >= 2 && n <= 256 && is_pow2(n) && k >= 1 && k < n && is_pow2(k))) return -1;
    int k_log = ilog2(k);
    uint8_t *d = (uint8_t *)malloc(k * L);
    if (!d) return -3;
    memcpy(d, message, k * L);
    aro_ifft_sharded(d, L, k_log, SUBSPACE[n - k]);
    for (size_t j = 0; j < n / k - 1; j++) {
        memcpy(parity + j * k * L, d, k * L);
        aro_fft_sharded(parity + j * k * L, L, k_log, SUBSPACE[j * k]);
    }
    free(d);
    return 0;
}

/* Erasure recovery by Lagrange interpolation on the survivors, column by
 * column: O(n * k^2) per column, for small test sizes only.  Output follows
 * the reference's convention for erased content g: received[p] = c_p ^ g_p. */
int aro_ext_recover_lagrange(size_t n, size_t k, uint8_t *received, size_t L,
                             const uint8_t *pos, size_t e) {
    build_tables();
    size_t T = n - k;
    if (!(n >= 2 && n <= 256 && is_pow2(n) && k >= 1 && k < n)) return -1;
    if (e > T) return 0;
    if (e == 0) return 1;
    uint8_t erased[256] = {0};
    for (size_t i = 0; i < e; i++) {
        if (pos[i] >= n || erased[pos[i]]) return -2;
        erased[pos[i]] = 1;
    }
    /* choose k survivors */
    uint8_t xs[256];
    size_t idx[256], m = 0;
    for (size_t i = 0; i < n && m < k; i++) if (!erased[i]) { idx[m] = i; xs[m] = SUBSPACE[i]; m++; }
    /* barycentric weights w_j = 1 / prod_{l != j} (x_j + x_l) */
    uint8_t w[256];
    for (size_t j = 0; j < k; j++) {
        uint8_t acc = 1;
        for (size_t l = 0; l < k; l++) if (l != j) acc = aro_mul_lut(acc, xs[j] ^ xs[l]);
        w[j] = INV[acc];
    }
    for (size_t q = 0; q < e; q++) {
        uint8_t x = SUBSPACE[pos[q]];
        /* coefficient of y_j: w_j * prod_{l != j} (x + x_l) */
        uint8_t coef[256];
        for (size_t j = 0; j < k; j++) {
            uint8_t acc = w[j];
            for (size_t l = 0; l < k; l++) if (l != j) acc = aro_mul_lut(acc, x ^ xs[l]);
            coef[j] = acc;
        }
        uint8_t *dst = received + (size_t)pos[q] * L;
        for (size_t c = 0; c < L; c++) {
            uint8_t v = 0;
            for (size_t j = 0; j < k; j++) v ^= aro_mul_lut(coef[j], received[idx[j] * L + c]);
            dst[c] ^= v;
        }
    }
    return 1;
}
