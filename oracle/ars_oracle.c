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

/* ---------------------------------------------------------------------------
 * Scalar error decoding: Codec::decode_systematic_scalar
 * (rs/src/codec.rs:341-420) with its syndrome (rs/src/codec.rs:266-301), its
 * key-equation solvers (HGCD rs/src/codec.rs:84-210, EEA rs/src/codec.rs:28-80)
 * and the novel-basis polynomial arithmetic they call
 * (rs/src/poly_arith.rs:56-116, 198-319).
 *
 * A polynomial is an array of NMAX coefficients in the basis X_j, of which the
 * first n are meaningful (the reference's [G; N]).  Where the reference would
 * panic (a slice or shift beyond N, a shift by a negative amount in its debug
 * build), the restatement returns ARO_PANIC up the call chain.
 */
#define NMAX 256
#define ARO_PANIC (-3)

static int pdeg(const uint8_t *p, size_t n) {
    for (size_t i = n; i-- > 0;)
        if (p[i]) return (int)i;
    return -1;
}

static void padd(uint8_t *dst, const uint8_t *a, const uint8_t *b, size_t n) {
    for (size_t i = 0; i < n; i++) dst[i] = a[i] ^ b[i];
}

/* rs/src/poly_arith.rs:198-231: product by FFT of the smallest power-of-two
 * size above deg a + deg b. */
static int poly_mul_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    int da = pdeg(a, n), db = pdeg(b, n);
    if (da < 0 || db < 0) {
        memset(out, 0, n);
        return 0;
    }
    size_t m = 1;
    while (m < (size_t)(da + db + 1)) m <<= 1;
    if (m > n) return ARO_PANIC; /* va[..m] with m > N */
    uint8_t va[NMAX], vb[NMAX];
    memcpy(va, a, n);
    memcpy(vb, b, n);
    int k = ilog2(m);
    aro_fft_scalar(va, k, 0);
    aro_fft_scalar(vb, k, 0);
    for (size_t i = 0; i < m; i++) va[i] = aro_mul(va[i], vb[i]);
    aro_ifft_scalar(va, k, 0);
    memcpy(out, va, n);
    return 0;
}

/* rs/src/poly_arith.rs:233-267: long division, one basis-aware product per
 * quotient term.  Returns 1 (Some), 0 (None: b = 0) or ARO_PANIC. */
static int poly_div_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *q, uint8_t *r) {
    int bd = pdeg(b, n);
    if (bd < 0) return 0;
    uint8_t rr[NMAX], qq[NMAX], term[NMAX], prod[NMAX];
    memcpy(rr, a, n);
    memset(qq, 0, n);
    const uint8_t binv = INV[b[bd]];
    for (;;) {
        int rd = pdeg(rr, n);
        if (rd < bd) break; /* also rd < 0 */
        int diff = rd - bd;
        uint8_t f = aro_mul(rr[rd], binv);
        qq[diff] ^= f;
        memset(term, 0, n);
        term[diff] = f;
        if (poly_mul_lnh(term, b, n, prod)) return ARO_PANIC;
        for (size_t i = 0; i < n; i++) rr[i] ^= prod[i];
    }
    memcpy(q, qq, n);
    memcpy(r, rr, n);
    return 1;
}

/* rs/src/poly_arith.rs:91-116 (poly::deriv_lnh), over all N coefficients. */
static void deriv_lnh(const uint8_t *c, size_t n, uint8_t *out) {
    memset(out, 0, n);
    if (n <= 1) return;
    memcpy(out, c, n);
    int m = ilog2(n);
    for (int j = 1; j <= m; j++) {
        size_t half = (size_t)1 << (j - 1), step = (size_t)1 << j;
        for (size_t s = 0; s < n; s += step)
            for (size_t i = 0; i < half; i++) out[s + i] ^= out[s + i + half];
    }
}

/* rs/src/poly_arith.rs:69-89 (poly::split_at, poly::shift_up). */
static int split_at(const uint8_t *p, int k, size_t n, uint8_t *lo, uint8_t *hi) {
    size_t pivot = (size_t)1 << k;
    if (pivot > n) return ARO_PANIC;
    uint8_t l[NMAX], h[NMAX];
    memset(l, 0, n);
    memset(h, 0, n);
    memcpy(l, p, pivot);
    memcpy(h, p + pivot, n - pivot);
    memcpy(lo, l, n);
    memcpy(hi, h, n);
    return 0;
}
static int shift_up(const uint8_t *p, int k, size_t n, uint8_t *out) {
    size_t shift = (size_t)1 << k;
    if (shift > n) return ARO_PANIC;
    uint8_t o[NMAX];
    memset(o, 0, n);
    memcpy(o + shift, p, n - shift);
    memcpy(out, o, n);
    return 0;
}

typedef struct {
    uint8_t m[4][NMAX]; /* row-major [m00, m01, m10, m11] */
} pmat;

/* rs/src/poly_arith.rs:292-319 (mat_vec_lnh, mat_mul_lnh). */
static int mat_vec(const pmat *M, const uint8_t *v0, const uint8_t *v1, size_t n, uint8_t *o0, uint8_t *o1) {
    uint8_t t0[NMAX], t1[NMAX], t2[NMAX], t3[NMAX];
    if (poly_mul_lnh(M->m[0], v0, n, t0) || poly_mul_lnh(M->m[1], v1, n, t1) ||
        poly_mul_lnh(M->m[2], v0, n, t2) || poly_mul_lnh(M->m[3], v1, n, t3))
        return ARO_PANIC;
    padd(o0, t0, t1, n);
    padd(o1, t2, t3, n);
    return 0;
}
static int mat_mul(const pmat *A, const pmat *B, size_t n, pmat *out) {
    pmat r;
    uint8_t t0[NMAX], t1[NMAX];
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            if (poly_mul_lnh(A->m[2 * i], B->m[j], n, t0) || poly_mul_lnh(A->m[2 * i + 1], B->m[2 + j], n, t1))
                return ARO_PANIC;
            padd(r.m[2 * i + j], t0, t1, n);
        }
    *out = r;
    return 0;
}

static void mat_identity(pmat *M, size_t n) {
    for (int i = 0; i < 4; i++) memset(M->m[i], 0, n);
    M->m[0][0] = 1;
    M->m[3][0] = 1;
}

/* rs/src/codec.rs:84-122 (poly_hgcd_middle). */
static int hgcd_middle(const uint8_t *p, int g, size_t n, uint8_t *p_ll, uint8_t *p_m) {
    if (g < 2) return ARO_PANIC; /* 1 << (g - 2) underflows */
    size_t q = (size_t)1 << (g - 2), h = (size_t)1 << (g - 1);
    if (q + 2 * h > n) return ARO_PANIC;
    const uint8_t c = SPV[g - 2][SUBSPACE[q]];
    uint8_t ll[NMAX], mm[NMAX];
    memset(ll, 0, n);
    memset(mm, 0, n);
    memcpy(ll, p, q);
    for (size_t i = 0; i < q; i++) {
        mm[i] = p[i + q] ^ aro_mul_lut(c, p[i + h]);
        mm[i + q] = p[i + h];
        mm[i + h] = p[i + h + q] ^ aro_mul_lut(c, p[i + 2 * h]);
        mm[i + h + q] = p[i + 2 * h];
    }
    memcpy(p_ll, ll, n);
    memcpy(p_m, mm, n);
    return 0;
}

/* rs/src/codec.rs:124-210 (hgcd, LNH Algorithm 5). */
static int hgcd(const uint8_t *a, const uint8_t *b, int g, size_t n, uint8_t *z0, uint8_t *z1, pmat *M) {
    if (g < 1) return ARO_PANIC; /* 1 << (g - 1) underflows */
    const int half = 1 << (g - 1);
    int bd = pdeg(b, n);
    if (bd < half) { /* lines 1-2 */
        memcpy(z0, a, n);
        memcpy(z1, b, n);
        mat_identity(M, n);
        return 0;
    }
    uint8_t a_l[NMAX], a_h[NMAX], b_l[NMAX], b_h[NMAX];
    if (split_at(a, g - 1, n, a_l, a_h) || split_at(b, g - 1, n, b_l, b_h)) return ARO_PANIC;
    uint8_t zh0[NMAX], zh1[NMAX];
    pmat mh;
    int rc = hgcd(a_h, b_h, g - 1, n, zh0, zh1, &mh);
    if (rc) return rc;
    uint8_t mv0[NMAX], mv1[NMAX], zm0[NMAX], zm1[NMAX], t[NMAX];
    if (mat_vec(&mh, a_l, b_l, n, mv0, mv1)) return ARO_PANIC;
    if (shift_up(zh0, g - 1, n, t)) return ARO_PANIC;
    padd(zm0, t, mv0, n);
    if (shift_up(zh1, g - 1, n, t)) return ARO_PANIC;
    padd(zm1, t, mv1, n);
    if (pdeg(zm1, n) < half) { /* lines 5-6 */
        memcpy(z0, zm0, n);
        memcpy(z1, zm1, n);
        *M = mh;
        return 0;
    }
    uint8_t qm[NMAX], rm[NMAX];
    rc = poly_div_lnh(zm0, zm1, n, qm, rm);
    if (rc != 1) return ARO_PANIC; /* .expect() */
    uint8_t zm1_ll[NMAX], zm1_m[NMAX], rm_ll[NMAX], rm_m[NMAX];
    if (hgcd_middle(zm1, g, n, zm1_ll, zm1_m) || hgcd_middle(rm, g, n, rm_ll, rm_m)) return ARO_PANIC;
    uint8_t ym0[NMAX], ym1[NMAX];
    pmat mm;
    rc = hgcd(zm1_m, rm_m, g - 1, n, ym0, ym1, &mm);
    if (rc) return rc;
    pmat swap, tmp, mr;
    for (int i = 0; i < 4; i++) memset(swap.m[i], 0, n);
    swap.m[1][0] = 1;
    swap.m[2][0] = 1;
    memcpy(swap.m[3], qm, n);
    if (mat_mul(&swap, &mh, n, &tmp) || mat_mul(&mm, &tmp, n, &mr)) return ARO_PANIC;
    uint8_t sg2[NMAX];
    memset(sg2, 0, n);
    if (((size_t)1 << (g - 2)) >= n) return ARO_PANIC;
    sg2[(size_t)1 << (g - 2)] = 1;
    if (mat_vec(&mm, zm1_ll, rm_ll, n, mv0, mv1)) return ARO_PANIC;
    uint8_t zr0[NMAX], zr1[NMAX];
    if (poly_mul_lnh(ym0, sg2, n, t)) return ARO_PANIC;
    padd(zr0, t, mv0, n);
    if (shift_up(ym1, g - 2, n, t)) return ARO_PANIC;
    padd(zr1, t, mv1, n);
    memcpy(z0, zr0, n);
    memcpy(z1, zr1, n);
    *M = mr;
    return 0;
}

/* rs/src/codec.rs:212-239 (solve_key_equation_hgcd).  Returns 1 (Some),
 * 0 (None) or ARO_PANIC; v1 = m11, lambda = m10 + m11 q_t. */
static int key_equation_hgcd(const uint8_t *syn, size_t n, size_t T, uint8_t *v1, uint8_t *lambda) {
    uint8_t st[NMAX], qt[NMAX], rt[NMAX], z0[NMAX], z1[NMAX], t[NMAX];
    memset(st, 0, n);
    st[T] = 1;
    int rc = poly_div_lnh(st, syn, n, qt, rt);
    if (rc != 1) return rc;
    pmat M;
    rc = hgcd(syn, rt, ilog2(T), n, z0, z1, &M);
    if (rc) return rc;
    memcpy(v1, M.m[3], n);
    if (poly_mul_lnh(M.m[3], qt, n, t)) return ARO_PANIC;
    padd(lambda, M.m[2], t, n);
    return 1;
}

/* rs/src/codec.rs:28-80 (solve_key_equation_eea and eea). */
static int key_equation_eea(const uint8_t *syn, size_t n, size_t T, uint8_t *v1_out, uint8_t *lambda) {
    uint8_t st[NMAX], qt[NMAX], rt[NMAX];
    memset(st, 0, n);
    st[T] = 1;
    int rc = poly_div_lnh(st, syn, n, qt, rt);
    if (rc != 1) return rc;
    uint8_t z0[NMAX], z1[NMAX], u0[NMAX], u1[NMAX], v0[NMAX], v1[NMAX];
    uint8_t q[NMAX], rem[NMAX], t[NMAX], nx[NMAX];
    memcpy(z0, syn, n);
    memcpy(z1, rt, n);
    memset(u0, 0, n);
    memset(u1, 0, n);
    memset(v0, 0, n);
    memset(v1, 0, n);
    u0[0] = 1;
    v1[0] = 1;
    const int target = (int)(T / 2);
    while (pdeg(z1, n) >= target) {
        if (poly_div_lnh(z0, z1, n, q, rem) != 1) return ARO_PANIC;
        memcpy(z0, z1, n);
        memcpy(z1, rem, n);
        if (poly_mul_lnh(q, u1, n, t)) return ARO_PANIC;
        padd(nx, u0, t, n);
        memcpy(u0, u1, n);
        memcpy(u1, nx, n);
        if (poly_mul_lnh(q, v1, n, t)) return ARO_PANIC;
        padd(nx, v0, t, n);
        memcpy(v0, v1, n);
        memcpy(v1, nx, n);
    }
    memcpy(v1_out, v1, n);
    if (poly_mul_lnh(v1, qt, n, t)) return ARO_PANIC;
    padd(lambda, u1, t, n);
    return 1;
}

/* rs/src/codec.rs:266-301: s = sum_i IFFT_T(chunk i, omega_{iT}), N entries
 * of which the first T are set. */
static void syndrome_scalar(const uint8_t *r, size_t n, size_t T, uint8_t *syn) {
    memset(syn, 0, n);
    uint8_t ws[NMAX];
    for (size_t i = 0; i < n / T; i++) {
        memcpy(ws, r + i * T, T);
        aro_ifft_scalar(ws, ilog2(T), SUBSPACE[i * T]);
        for (size_t j = 0; j < T; j++) syn[j] ^= ws[j];
    }
}

/* The formal derivative in the basis X (an extension, not the reference's
 * operator): every s_b has x-coefficient 1 (gf2p8/src/poly_11d.rs:204-213),
 * so X_j' = sum_{b in j} X_{j - 2^b} and D'[i] = xor_{b not in i} D[i | 2^b]. */
static void formal_deriv(const uint8_t *c, size_t n, uint8_t *out) {
    const int nb = ilog2(n);
    for (size_t i = 0; i < n; i++) {
        uint8_t v = 0;
        for (int b = 0; b < nb; b++)
            if (!((i >> b) & 1)) v ^= c[i | ((size_t)1 << b)];
        out[i] = v;
    }
}

/* rs/src/codec.rs:341-420.  Returns 1 (true), 0 (false), -1 for a code
 * outside the reference's contract, or ARO_PANIC.  *kind: 0 zero syndrome,
 * 1 true after the key equation, 2 false, 3 panic.  formal != 0 replaces the
 * reference's deriv_lnh by the formal derivative in step 4 (extension). */
int aro_decode_systematic_scalar(size_t n, size_t T, uint8_t *received, int *kind, int formal) {
    build_tables();
    int dummy;
    if (!kind) kind = &dummy;
    if (!valid_ref_code(n, T)) return -1;
    const int t_log = ilog2(T);
    uint8_t syn[NMAX];
    syndrome_scalar(received, n, T, syn);
    int zero = 1;
    for (size_t j = 0; j < T; j++) zero &= syn[j] == 0;
    if (zero) {
        *kind = 0;
        return 1;
    }
    uint8_t v1[NMAX], lambda[NMAX];
    int rc = key_equation_hgcd(syn, n, T, v1, lambda);
    if (rc == ARO_PANIC) {
        *kind = 3;
        return ARO_PANIC;
    }
    if (rc == 0) {
        *kind = 2;
        return 0;
    }
    *kind = 1;
    int dl = pdeg(lambda, n);
    if (dl < 0) return 1; /* zero locator */
    /* step 3: root search, one T-point FFT per chunk, first dl roots */
    size_t idx[NMAX];
    int cnt = 0;
    uint8_t ev[NMAX];
    for (size_t c = 0; c < n / T && cnt < dl; c++) {
        memcpy(ev, lambda, n);
        aro_fft_scalar(ev, t_log, SUBSPACE[c * T]);
        for (size_t off = 0; off < T; off++)
            if (ev[off] == 0) {
                idx[cnt++] = c * T + off;
                if (cnt == dl) break;
            }
    }
    if (cnt < dl) {
        *kind = 2;
        return 0;
    }
    /* step 4: error values q(w_p) / lambda'(w_p), lambda' = deriv_lnh */
    uint8_t lp[NMAX], qe[NMAX], le[NMAX];
    if (formal)
        formal_deriv(lambda, n, lp);
    else
        deriv_lnh(lambda, n, lp);
    for (size_t c = 0; c < n / T; c++) {
        int any = 0;
        for (int i = 0; i < cnt; i++) any |= idx[i] / T == c;
        if (!any) continue;
        memcpy(qe, v1, n);
        memcpy(le, lp, n);
        aro_fft_scalar(qe, t_log, SUBSPACE[c * T]);
        aro_fft_scalar(le, t_log, SUBSPACE[c * T]);
        for (int i = 0; i < cnt; i++) {
            if (idx[i] / T != c) continue;
            size_t off = idx[i] % T;
            if (le[off] == 0) continue;
            received[idx[i]] ^= aro_mul_lut(qe[off], INV[le[off]]);
        }
    }
    return 1;
}

/* Sharded form (no reference counterpart; README.md:58 lists sharded error
 * correction as not implemented): byte column j of the n x L codeword is the
 * received word decode_systematic_scalar sees.  status[j] = its *kind.
 * Returns 1 when every column returned true, 0 when one returned false,
 * ARO_PANIC when one would have panicked, -1 for a bad code. */
int aro_decode_sharded(size_t n, size_t T, uint8_t *received, size_t L, uint8_t *status, int formal) {
    build_tables();
    if (!valid_ref_code(n, T)) return -1;
    int all = 1, panic = 0;
    uint8_t col[NMAX];
    for (size_t j = 0; j < L; j++) {
        for (size_t i = 0; i < n; i++) col[i] = received[i * L + j];
        int kind;
        int rc = aro_decode_systematic_scalar(n, T, col, &kind, formal);
        if (status) status[j] = (uint8_t)kind;
        if (rc == 1)
            for (size_t i = 0; i < n; i++) received[i * L + j] = col[i];
        else if (rc == 0)
            all = 0;
        else
            panic = 1;
    }
    return panic ? ARO_PANIC : all;
}

/* Exported pieces, for pinning each step on its own. */
int aro_poly_mul_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *out) {
    build_tables();
    return poly_mul_lnh(a, b, n, out);
}
int aro_poly_div_lnh(const uint8_t *a, const uint8_t *b, size_t n, uint8_t *q, uint8_t *r) {
    build_tables();
    return poly_div_lnh(a, b, n, q, r);
}
void aro_deriv_lnh(const uint8_t *c, size_t n, uint8_t *out) { deriv_lnh(c, n, out); }
void aro_formal_deriv(const uint8_t *c, size_t n, uint8_t *out) { formal_deriv(c, n, out); }
void aro_syndrome_scalar(const uint8_t *r, size_t n, size_t T, uint8_t *syn) {
    build_tables();
    syndrome_scalar(r, n, T, syn);
}
int aro_key_equation(const uint8_t *syn, size_t n, size_t T, int use_eea, uint8_t *v1, uint8_t *lambda) {
    build_tables();
    return use_eea ? key_equation_eea(syn, n, T, v1, lambda) : key_equation_hgcd(syn, n, T, v1, lambda);
}
