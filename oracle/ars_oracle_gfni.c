/*
 * ars_oracle_gfni.c -- CPU ORACLE, AVX-512BW + GFNI restatement of the
 * reference's native SIMD kernel (rs/src/kernel/gfni_kernel.rs).
 *
 * TEST INFRASTRUCTURE ONLY.  Used (a) as the CPU baseline that bench.py times
 * beside the GPU path, and (b) in tests, where it must equal the LUT
 * restatement in ars_oracle.c byte for byte (the reference's own check,
 * rs/src/kernel/gfni_kernel.rs:324-387).  It is never linked into the product.
 *
 * The codec glue (which transforms run, in which order) is the same as in
 * ars_oracle.c (rs/src/codec.rs:241-265, 479-569); only the butterfly and
 * scale loops change.  Zero twiddles take the XOR-only path that the
 * reference's generated unrolled kernels take (rs/build.rs:36-46, 86-96).
 */
#define _POSIX_C_SOURCE 199309L /* clock_gettime for the timing helper */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <immintrin.h>

void aro_tables(uint8_t *exp, uint8_t *log, uint8_t *inv, uint8_t *basis,
                uint8_t *subspace, uint8_t *spv, uint8_t *spcoeff, uint64_t *gfni);
uint8_t aro_mul_lut(uint8_t a, uint8_t b);

static uint8_t G_BASIS[8], G_SUBSPACE[256], G_SPV[9][256], G_INV[256];
static uint64_t G_GFNI[256];
static int g_ready = 0;

static void g_init(void) {
    if (g_ready) return;
    aro_tables(NULL, NULL, G_INV, G_BASIS, G_SUBSPACE, &G_SPV[0][0], NULL, G_GFNI);
    g_ready = 1;
}

int aro_gfni_supported(void) {
    __builtin_cpu_init();
    return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
           __builtin_cpu_supports("gfni");
}

#define TGT __attribute__((target("avx512f,avx512bw,gfni")))

/* rs/src/kernel/gfni_kernel.rs:14-44: (a, b) <- (a + t*b, b + a + t*b). */
TGT static void bfly_fwd(uint8_t *a, uint8_t *b, size_t len, __m512i mat, int zero) {
    size_t i = 0;
    for (; i + 64 <= len; i += 64) {
        __m512i va = _mm512_loadu_si512((const void *)(a + i));
        __m512i vb = _mm512_loadu_si512((const void *)(b + i));
        if (!zero) va = _mm512_xor_si512(va, _mm512_gf2p8affine_epi64_epi8(vb, mat, 0));
        vb = _mm512_xor_si512(vb, va);
        if (!zero) _mm512_storeu_si512((void *)(a + i), va);
        _mm512_storeu_si512((void *)(b + i), vb);
    }
    if (i < len) {
        __mmask64 k = (1ULL << (len - i)) - 1;
        __m512i va = _mm512_maskz_loadu_epi8(k, a + i);
        __m512i vb = _mm512_maskz_loadu_epi8(k, b + i);
        if (!zero) va = _mm512_xor_si512(va, _mm512_gf2p8affine_epi64_epi8(vb, mat, 0));
        vb = _mm512_xor_si512(vb, va);
        _mm512_mask_storeu_epi8(a + i, k, va);
        _mm512_mask_storeu_epi8(b + i, k, vb);
    }
}

/* rs/src/kernel/gfni_kernel.rs:46-76: (g0, g1) <- (g0 + t*(g0+g1), g0+g1). */
TGT static void bfly_inv(uint8_t *a, uint8_t *b, size_t len, __m512i mat, int zero) {
    size_t i = 0;
    for (; i + 64 <= len; i += 64) {
        __m512i va = _mm512_loadu_si512((const void *)(a + i));
        __m512i vb = _mm512_loadu_si512((const void *)(b + i));
        vb = _mm512_xor_si512(vb, va);
        if (!zero) va = _mm512_xor_si512(va, _mm512_gf2p8affine_epi64_epi8(vb, mat, 0));
        if (!zero) _mm512_storeu_si512((void *)(a + i), va);
        _mm512_storeu_si512((void *)(b + i), vb);
    }
    if (i < len) {
        __mmask64 k = (1ULL << (len - i)) - 1;
        __m512i va = _mm512_maskz_loadu_epi8(k, a + i);
        __m512i vb = _mm512_maskz_loadu_epi8(k, b + i);
        vb = _mm512_xor_si512(vb, va);
        if (!zero) va = _mm512_xor_si512(va, _mm512_gf2p8affine_epi64_epi8(vb, mat, 0));
        _mm512_mask_storeu_epi8(a + i, k, va);
        _mm512_mask_storeu_epi8(b + i, k, vb);
    }
}

/* rs/src/kernel/gfni_kernel.rs:78-107 (the generic recursion) */
TGT static void fft_rec(uint8_t *shards, size_t L, int k, uint8_t beta) {
    if (k == 0) return;
    size_t half = (size_t)1 << (k - 1);
    uint8_t tw = G_SPV[k - 1][beta];
    __m512i mat = _mm512_set1_epi64((long long)G_GFNI[tw]);
    for (size_t i = 0; i < half; i++) bfly_fwd(shards + i * L, shards + (i + half) * L, L, mat, tw == 0);
    fft_rec(shards, L, k - 1, beta);
    fft_rec(shards + half * L, L, k - 1, beta ^ G_BASIS[k - 1]);
}

/* rs/src/kernel/gfni_kernel.rs:109-150 */
TGT static void ifft_rec(uint8_t *shards, size_t L, int k, uint8_t beta) {
    if (k == 0) return;
    size_t half = (size_t)1 << (k - 1);
    ifft_rec(shards, L, k - 1, beta);
    ifft_rec(shards + half * L, L, k - 1, beta ^ G_BASIS[k - 1]);
    uint8_t tw = G_SPV[k - 1][beta];
    __m512i mat = _mm512_set1_epi64((long long)G_GFNI[tw]);
    for (size_t i = 0; i < half; i++) bfly_inv(shards + i * L, shards + (i + half) * L, L, mat, tw == 0);
}

/* The reference's generated kernels for beta = 0 (rs/build.rs:386-436,
 * 474-527; dispatched at rs/src/kernel/gfni_kernel.rs:200-240): the same
 * butterflies in the same order as the recursion -- preorder for the FFT,
 * postorder for the IFFT -- as one straight sequence with each butterfly's
 * matrix fixed in advance (there, compile-time constants).  Here the
 * sequence for each size is built once: (first shard, half width, matrix,
 * zero twiddle). */
typedef struct {
    uint32_t off, half;
    uint64_t mat;
    int zero;
} B0Step;
static B0Step B0_FFT[9][255], B0_IFFT[9][255];
static int B0_LEN[9], b0_ready = 0, g_unrolled = 1;

static void b0_build(B0Step *seq, int *len, int k, uint8_t beta, uint32_t off, int pre) {
    if (k == 0) return;
    uint32_t half = 1u << (k - 1);
    uint8_t tw = G_SPV[k - 1][beta];
    B0Step st = {off, half, G_GFNI[tw], tw == 0};
    if (pre) seq[(*len)++] = st;
    b0_build(seq, len, k - 1, beta, off, pre);
    b0_build(seq, len, k - 1, beta ^ G_BASIS[k - 1], off + half, pre);
    if (!pre) seq[(*len)++] = st;
}

static void b0_init(void) {
    if (b0_ready) return;
    for (int k = 1; k <= 8; k++) {
        int a = 0, b = 0;
        b0_build(B0_FFT[k], &a, k, 0, 0, 1);
        b0_build(B0_IFFT[k], &b, k, 0, 0, 0);
        B0_LEN[k] = a;
    }
    b0_ready = 1;
}

/* 1 (default): beta = 0 transforms take the fixed sequences above, as the
 * reference's GfniKernel dispatches them; 0: the generic recursion. */
void aro_gfni_set_unrolled(int on) { g_unrolled = on != 0; }

TGT void aro_gfni_fft_sharded(uint8_t *shards, size_t L, int k, uint8_t beta) {
    g_init();
    if (k == 0) return;
    if (beta == 0 && g_unrolled && k <= 8) {
        b0_init();
        for (int i = 0; i < B0_LEN[k]; i++) {
            const B0Step st = B0_FFT[k][i];
            __m512i mat = _mm512_set1_epi64((long long)st.mat);
            uint8_t *blk = shards + (size_t)st.off * L;
            for (size_t j = 0; j < st.half; j++) bfly_fwd(blk + j * L, blk + (j + st.half) * L, L, mat, st.zero);
        }
        return;
    }
    fft_rec(shards, L, k, beta);
}

TGT void aro_gfni_ifft_sharded(uint8_t *shards, size_t L, int k, uint8_t beta) {
    g_init();
    if (k == 0) return;
    if (beta == 0 && g_unrolled && k <= 8) {
        b0_init();
        for (int i = 0; i < B0_LEN[k]; i++) {
            const B0Step st = B0_IFFT[k][i];
            __m512i mat = _mm512_set1_epi64((long long)st.mat);
            uint8_t *blk = shards + (size_t)st.off * L;
            for (size_t j = 0; j < st.half; j++) bfly_inv(blk + j * L, blk + (j + st.half) * L, L, mat, st.zero);
        }
        return;
    }
    ifft_rec(shards, L, k, beta);
}

/* rs/src/kernel/gfni_kernel.rs:152-173 */
TGT static void scale_gfni(const uint8_t *src, uint8_t *dst, size_t len, uint8_t c) {
    __m512i mat = _mm512_set1_epi64((long long)G_GFNI[c]);
    size_t i = 0;
    for (; i + 64 <= len; i += 64)
        _mm512_storeu_si512((void *)(dst + i),
                            _mm512_gf2p8affine_epi64_epi8(_mm512_loadu_si512((const void *)(src + i)), mat, 0));
    if (i < len) {
        __mmask64 k = (1ULL << (len - i)) - 1;
        _mm512_mask_storeu_epi8(dst + i, k,
                                _mm512_gf2p8affine_epi64_epi8(_mm512_maskz_loadu_epi8(k, src + i), mat, 0));
    }
}

static int is_pow2(size_t v) { return v && !(v & (v - 1)); }
static int ilog2(size_t v) { int r = 0; while ((size_t)1 << r < v) r++; return r; }

/* rs/src/codec.rs:241-265 over the GFNI kernel. */
TGT int aro_gfni_encode_systematic_sharded(size_t n, size_t T, const uint8_t *message,
                                           uint8_t *parity, uint8_t *workspace, size_t L) {
    g_init();
    if (!(n >= 2 && n <= 256 && is_pow2(n) && T >= 1 && T < n && is_pow2(T))) return -1;
    int t_log = ilog2(T);
    size_t k = n - T;
    for (size_t i = 0; i < k / T; i++) {
        memcpy(workspace, message + i * T * L, T * L);
        aro_gfni_ifft_sharded(workspace, L, t_log, G_SUBSPACE[(i + 1) * T]);
        size_t j = 0, len = T * L;
        for (; j + 64 <= len; j += 64)
            _mm512_storeu_si512((void *)(parity + j),
                                _mm512_xor_si512(_mm512_loadu_si512((const void *)(parity + j)),
                                                 _mm512_loadu_si512((const void *)(workspace + j))));
        for (; j < len; j++) parity[j] ^= workspace[j];
    }
    aro_gfni_fft_sharded(parity, L, t_log, 0);
    return 0;
}

/* rs/src/codec.rs:479-569 over the GFNI kernel (lambda / denominators as in
 * rs/src/codec.rs:436-465). */
TGT int aro_gfni_recover_erasure_shards(size_t n, size_t T, uint8_t *received, uint8_t *workspace,
                                        size_t L, const uint8_t *pos, size_t e) {
    g_init();
    if (!(n >= 2 && n <= 256 && is_pow2(n) && T >= 1 && T < n && is_pow2(T))) return -1;
    if (e > T) return 0;
    if (e == 0) return 1;
    for (size_t i = 0; i < e; i++) if (pos[i] >= n) return -2;
    int t_log = ilog2(T), n_log = ilog2(n);
    uint8_t pts[256], lambda[257], denoms[256];
    for (size_t i = 0; i < e; i++) pts[i] = G_SUBSPACE[pos[i]];
    memset(lambda, 0, sizeof lambda);
    lambda[0] = 1;
    for (size_t k = 0; k < e; k++) {
        for (size_t j = k + 1; j >= 1; j--) lambda[j] = lambda[j - 1] ^ aro_mul_lut(pts[k], lambda[j]);
        lambda[0] = aro_mul_lut(pts[k], lambda[0]);
    }
    for (size_t j = 0; j < e; j++) {
        uint8_t acc = 1;
        for (size_t i = 0; i < e; i++) if (i != j) acc = aro_mul_lut(acc, pts[j] ^ pts[i]);
        denoms[j] = acc;
    }
    memcpy(workspace, received, T * L);
    aro_gfni_ifft_sharded(workspace, L, t_log, 0);
    for (size_t c = 1; c < n / T; c++) {
        uint8_t *tmp = workspace + T * L;
        memcpy(tmp, received + c * T * L, T * L);
        aro_gfni_ifft_sharded(tmp, L, t_log, G_SUBSPACE[c * T]);
        for (size_t j = 0; j < T * L; j++) workspace[j] ^= tmp[j];
    }
    memset(workspace + T * L, 0, (n - T) * L);
    uint8_t lev[256];
    for (size_t i = 0; i < n; i++) {
        uint8_t p = G_SUBSPACE[i], v = lambda[e];
        for (size_t j = e; j-- > 0;) v = aro_mul_lut(v, p) ^ lambda[j];
        lev[i] = v;
    }
    aro_gfni_fft_sharded(workspace, L, n_log, 0);
    for (size_t i = 0; i < n; i++) scale_gfni(workspace + i * L, workspace + i * L, L, lev[i]);
    aro_gfni_ifft_sharded(workspace, L, n_log, 0);
    memmove(workspace, workspace + T * L, e * L);
    memset(workspace + e * L, 0, (n - e) * L);
    aro_gfni_fft_sharded(workspace, L, n_log, 0);
    for (size_t i = 0; i < e; i++)
        scale_gfni(workspace + (size_t)pos[i] * L, received + (size_t)pos[i] * L, L, G_INV[denoms[i]]);
    return 1;
}

/* Timing helper for the CPU baseline (not part of the restatement): runs
 * encode (zeroed parity) and recover (erased shards zeroed) of one seeded
 * (n, T, L) codeword `reps` times on buffers allocated and touched once, as
 * the reference's criterion benches do (rs/benches/), and writes the
 * median nanoseconds of each to out_ns[0], out_ns[1].  Returns 0, or -1 on
 * bad arguments or allocation failure. */
#include <stdlib.h>
#include <time.h>

static double now_ns(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e9 + (double)ts.tv_nsec;
}
static int cmp_d(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int aro_gfni_time_codec(size_t n, size_t T, size_t L, const uint8_t *pos, size_t e, int reps, size_t offset,
                        double *out_ns) {
    if (!aro_gfni_supported() || reps < 1 || n < 2 || n > 256 || T >= n || e > T || offset >= 64) return -1;
    g_init();
    size_t k = n - T;
    /* message, parity and codeword start `offset` bytes past a 64-byte
     * boundary (the reference's benches time offsets 0 and 1,
     * rs/benches/encode_systematic_sharded.rs:51-73) */
    uint8_t *mb = aligned_alloc(64, k * L + 128), *pb = aligned_alloc(64, T * L + 128),
            *rb = aligned_alloc(64, n * L + 128), *ws = aligned_alloc(64, n * L + 64);
    uint8_t *msg = mb ? mb + offset : NULL, *par = pb ? pb + offset : NULL, *rx = rb ? rb + offset : NULL;
    double *te = malloc(sizeof(double) * reps), *tr = malloc(sizeof(double) * reps);
    int rc = -1;
    if (!msg || !par || !rx || !ws || !te || !tr) goto done;
    uint64_t s = 0x9e3779b97f4a7c15ull;
    for (size_t i = 0; i < k * L; i++) {
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        msg[i] = (uint8_t)s;
    }
    memset(ws, 1, n * L);
    memset(rx, 1, n * L);
    for (int r = 0; r < reps; r++) {
        memset(par, 0, T * L);
        double a = now_ns();
        if (aro_gfni_encode_systematic_sharded(n, T, msg, par, ws, L)) goto done;
        double b = now_ns();
        memcpy(rx, par, T * L);
        memcpy(rx + T * L, msg, k * L);
        for (size_t i = 0; i < e; i++) memset(rx + (size_t)pos[i] * L, 0, L);
        double c = now_ns();
        if (aro_gfni_recover_erasure_shards(n, T, rx, ws, L, pos, e) != 1) goto done;
        double d = now_ns();
        te[r] = b - a;
        tr[r] = d - c;
    }
    qsort(te, reps, sizeof(double), cmp_d);
    qsort(tr, reps, sizeof(double), cmp_d);
    out_ns[0] = te[reps / 2];
    out_ns[1] = tr[reps / 2];
    rc = 0;
done:
    free(mb); free(pb); free(rb); free(ws); free(te); free(tr);
    return rc;
}
