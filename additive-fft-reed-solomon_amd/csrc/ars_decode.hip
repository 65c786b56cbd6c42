// ars_decode.hip -- device side of sharded error decoding.
//
// Every byte column of an n x L codeword is one received word, decoded as
// Codec::decode_systematic_scalar decodes it (rs/src/codec.rs:341-420).  The
// device does the FFT-heavy steps; the host solves the key equation of the
// few columns that need it (ars_keyeq.hpp):
//   k_syndrome_*   s = sum_c IFFT_T(chunk c, omega_{cT}) per column
//                  (rs/src/codec.rs:266-301), in the encode tile layout; a
//                  column whose syndrome is not zero is appended to a list
//                  with its T syndrome bytes.  Clean columns write nothing,
//                  so a scrub of clean data only reads the codeword.
//   k_decode_columns  one wave per listed column: its key equation (the
//                  same monomial-basis Euclid as the host solver, ars_keyeq.hpp),
//                  then the root search (one T-point FFT of lambda per chunk,
//                  rs/src/codec.rs:383-396, the first deg(lambda) roots in index
//                  order) and the error values v(w_p) / lambda'(w_p)
//                  (rs/src/codec.rs:405-431), applied in place when deg(lambda)
//                  roots were found.
//   k_decode_tail  the root-search half alone, for plans the host solved.
#include <hip/hip_runtime.h>
#include <cstdint>

#include "ars_decode.hpp"
#include "ars_device.hpp"
#include "ars_tiles.hpp"

namespace ars {

namespace {

__device__ __forceinline__ void stage_tw(MulTab *s_tw, const MulTab *__restrict__ g_tw) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(g_tw);
    uint32_t *dst = reinterpret_cast<uint32_t *>(s_tw);
    for (uint32_t i = threadIdx.x; i < (uint32_t)kTwiddleEntries * 6u; i += blockDim.x) dst[i] = src[i];
}

// Byte mask of the quad's listed columns: valid (b < nb) and non-zero.
__device__ __forceinline__ uint32_t nonzero_bytes(uint32_t orw, uint32_t nb) {
    uint32_t m = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if ((uint32_t)b < nb && ((orw >> (8 * b)) & 0xffu)) m |= 1u << b;
    return m;
}

// Appends the columns of `mask` (quad starting at col0); returns their first
// list index.
__device__ __forceinline__ uint32_t list_columns(const SynOut &so, uint32_t mask, uint32_t col0) {
    const uint32_t base = atomicAdd(so.count, (uint32_t)__popc(mask));
    uint32_t k = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (mask & (1u << b)) so.cols[base + k++] = col0 + (uint32_t)b;
    return base;
}

template <int TL, bool AL>
__global__ void __launch_bounds__(256) k_syndrome_small(const uint8_t *__restrict__ rx, uint32_t L,
                                                        uint32_t chunks, const MulTab *__restrict__ g_tw,
                                                        SynOut so) {
    constexpr int T = 1 << TL;
    __shared__ MulTab tw[kTwiddleEntries];
    stage_tw(tw, g_tw);
    __syncthreads();
    const uint32_t quad = blockIdx.x * blockDim.x + threadIdx.x;
    BufIO<AL> io = make_bufio<AL>(rx, chunks * T * L, const_cast<uint8_t *>(rx), 0u, quad, L);
    uint32_t acc[T], x[T];
#pragma unroll
    for (int j = 0; j < T; ++j) acc[j] = 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
#pragma unroll
        for (int j = 0; j < T; ++j) x[j] = io.ld_in(c * T, j);
        run_levels<TL, 0, 0, TL, true, false>(x, c * T, tw);
#pragma unroll
        for (int j = 0; j < T; ++j) acc[j] ^= x[j];
    }
    uint32_t orw = 0;
#pragma unroll
    for (int j = 0; j < T; ++j) orw |= acc[j];
    const uint32_t mask = nonzero_bytes(orw, io.nb);
    if (mask) {
        uint32_t k = list_columns(so, mask, quad * 4u);
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!(mask & (1u << b))) continue;
            uint8_t *dst = so.syn + (size_t)k++ * T;
#pragma unroll
            for (int j = 0; j < T; ++j) dst[j] = (uint8_t)(acc[j] >> (8 * b));
        }
    }
}

// Tile variant for T >= 32 (see ars_tiles.hpp): each chunk's IFFT runs levels
// 0..3 in registers, one LDS exchange, levels 4..TL-1 in registers; the
// accumulated syndrome stays in the phase-B layout (lane q holds rows
// h*16 + q*GB + m).  The lanes of a quad OR their rows into LDS, lane q = 0
// lists the quad's non-zero columns, and every lane then writes its rows of
// those columns.
template <int TL, int PQ, int NBUF, bool AL>
__global__ void __launch_bounds__((Tile<TL, PQ>::THREADS))
    k_syndrome_tile(const uint8_t *__restrict__ rx, uint32_t L, uint32_t chunks,
                    const MulTab *__restrict__ g_tw, SynOut so) {
    using TT = Tile<TL, PQ>;
    constexpr int T = TT::N, Q = TT::Q, GB = TT::GB, RB = TT::RB;
    __shared__ uint32_t tile[NBUF * TT::WORDS];
    __shared__ MulTab tw[kTwiddleEntries];
    __shared__ uint32_t s_or[PQ], s_base[PQ];
    stage_tw(tw, g_tw);
    if (threadIdx.x < PQ) s_or[threadIdx.x] = 0u;
    __syncthreads();
    // with 64 quads a wave holds one shard group: wave-uniform twiddle indices,
    // the tables' high words by scalar loads (uni_levels)
    constexpr bool UT = PQ % 64 == 0;
    const int p = threadIdx.x % PQ;
    const int q = UT ? __builtin_amdgcn_readfirstlane((int)(threadIdx.x / PQ)) : (int)(threadIdx.x / PQ);
    const uint32_t quad = blockIdx.x * PQ + p;
    BufIO<AL> io = make_bufio<AL>(rx, chunks * T * L, const_cast<uint8_t *>(rx), 0u, quad, L);
    uint32_t acc[GB][Q], y[GB][Q], x[16];
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h) acc[m][h] = 0u;
    for (uint32_t c = 0; c < chunks; ++c) {
        const uint32_t B = c * T;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[j] = io.ld_in(B + q * 16, j);
        uni_levels<4, 0, 0, 4, true, false, UT>(x, B + q * 16, tw, g_tw);
        uint32_t *tb = tile + (NBUF == 2 ? (c & 1) * TT::WORDS : 0);
        if (NBUF == 1 && c > 0) __syncthreads();
        TT::put_a(tb, x, q, p);
        __syncthreads();
        TT::get_b(tb, y, q, p);
        // chunk 0 sits at beta = 0: its phase-B blocks at offset 0 have the
        // zero twiddle, XOR-only at compile time
        if (c == 0) {
#pragma unroll
            for (int m = 0; m < GB; ++m) uni_levels<RB, 4, 4, TL, true, true, UT>(y[m], 0u, tw, g_tw);
        } else {
#pragma unroll
            for (int m = 0; m < GB; ++m) uni_levels<RB, 4, 4, TL, true, false, UT>(y[m], B, tw, g_tw);
        }
#pragma unroll
        for (int m = 0; m < GB; ++m)
#pragma unroll
            for (int h = 0; h < Q; ++h) acc[m][h] ^= y[m][h];
    }
    uint32_t orw = 0;
#pragma unroll
    for (int m = 0; m < GB; ++m)
#pragma unroll
        for (int h = 0; h < Q; ++h) orw |= acc[m][h];
    if (orw) atomicOr(&s_or[p], orw);
    __syncthreads();
    const uint32_t mask = nonzero_bytes(s_or[p], io.nb);
    if (q == 0 && mask) s_base[p] = list_columns(so, mask, quad * 4u);
    __syncthreads();
    if (mask) {
        uint32_t k = s_base[p];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            if (!(mask & (1u << b))) continue;
            uint8_t *dst = so.syn + (size_t)k++ * T;
#pragma unroll
            for (int m = 0; m < GB; ++m)
#pragma unroll
                for (int h = 0; h < Q; ++h) dst[h * 16 + q * GB + m] = (uint8_t)(acc[m][h] >> (8 * b));
        }
    }
}

__device__ __forceinline__ uint8_t lmul(const uint8_t *ex, const uint8_t *lg, uint8_t a, uint8_t b) {
    return (a && b) ? ex[lg[a] + lg[b]] : (uint8_t)0;
}

// LDS tables of a one-wave block: exp / log and the raw twiddles s_l(omega_o)
// in the ars_gf.hpp layout (byte 1 of a MulTab's first table is c * 1 = c).
struct WaveTabs {
    uint8_t ex[512], lg[256], tw[256];
};

__device__ __forceinline__ void load_wave_tabs(WaveTabs &t, const uint8_t *__restrict__ gtab,
                                               const MulTab *__restrict__ g_tw) {
    for (uint32_t i = threadIdx.x; i < 512; i += 64) t.ex[i] = gtab[i];
    for (uint32_t i = threadIdx.x; i < 256; i += 64) t.lg[i] = gtab[512 + i];
    for (uint32_t i = threadIdx.x; i < (uint32_t)kTwiddleEntries; i += 64) t.tw[i] = (uint8_t)(g_tw[i].w[0] >> 8);
}

// Root search and error values of one column, by the wave of a 64-thread
// block: lam, v, lp hold the first T coefficients (basis X) of lambda, v and
// the derivative; dl = deg lambda (< 0: zero locator, nothing changes).  For
// each chunk c, one forward T-point FFT of the three at beta = omega_{cT}
// (rs/src/kernel/lut_kernel.rs:11-18, 28-54) in `ev`; the first dl roots in
// index order are kept (rs/src/codec.rs:383-396) with their values
// v(w_p) / lp(w_p) (0 where lp(w_p) = 0, rs/src/codec.rs:420-424).  If dl roots
// were found the values are XORed into the column.  Returns true unless the
// column fails (too few roots).
__device__ bool tail_column(const WaveTabs &t, const uint8_t *lam, const uint8_t *v, const uint8_t *lp, int dl,
                            uint32_t n, uint32_t tl, uint8_t *rx, uint32_t L, uint32_t col, uint8_t (*ev)[128],
                            uint8_t *rpos, uint8_t *rval) {
    const uint32_t lane = threadIdx.x;
    if (dl < 0) return true;
    const uint32_t T = 1u << tl, need = (uint32_t)dl;
    uint32_t cnt = 0;  // roots found so far (uniform)
    for (uint32_t c = 0; c < n / T && cnt < need; ++c) {
        for (uint32_t i = lane; i < T; i += 64) {
            ev[0][i] = lam[i];
            ev[1][i] = v[i];
            ev[2][i] = lp[i];
        }
        __syncthreads();
        for (int l = (int)tl - 1; l >= 0; --l) {
            for (uint32_t i = lane; i < T / 2; i += 64) {
                const uint32_t blk = i >> l, pos = i & ((1u << l) - 1u);
                const uint32_t a = (blk << (l + 1)) + pos, b = a + (1u << l);
                const uint32_t o = c * T + (blk << (l + 1));
                const uint8_t tw = t.tw[lvl_off(l) + (o >> (l + 1))];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const uint8_t xa = ev[k][a] ^ lmul(t.ex, t.lg, tw, ev[k][b]);
                    ev[k][a] = xa;
                    ev[k][b] ^= xa;
                }
            }
            __syncthreads();
        }
        for (uint32_t h = 0; h < T; h += 64) {
            const uint32_t off = h + lane;
            const bool root = off < T && ev[0][off] == 0;
            const uint64_t m = __ballot(root);
            const uint32_t rank = cnt + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (root && rank < need) {
                const uint8_t d = ev[2][off];
                rpos[rank] = (uint8_t)(c * T + off);
                rval[rank] = d ? lmul(t.ex, t.lg, ev[1][off], t.ex[255 - t.lg[d]]) : (uint8_t)0;
            }
            cnt += (uint32_t)__popcll(m);
        }
        __syncthreads();
    }
    if (cnt < need) return false;
    for (uint32_t r = lane; r < need; r += 64) rx[(size_t)rpos[r] * L + col] ^= rval[r];
    __syncthreads();
    return true;
}

// One wave per listed column (blockIdx.x = list index), plans from the host.
__global__ void __launch_bounds__(64) k_decode_tail(uint8_t *__restrict__ rx, uint32_t L, uint32_t n, uint32_t tl,
                                                    const uint32_t *__restrict__ cols,
                                                    const DecPlan *__restrict__ plans,
                                                    uint8_t *__restrict__ status, uint32_t *__restrict__ failed,
                                                    const uint8_t *__restrict__ gtab,
                                                    const MulTab *__restrict__ g_tw) {
    __shared__ WaveTabs tabs;
    __shared__ uint8_t pl[3][128];
    __shared__ uint8_t ev[3][128];
    __shared__ uint8_t rpos[256], rval[256];
    load_wave_tabs(tabs, gtab, g_tw);
    const DecPlan *p = plans + blockIdx.x;
    const uint32_t col = cols[blockIdx.x];
    for (uint32_t i = threadIdx.x; i < 128; i += 64) {
        pl[0][i] = p->lam[i];
        pl[1][i] = p->v[i];
        pl[2][i] = p->lp[i];
    }
    const int dl = p->deg;
    __syncthreads();
    const bool ok = tail_column(tabs, pl[0], pl[1], pl[2], dl, n, tl, rx, L, col, ev, rpos, rval);
    if (threadIdx.x == 0) {
        if (status) status[col] = ok ? 1 : 2;
        if (!ok) atomicAdd(failed, 1u);
    }
}

// ---------------------------------------------------------------------------
// The key equation on the device: the same monomial-basis Euclid as the host
// solver (ars_keyeq.hpp), one wave per column, polynomials as LDS byte arrays
// of kCap coefficients with the lanes striding over them.  Degrees are
// wave-uniform values found by ballot.

constexpr int kCap = 264;  // coefficient slots (degrees < 257 and slack)

// Degree of p, given that no coefficient above slot hi is non-zero (slots
// above hi may hold stale values and are not read).  Scans 64-slot windows
// from the top down and stops at the first one that holds a non-zero.
__device__ __forceinline__ int wave_deg(const uint8_t *p, int hi = kCap - 1) {
    for (int s = hi & ~63; s >= 0; s -= 64) {
        const int i = s + (int)threadIdx.x;
        const uint64_t m = __ballot(i <= hi && p[i] != 0);
        if (m) return s + 63 - (int)__clzll(m);
    }
    return -1;
}

__device__ __forceinline__ void wave_fill(uint8_t *p, uint8_t c0, int hi = kCap - 1) {
    for (int i = threadIdx.x; i <= hi; i += 64) p[i] = i == 0 ? c0 : (uint8_t)0;
}

// r <- r mod b, q <- r div b (deg b = db >= 0).  Returns deg of the new r.
// q is written in slots 0 .. dr - db (the quotient's degree); each step
// cancels r's leading term, so the new degree is searched below it.
__device__ int wave_divmod(const WaveTabs &t, uint8_t *r, int dr, const uint8_t *b, int db, uint8_t *q) {
    wave_fill(q, 0, dr - db);
    __syncthreads();
    const uint8_t binv = t.ex[255 - t.lg[b[db]]];
    for (int guard = 0; dr >= db && guard < kCap; ++guard) {
        const int sh = dr - db;
        const uint8_t f = lmul(t.ex, t.lg, r[dr], binv);  // non-zero: r[dr] is r's leading term
        const uint32_t lf = t.lg[f];
        __syncthreads();  // every lane has read r[dr]
        if (threadIdx.x == 0) q[sh] = f;
        for (int j = threadIdx.x; j <= db; j += 64) {
            const uint8_t bj = b[j];
            if (bj) r[sh + j] ^= t.ex[lf + t.lg[bj]];
        }
        __syncthreads();
        dr = wave_deg(r, dr - 1);
    }
    return dr;
}

// out = a * b + c (out distinct from a, b, c; dc = deg c).  Returns deg out.
// Only slots up to max(dc, da + db) are read from c and written to out; the
// callers read every polynomial only up to its degree.  The leading
// coefficients are non-zero, so deg out = max(da + db, dc) unless the two
// tie (then the leading terms may cancel and the degree is searched).
__device__ int wave_muladd(const WaveTabs &t, const uint8_t *a, int da, const uint8_t *b, int db, const uint8_t *c,
                           int dc, uint8_t *out) {
    const int dp = (da >= 0 && db >= 0) ? da + db : -1;
    const int hi = min(dp > dc ? dp : dc, kCap - 1);
    for (int i = threadIdx.x; i <= hi; i += 64) {
        uint8_t acc = i <= dc ? c[i] : (uint8_t)0;
        if (i <= dp) {
            const int j0 = i - db > 0 ? i - db : 0, j1 = i < da ? i : da;
            for (int j = j0; j <= j1; ++j) acc ^= lmul(t.ex, t.lg, a[j], b[i - j]);
        }
        out[i] = acc;
    }
    __syncthreads();
    return dp == dc ? wave_deg(out, hi) : hi;
}

// The two cofactor updates of one Euclid step, which share the quotient a:
// out1 = a * b1 + c1 and out2 = a * b2 + c2 in one pass (one barrier, a's
// terms read once per slot, the two chains interleaved).  out may be c (a
// lane reads c[i] before it writes out[i], at its own slots only), not a or
// b.  Degrees as in wave_muladd, returned in d1 and d2.
__device__ void wave_muladd2(const WaveTabs &t, const uint8_t *a, int da, const uint8_t *b1, int db1,
                             const uint8_t *c1, int dc1, uint8_t *out1, int &d1, const uint8_t *b2, int db2,
                             const uint8_t *c2, int dc2, uint8_t *out2, int &d2) {
    const int dp1 = (da >= 0 && db1 >= 0) ? da + db1 : -1, dp2 = (da >= 0 && db2 >= 0) ? da + db2 : -1;
    const int hi1 = min(dp1 > dc1 ? dp1 : dc1, kCap - 1), hi2 = min(dp2 > dc2 ? dp2 : dc2, kCap - 1);
    const int hi = hi1 > hi2 ? hi1 : hi2;
    for (int i = threadIdx.x; i <= hi; i += 64) {
        uint8_t acc1 = i <= dc1 ? c1[i] : (uint8_t)0, acc2 = i <= dc2 ? c2[i] : (uint8_t)0;
        if (da >= 0) {
            const int j0 = i - (db1 > db2 ? db1 : db2) > 0 ? i - (db1 > db2 ? db1 : db2) : 0, j1 = i < da ? i : da;
            for (int j = j0; j <= j1; ++j) {
                const uint8_t aj = a[j];
                if (!aj) continue;
                const uint32_t la = t.lg[aj];
                const int k = i - j;
                if (k <= db1) {
                    const uint8_t x = b1[k];
                    if (x) acc1 ^= t.ex[la + t.lg[x]];
                }
                if (k <= db2) {
                    const uint8_t x = b2[k];
                    if (x) acc2 ^= t.ex[la + t.lg[x]];
                }
            }
        }
        if (i <= hi1) out1[i] = acc1;
        if (i <= hi2) out2[i] = acc2;
    }
    __syncthreads();
    d1 = dp1 == dc1 ? wave_deg(out1, hi1) : hi1;
    d2 = dp2 == dc2 ? wave_deg(out2, hi2) : hi2;
}

// The basis table xl (256 x 256) holds the monomial coefficients of X_j in
// the log domain: xl[j][i] = log(coefficient of x^i in X_j), kLogZero for 0.
constexpr uint8_t kLogZero = 255;

// a * X coefficient, a given by its log la (kLogZero for 0).
__device__ __forceinline__ uint8_t lmul_log(const WaveTabs &t, uint8_t la, uint8_t lx) {
    return (la != kLogZero && lx != kLogZero) ? t.ex[la + lx] : (uint8_t)0;
}

// Basis X coefficients of the monomial polynomial r (deg dr, destroyed) into
// out, by back-substitution on the monic rows X_d.  Row d has zeros above
// column d, so every lane updates its four slots of r unconditionally.  A
// lane's slots are 4 lane .. 4 lane + 3, so it reads each row as one dword
// (global byte loads cost the texture unit one lane per cycle); row d - 1 is
// loaded while row d is applied.
__device__ void wave_to_novel(const WaveTabs &t, uint8_t *r, int dr, const uint8_t *__restrict__ xl,
                              uint8_t *out) {
    const int lane = threadIdx.x;
    wave_fill(out, 0);
    int d = dr < 255 ? dr : 255;
    constexpr uint32_t kZeroRow = 0xffffffffu;  // four log-0 entries
    uint32_t cur = d >= 0 ? *reinterpret_cast<const uint32_t *>(xl + d * 256 + 4 * lane) : kZeroRow;
    __syncthreads();
    for (; d >= 0; --d) {
        const uint8_t a = r[d];
        const uint32_t nxt = d > 0 ? *reinterpret_cast<const uint32_t *>(xl + (d - 1) * 256 + 4 * lane) : kZeroRow;
        __syncthreads();
        if (a) {
            const uint8_t la = t.lg[a];
            if (lane == 0) out[d] = a;
#pragma unroll
            for (int b = 0; b < 4; ++b) r[4 * lane + b] ^= lmul_log(t, la, (uint8_t)(cur >> (8 * b)));
        }
        __syncthreads();
        cur = nxt;
    }
}

// One wave per listed column, the whole decode of that column: the monomial
// form of the syndrome, X_T = q_t s + r_t, the remainder sequence of (s, r_t)
// to the first remainder of degree < T/2 with its cofactors, lambda = u + v q_t,
// back to the basis X, the derivative (ARS_DEC_FORMAL_DERIVATIVE: formal;
// otherwise the reference's deriv_lnh), then tail_column.  Grid-stride over
// the list; its length is read on the device, so no host round trip.
__global__ void __launch_bounds__(64) k_decode_columns(uint8_t *__restrict__ rx, uint32_t L, uint32_t n, uint32_t tl,
                                                       const uint32_t *__restrict__ count,
                                                       const uint32_t *__restrict__ cols,
                                                       const uint8_t *__restrict__ syn, uint32_t formal,
                                                       uint8_t *__restrict__ status, uint32_t *__restrict__ failed,
                                                       const uint8_t *__restrict__ gtab,
                                                       const MulTab *__restrict__ g_tw,
                                                       const uint8_t *__restrict__ xt) {
    // xt: the basis table in the log domain (see wave_to_novel)
    enum { S, R, QT, Q, U0, U1, V0, V1, TT, LX, VX, LP, NPOLY };
    __shared__ WaveTabs tabs;
    __shared__ uint8_t P[NPOLY][kCap];
    __shared__ uint8_t sx[128];
    __shared__ uint8_t ev[3][128];
    __shared__ uint8_t rpos[256], rval[256];
    load_wave_tabs(tabs, gtab, g_tw);
    __syncthreads();
    const uint32_t lane = threadIdx.x, T = 1u << tl, total = *count;
    for (uint32_t li = blockIdx.x; li < total; li += gridDim.x) {
        const uint32_t col = cols[li];
        for (uint32_t i = lane; i < T; i += 64) {
            const uint8_t v = syn[(size_t)li * T + i];
            sx[i] = v ? tabs.lg[v] : kLogZero;  // logs of the syndrome coefficients
        }
        __syncthreads();
        // s = sum_j sx[j] X_j.  Global byte loads cost the texture unit one lane
        // per cycle, so the basis rows are read as dwords: a lane takes the 4
        // adjacent slots 4 (lane mod 32) .. + 3 of every other row (the two
        // half-waves split the rows and add their partial sums at the end).
        // Row j is zero above column j, so slots >= T sum to 0.
        {
            const int half = (int)(lane >> 5), q4 = (int)(lane & 31u) * 4;
            uint32_t acc = 0;
            if (q4 < (int)T) {
                for (int j = half; j < (int)T; j += 2) {
                    const uint32_t la = sx[j];
                    const uint32_t w = *reinterpret_cast<const uint32_t *>(xt + j * 256 + q4);
                    if (la != kLogZero) {
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const uint32_t lx = (w >> (8 * b)) & 0xffu;
                            if (lx != kLogZero) acc ^= (uint32_t)tabs.ex[la + lx] << (8 * b);
                        }
                    }
                }
            }
            acc ^= (uint32_t)__shfl_xor((int)acc, 32);
            if (half == 0) {
#pragma unroll
                for (int b = 0; b < 4; ++b) P[S][q4 + b] = (uint8_t)(acc >> (8 * b));
            }
            for (int i = 128 + (int)lane; i < kCap; i += 64) P[S][i] = 0;
        }
        __syncthreads();
        const int ds = wave_deg(P[S], (int)T - 1);
        int dl = -1;
        if (ds >= 0) {
            // X_T = q_t s + r_t
            {
                // row T as dwords (slots 0 .. 255; entries above column T are log 0)
                const uint32_t w = *reinterpret_cast<const uint32_t *>(xt + T * 256 + 4 * lane);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const uint32_t l = (w >> (8 * b)) & 0xffu;
                    P[R][4 * lane + b] = l == kLogZero ? (uint8_t)0 : tabs.ex[l];
                }
                for (int i = 256 + (int)lane; i < kCap; i += 64) P[R][i] = 0;
            }
            __syncthreads();
            const int drt = wave_divmod(tabs, P[R], (int)T, P[S], ds, P[QT]);
            const int dqt = (int)T - ds;
            // remainder sequence; the buffers rotate through the index variables
            int z0 = S, z1 = R, dz0 = ds, dz1 = drt;
            int u0 = U0, u1 = U1, v0 = V0, v1 = V1, tt = TT;
            wave_fill(P[u0], 1);
            wave_fill(P[u1], 0);
            wave_fill(P[v0], 0);
            wave_fill(P[v1], 1);
            __syncthreads();
            int du0 = 0, du1 = -1, dv0 = -1, dv1 = 0;
            for (int guard = 0; dz1 >= (int)(T / 2) && guard < kCap; ++guard) {
                const int dr = wave_divmod(tabs, P[z0], dz0, P[z1], dz1, P[Q]);
                const int dq = dz0 - dz1;
                const int zr = z0;
                z0 = z1;
                dz0 = dz1;
                z1 = zr;
                dz1 = dr;
                // u' = q u1 + u0 and v' = q v1 + v0, each in place over u0 / v0
                // (a lane reads c[i] and writes out[i] only at its own slots)
                int dnu, dnv;
                wave_muladd2(tabs, P[Q], dq, P[u1], du1, P[u0], du0, P[u0], dnu, P[v1], dv1, P[v0], dv0, P[v0],
                             dnv);
                int f = u0;
                u0 = u1;
                du0 = du1;
                u1 = f;
                du1 = dnu;
                f = v0;
                v0 = v1;
                dv0 = dv1;
                v1 = f;
                dv1 = dnv;
            }
            // lambda = u + v q_t into P[tt]
            dl = wave_muladd(tabs, P[v1], dv1, P[QT], dqt, P[u1], du1, P[tt]);
            wave_to_novel(tabs, P[tt], dl, xt, P[LX]);
            wave_to_novel(tabs, P[v1], dv1, xt, P[VX]);
            // derivative in the basis X over n coefficients
            if (formal) {
                for (uint32_t i = lane; i < n; i += 64) {
                    uint8_t acc = 0;
                    for (uint32_t b = 1; b < n; b <<= 1)
                        if (!(i & b)) acc ^= P[LX][i | b];
                    P[LP][i] = acc;
                }
                __syncthreads();
            } else {
                for (uint32_t i = lane; i < n; i += 64) P[LP][i] = P[LX][i];
                __syncthreads();
                for (uint32_t half = 1; half < n; half <<= 1) {
                    for (uint32_t i = lane; i < n; i += 64)
                        if (!(i & half)) P[LP][i] ^= P[LP][i + half];
                    __syncthreads();
                }
            }
        }
        const bool ok = tail_column(tabs, P[LX], P[VX], P[LP], dl, n, tl, rx, L, col, ev, rpos, rval);
        if (lane == 0) {
            if (status) status[col] = ok ? 1 : 2;
            if (!ok) atomicAdd(failed, 1u);
        }
        __syncthreads();
    }
}

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

template <int TL, bool AL>
hipError_t syn_one(const uint8_t *rx, uint32_t L, uint32_t chunks, const MulTab *tw, const SynOut &so,
                   hipStream_t s) {
    const uint64_t quads = ((uint64_t)L + 3) / 4;
    if constexpr (TL <= 4) {
        hipLaunchKernelGGL((k_syndrome_small<TL, AL>), dim3(ceil_div(quads, 256)), dim3(256), 0, s, rx, L, chunks,
                           tw, so);
    } else {
        // the encode tile's default shape: 64 quads (scalar table words) and
        // one LDS buffer
        constexpr int PQ = 64, NB = 1;
        hipLaunchKernelGGL((k_syndrome_tile<TL, PQ, NB, AL>), dim3(ceil_div(quads, PQ)),
                           dim3(Tile<TL, PQ>::THREADS), 0, s, rx, L, chunks, tw, so);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_syndrome(int tl, bool al, const uint8_t *rx, uint32_t L, uint32_t chunks, const MulTab *tw,
                           const SynOut &so, hipStream_t s) {
#define SY_CASE(K) \
    case K: return al ? syn_one<K, true>(rx, L, chunks, tw, so, s) : syn_one<K, false>(rx, L, chunks, tw, so, s);
    switch (tl) {
        SY_CASE(1) SY_CASE(2) SY_CASE(3) SY_CASE(4) SY_CASE(5) SY_CASE(6) SY_CASE(7)
        default: return hipErrorInvalidValue;
    }
#undef SY_CASE
}

hipError_t launch_decode_columns(uint8_t *rx, uint32_t L, uint32_t n, int tl, const uint32_t *count,
                                 const uint32_t *cols, const uint8_t *syn, bool formal, uint32_t grid,
                                 uint8_t *status, uint32_t *failed, const uint8_t *gtab, const MulTab *tw,
                                 const uint8_t *xt, hipStream_t s) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_columns, dim3(grid), dim3(64), 0, s, rx, L, n, (uint32_t)tl, count, cols, syn,
                       formal ? 1u : 0u, status, failed, gtab, tw, xt);
    return hipGetLastError();
}

hipError_t launch_decode_tail(uint8_t *rx, uint32_t L, uint32_t n, int tl, const uint32_t *cols,
                              const DecPlan *plans, uint32_t count, uint8_t *status, uint32_t *failed,
                              const uint8_t *gtab, const MulTab *tw, hipStream_t s) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_decode_tail, dim3(count), dim3(64), 0, s, rx, L, n, (uint32_t)tl, cols, plans, status,
                       failed, gtab, tw);
    return hipGetLastError();
}

}  // namespace ars
