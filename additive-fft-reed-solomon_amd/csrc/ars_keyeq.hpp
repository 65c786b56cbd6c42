// ars_keyeq.hpp -- host side of sharded error decoding: the key equation of
// one byte column.
//
// For a column whose syndrome s (T coefficients in the novel basis X_j) is
// not zero, the reference's decoder (rs/src/codec.rs:212-239, 341-420) works
// with s_t = X_T = q_t s + r_t, the Euclidean remainder sequence of (s, r_t)
// stopped at the first remainder of degree < T/2, its cofactors (u, v), the
// error locator lambda = u + v q_t, and v itself as the numerator of the
// error values.  Polynomial identities do not depend on the basis the
// coefficients are written in, and division with remainder is unique, so the
// sequence is run here in the monomial basis with plain long division; only
// the results go back to the basis X, where the device evaluates them by FFT.
//
// HIP-free so it builds and is tested with g++ against the oracle.
#pragma once
#include <cstdint>
#include <cstring>

#include "ars_gf.hpp"

namespace ars {

// Monomial-basis polynomial of degree < 257.
struct MPoly {
    int deg = -1;  // -1 for the zero polynomial
    uint8_t c[257];
    MPoly() { std::memset(c, 0, sizeof c); }
    void trim() {
        while (deg >= 0 && c[deg] == 0) --deg;
    }
};

// Field product table and the monomial coefficients of X_j = prod_{b in j}
// s_b(x) (monic of degree j), built once per process.
struct KeyTables {
    uint8_t mul[256][256];
    uint8_t x[256][256];  // x[j][i]: coefficient of x^i in X_j

    KeyTables() {
        const GfTables &g = gf();
        for (int a = 0; a < 256; ++a)
            for (int b = 0; b < 256; ++b) mul[a][b] = g.mul((uint8_t)a, (uint8_t)b);
        // s_b in monomial form: s_0 = x, s_{b+1} = s_b^2 + s_b(v_b) s_b; squaring a
        // linearised polynomial squares its coefficients and doubles exponents.
        uint8_t s[8][129];
        std::memset(s, 0, sizeof s);
        s[0][1] = 1;
        for (int b = 0; b < 7; ++b) {
            const uint8_t k = g.spv[b][g.basis[b]];
            for (int i = 0; i <= 64; ++i) {
                if (!s[b][i]) continue;
                s[b + 1][2 * i] ^= mul[s[b][i]][s[b][i]];
                s[b + 1][i] ^= mul[k][s[b][i]];
            }
        }
        std::memset(x, 0, sizeof x);
        x[0][0] = 1;
        for (int j = 1; j < 256; ++j) {
            const int top = 31 - __builtin_clz((unsigned)j);
            const int rest = j ^ (1 << top);  // X_j = X_rest * s_top
            for (int i = 0; i <= rest; ++i) {
                const uint8_t a = x[rest][i];
                if (!a) continue;
                for (int e = 0; e <= (1 << top); ++e)
                    if (s[top][e]) x[j][i + e] ^= mul[a][s[top][e]];
            }
        }
    }
};

inline const KeyTables &key_tables() {
    static const KeyTables t;
    return t;
}

// sum_{j < cnt} v[j] X_j
inline void mono_from_novel(const uint8_t *v, int cnt, MPoly &out) {
    const KeyTables &kt = key_tables();
    out = MPoly();
    for (int j = 0; j < cnt; ++j) {
        const uint8_t a = v[j];
        if (!a) continue;
        for (int i = 0; i <= j; ++i) out.c[i] ^= kt.mul[a][kt.x[j][i]];
    }
    out.deg = cnt - 1;
    out.trim();
}

// Coefficients in the basis X (256 entries) of p, by back-substitution on
// the monic rows X_d.
inline void novel_from_mono(const MPoly &p, uint8_t *out) {
    const KeyTables &kt = key_tables();
    uint8_t r[257];
    std::memcpy(r, p.c, sizeof r);
    std::memset(out, 0, 256);
    // every polynomial of the key equation has degree < 3T/2 <= 192
    for (int d = p.deg < 255 ? p.deg : 255; d >= 0; --d) {
        const uint8_t a = r[d];
        out[d] = a;
        if (!a) continue;
        for (int i = 0; i <= d; ++i) r[i] ^= kt.mul[a][kt.x[d][i]];
    }
}

inline void mp_mul(const MPoly &a, const MPoly &b, MPoly &out) {
    const KeyTables &kt = key_tables();
    MPoly r;
    if (a.deg >= 0 && b.deg >= 0) {
        for (int i = 0; i <= a.deg; ++i) {
            const uint8_t ai = a.c[i];
            if (!ai) continue;
            const uint8_t *row = kt.mul[ai];
            for (int j = 0; j <= b.deg && i + j < 257; ++j) r.c[i + j] ^= row[b.c[j]];
        }
        r.deg = a.deg + b.deg < 256 ? a.deg + b.deg : 256;
        r.trim();
    }
    out = r;
}

inline void mp_add(const MPoly &a, const MPoly &b, MPoly &out) {
    MPoly r;
    const int d = a.deg > b.deg ? a.deg : b.deg;
    for (int i = 0; i <= d; ++i) r.c[i] = a.c[i] ^ b.c[i];
    r.deg = d;
    r.trim();
    out = r;
}

// a = q b + r, deg r < deg b; b must not be zero.
inline void mp_divmod(const MPoly &a, const MPoly &b, MPoly &q, MPoly &r) {
    const KeyTables &kt = key_tables();
    const GfTables &g = gf();
    MPoly qq, rr = a;
    const uint8_t binv = g.inv[b.c[b.deg]];
    while (rr.deg >= b.deg) {
        const int sh = rr.deg - b.deg;
        const uint8_t f = kt.mul[rr.c[rr.deg]][binv];
        qq.c[sh] = f;
        if (qq.deg < sh) qq.deg = sh;
        const uint8_t *row = kt.mul[f];
        for (int j = 0; j <= b.deg; ++j) rr.c[sh + j] ^= row[b.c[j]];
        rr.trim();
    }
    qq.trim();
    q = qq;
    r = rr;
}

// What the device's root-search step needs for one column: the first T
// coefficients (basis X) of lambda, of v and of the derivative used for the
// error values, and deg lambda.
struct DecPlan {
    uint8_t lam[128];
    uint8_t v[128];
    uint8_t lp[128];
    int16_t deg;  // deg lambda; -1: zero locator (the column is left as is)
    uint8_t pad[2];
};
static_assert(sizeof(DecPlan) == 388, "DecPlan layout is shared with the device");

// formal = false: the reference's poly::deriv_lnh (rs/src/poly_arith.rs:91-116),
// res[i] = sum over s < n with s a bit-superset of i of c[s].  formal = true:
// the formal derivative, res[i] = xor over bits b not in i of c[i | 2^b]
// (every s_b has x-coefficient 1, so X_j' = sum_{b in j} X_{j - 2^b}).
inline void lnh_derivative(const uint8_t *c, uint32_t n, bool formal, uint8_t *out) {
    if (formal) {
        for (uint32_t i = 0; i < n; ++i) {
            uint8_t v = 0;
            for (uint32_t b = 1; b < n; b <<= 1)
                if (!(i & b)) v ^= c[i | b];
            out[i] = v;
        }
        return;
    }
    std::memcpy(out, c, n);
    for (uint32_t half = 1; half < n; half <<= 1)
        for (uint32_t s = 0; s < n; s += 2 * half)
            for (uint32_t i = 0; i < half; ++i) out[s + i] ^= out[s + i + half];
}

// Fills `plan` for a column with syndrome `syn` (T coefficients, basis X).
// Returns false if the syndrome is zero (nothing to solve).
inline bool solve_key_equation(const uint8_t *syn, uint32_t n, uint32_t T, bool formal, DecPlan &plan) {
    MPoly s;
    mono_from_novel(syn, (int)T, s);
    if (s.deg < 0) return false;
    const KeyTables &kt = key_tables();
    MPoly st;  // X_T in the monomial basis
    for (uint32_t i = 0; i <= T; ++i) st.c[i] = kt.x[T][i];
    st.deg = (int)T;
    MPoly qt, rt;
    mp_divmod(st, s, qt, rt);
    // remainder sequence from (s, r_t) to the first remainder of degree < T/2,
    // with cofactors u, v of s and r_t
    MPoly z0 = s, z1 = rt, u0, u1, v0, v1, q, rem, t;
    u0.c[0] = 1;
    u0.deg = 0;
    v1.c[0] = 1;
    v1.deg = 0;
    while (z1.deg >= (int)(T / 2)) {
        mp_divmod(z0, z1, q, rem);
        z0 = z1;
        z1 = rem;
        mp_mul(q, u1, t);
        mp_add(u0, t, t);
        u0 = u1;
        u1 = t;
        mp_mul(q, v1, t);
        mp_add(v0, t, t);
        v0 = v1;
        v1 = t;
    }
    MPoly lam;
    mp_mul(v1, qt, t);
    mp_add(u1, t, lam);
    uint8_t lx[256], vx[256], dx[256];
    novel_from_mono(lam, lx);
    novel_from_mono(v1, vx);
    lnh_derivative(lx, n, formal, dx);
    std::memset(&plan, 0, sizeof plan);
    std::memcpy(plan.lam, lx, T);
    std::memcpy(plan.v, vx, T);
    std::memcpy(plan.lp, dx, T);
    plan.deg = (int16_t)lam.deg;
    return true;
}

}  // namespace ars
