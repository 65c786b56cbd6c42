// ars_gf.hpp -- host-side GF(2^8) / Cantor-basis tables for the product path.
//
// The field is GF(2)[x]/(x^8+x^4+x^3+x^2+1) (0x11d) with generator 2, the
// evaluation points are the span of the Cantor basis v_0 = 1,
// v_i^2 + v_i = v_{i-1} (trace-0 root, smallest first), and the subspace
// polynomials are s_0(x) = x, s_{j+1} = s_j (s_j + s_j(v_j)).  These are the
// definitions of gf2p8/src/field.rs:22-37, 68-102, 122-198, 218-235, 312-336
// and gf2p8/src/poly_11d.rs:18-21; the tables are recomputed here from the
// mathematics so the library has no dependency on test code.
#pragma once
#include <cstdint>
#include <cstring>

namespace ars {

// A constant multiplier in the form the device kernels consume: three 8-byte
// lookup tables for v_perm_b32, one per bit chunk of the multiplicand
// (bits 0-2, 3-5, 6-7).  By linearity of multiplication by a constant over
// GF(2), c*x = T0[x & 7] ^ T1[(x >> 3) & 7] ^ T2[x >> 6].
struct alignas(8) MulTab {
    uint32_t w[6];  // w0,w1: T0 (lo/hi 4 bytes); w2,w3: T1; w4,w5: T2 (w5 = 0)
};

struct GfTables {
    uint8_t exp[512];
    uint8_t log[256];
    uint8_t inv[256];
    uint8_t basis[8];
    uint8_t sub[256];       // omega_i
    uint8_t spv[9][256];    // s_j(x)

    static uint8_t mul_slow(uint8_t a, uint8_t b) {
        uint16_t x = a, r = 0;
        for (int i = 0; i < 8; ++i) {
            if (b & (1u << i)) r ^= x;
            x <<= 1;
            if (x & 0x100) x ^= 0x11d;
        }
        return (uint8_t)r;
    }
    static int trace(uint8_t x) {
        uint8_t t = x, s = x;
        for (int i = 0; i < 7; ++i) { t = mul_slow(t, t); s ^= t; }
        return s & 1;
    }
    uint8_t mul(uint8_t a, uint8_t b) const {
        if (!a || !b) return 0;
        return exp[log[a] + log[b]];
    }

    GfTables() {
        uint8_t x = 1;
        for (int i = 0; i < 255; ++i) { exp[i] = x; log[x] = (uint8_t)i; x = mul_slow(x, 2); }
        for (int i = 255; i < 512; ++i) exp[i] = exp[i - 255];
        log[0] = 0;
        inv[0] = 0;
        for (int a = 1; a < 256; ++a) inv[a] = exp[255 - log[a]];
        basis[0] = 1;
        for (int i = 1; i < 8; ++i) {
            uint8_t root = 0;
            for (int r = 0; r < 256; ++r) {
                if ((uint8_t)(mul_slow((uint8_t)r, (uint8_t)r) ^ r) == basis[i - 1]) { root = (uint8_t)r; break; }
            }
            basis[i] = trace(root) ? (uint8_t)(root ^ 1) : root;
        }
        for (int i = 0; i < 256; ++i) {
            uint8_t p = 0;
            for (int b = 0; b < 8; ++b) if ((i >> b) & 1) p ^= basis[b];
            sub[i] = p;
        }
        for (int i = 0; i < 256; ++i) spv[0][i] = (uint8_t)i;
        for (int j = 0; j < 8; ++j) {
            uint8_t bj = spv[j][basis[j]];
            for (int i = 0; i < 256; ++i) spv[j + 1][i] = mul(spv[j][i], (uint8_t)(spv[j][i] ^ bj));
        }
    }

    MulTab multab(uint8_t c) const {
        MulTab m;
        uint8_t t[3][8];
        for (int v = 0; v < 8; ++v) {
            t[0][v] = mul(c, (uint8_t)v);
            t[1][v] = mul(c, (uint8_t)(v << 3));
            t[2][v] = v < 4 ? mul(c, (uint8_t)(v << 6)) : 0;
        }
        for (int k = 0; k < 3; ++k) {
            m.w[2 * k] = t[k][0] | (t[k][1] << 8) | (t[k][2] << 16) | ((uint32_t)t[k][3] << 24);
            m.w[2 * k + 1] = t[k][4] | (t[k][5] << 8) | (t[k][6] << 16) | ((uint32_t)t[k][7] << 24);
        }
        return m;
    }
};

// Twiddle table layout shared by host and device: level l (0..7) holds the
// multipliers s_l(omega_{j * 2^(l+1)}) for j < 128 >> l, starting at
// entry 256 - (256 >> l).  255 entries in total; entry j = 0 of each level is
// the zero multiplier (s_l vanishes on V_l).
constexpr int kTwiddleEntries = 255;
constexpr int lvl_off(int l) { return 256 - (256 >> l); }

inline const GfTables& gf() {
    static const GfTables t;
    return t;
}

}  // namespace ars
