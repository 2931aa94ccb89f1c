// The cosine and sine the reference's golden spiral calls (layout/force/barnes_hut/seed.rs:1-7:
// libm::cos, libm::sin of the Rust libm crate 0.2.16, Cargo.lock:271-273), host side: that
// crate ports musl's math library, i.e. Sun's fdlibm. Arguments within 2^20 pi/2 are reduced
// as rem_pio2 does (the first quadrants directly, then rint(x 2/pi) and up to three rounds of
// pi/2 split into 33-bit pieces); larger ones by an exact reduction (x 2/pi over 256 bits of
// 2/pi, the remainder times 256 bits of pi/2, rounded to a head and a tail), where fdlibm's
// Payne-Hanek routine is expected to agree except in its last-bit cases. Then the kernels
// __cos / __sin: fdlibm's degree-14 and degree-13 minimax polynomials on [-pi/4, pi/4].
#pragma once
#include <stdint.h>
#include <string.h>

#include "pi_bits.h"

namespace gr_libm {

static inline uint64_t bits_of(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}

// fdlibm __cos(x, y): x in [-pi/4, pi/4], y its tail
static inline double k_cos(double x, double y) {
    const double c1 = 4.16666666666666019037e-02, c2 = -1.38888888888741095749e-03,
                 c3 = 2.48015872894767294178e-05, c4 = -2.75573143513906633035e-07,
                 c5 = 2.08757232129817482790e-09, c6 = -1.13596475577881948265e-11;
    const double z = x * x, w = z * z;
    const double r = z * (c1 + z * (c2 + z * c3)) + w * w * (c4 + z * (c5 + z * c6));
    const double hz = 0.5 * z, one_minus = 1.0 - hz;
    return one_minus + (((1.0 - one_minus) - hz) + (z * r - x * y));
}

// fdlibm __sin(x, y, iy): iy = 0 when the tail y is zero
static inline double k_sin(double x, double y, int iy) {
    const double s1 = -1.66666666666666324348e-01, s2 = 8.33333333332248946124e-03,
                 s3 = -1.98412698298579493134e-04, s4 = 2.75573137070700676789e-06,
                 s5 = -2.50507602534068634195e-08, s6 = 1.58969099521155010221e-10;
    const double z = x * x, w = z * z;
    const double r = s2 + z * (s3 + z * s4) + z * w * (s5 + z * s6);
    const double v = z * x;
    return iy == 0 ? x + v * (s1 + z * r) : x - ((z * (0.5 * y - v * r) - y) - v * s1);
}

// pi/2 = P1 + P1T = P1 + P2 + P2T = P1 + P2 + P3 + P3T, each Pk holding 33 bits
constexpr double TOINT = 6755399441055744.0, INVPIO2 = 6.36619772367581382433e-01;
constexpr double P1 = 1.57079632673412561417e+00, P1T = 6.07710050650619224932e-11;
constexpr double P2 = 6.07710050630396597660e-11, P2T = 2.02226624879595063154e-21;
constexpr double P3 = 2.02226624871116645580e-21, P3T = 8.47842766036889956997e-32;

static inline int reduce_medium(double x, uint32_t ix, double &y0, double &y1) {
    const double fn = x * INVPIO2 + TOINT - TOINT;
    const int n = (int)fn;
    double r = x - fn * P1, w = fn * P1T;
    y0 = r - w;
    const int ex = (int)(ix >> 20);
    if (ex - (int)((bits_of(y0) >> 52) & 0x7ff) > 16) {
        double t = r;
        w = fn * P2;
        r = t - w;
        w = fn * P2T - ((t - r) - w);
        y0 = r - w;
        if (ex - (int)((bits_of(y0) >> 52) & 0x7ff) > 49) {
            t = r;
            w = fn * P3;
            r = t - w;
            w = fn * P3T - ((t - r) - w);
            y0 = r - w;
        }
    }
    y1 = (r - y0) - w;
    return n;
}

// ---- exact reduction past 2^20 pi/2: little-endian 64-bit limbs ----

// out[0 .. na + nb) = a * b
static inline void mul_limbs(const uint64_t *a, int na, const uint64_t *b, int nb, uint64_t *out) {
    for (int i = 0; i < na + nb; i++) out[i] = 0;
    for (int i = 0; i < na; i++) {
        unsigned __int128 carry = 0;
        for (int j = 0; j < nb; j++) {
            carry += (unsigned __int128)a[i] * b[j] + out[i + j];
            out[i + j] = (uint64_t)carry;
            carry >>= 64;
        }
        out[i + nb] = (uint64_t)carry;
    }
}

static inline int bit_at(const uint64_t *v, int k) { return (int)((v[k >> 6] >> (k & 63)) & 1u); }

// v >> s, keeping nout limbs
static inline void shr_limbs(const uint64_t *v, int nv, int s, uint64_t *out, int nout) {
    for (int i = 0; i < nout; i++) {
        const int lo = i + (s >> 6), sh = s & 63;
        uint64_t w = lo < nv ? v[lo] >> sh : 0;
        if (sh && lo + 1 < nv) w |= v[lo + 1] << (64 - sh);
        out[i] = w;
    }
}

static inline int top_bit(const uint64_t *v, int nv) {
    for (int i = nv - 1; i >= 0; i--)
        if (v[i]) return 64 * i + 63 - __builtin_clzll(v[i]);
    return -1;
}

// The double nearest v * 2^scale (v < 2^448), ties to even; *mant_out, *shift_out give it
// back as mant << shift in v's units.
static inline double nearest(const uint64_t *v, int nv, int scale, uint64_t *mant_out,
                             int *shift_out) {
    const int t = top_bit(v, nv);
    if (t < 0) {
        *mant_out = 0;
        *shift_out = 0;
        return 0.0;
    }
    int shift = t > 52 ? t - 52 : 0;
    uint64_t m[1];
    shr_limbs(v, nv, shift, m, 1);
    uint64_t mant = m[0] & ((1ull << 53) - 1);
    if (shift > 0 && bit_at(v, shift - 1)) {
        const int below = shift - 1;  // bits [0, below) decide a tie
        bool sticky = false;
        for (int i = 0; i < (below >> 6) && !sticky; i++) sticky = v[i] != 0;
        if (!sticky && (below & 63)) sticky = (v[below >> 6] & ((1ull << (below & 63)) - 1)) != 0;
        if (sticky || (mant & 1u)) mant++;
        if (mant >> 53) {
            mant >>= 1;
            shift++;
        }
    }
    *mant_out = mant;
    *shift_out = shift;
    return __builtin_ldexp((double)mant, shift + scale);
}

// |x| in [2^20 pi/2, 2^60): n and y0 + y1 = x - n pi/2, the remainder computed exactly from
// 256 bits of each constant and rounded to a head and a tail.
static inline int reduce_large(double x, double &y0, double &y1) {
    const uint64_t u = bits_of(x);
    const bool neg = u >> 63;
    const int e = (int)((u >> 52) & 0x7ff) - 1075;        // |x| = mant * 2^e
    const uint64_t mant[1] = {(u & ((1ull << 52) - 1)) | (1ull << 52)};
    uint64_t t2[4], p[5];
    for (int i = 0; i < 4; i++) t2[i] = GR_TWO_OVER_PI_256[3 - i];
    mul_limbs(mant, 1, t2, 4, p);                          // |x| 2/pi = p * 2^(e - 256)
    const int s = 256 - e;                                 // fraction bits (e < 0 here)
    uint64_t ip[1];
    shr_limbs(p, 5, s, ip, 1);
    unsigned q = (unsigned)(ip[0] & 3u);
    // the top 192 bits of the fraction: g = G * 2^-192
    uint64_t g[3];
    shr_limbs(p, 5, s - 192, g, 3);
    bool flip = false;
    if (g[2] >> 63) {                                      // g >= 1/2: the next quadrant
        q++;
        flip = true;
        unsigned __int128 borrow = 0;                      // G = 2^192 - G
        for (int i = 0; i < 3; i++) {
            const unsigned __int128 d = (unsigned __int128)0 - g[i] - borrow;
            g[i] = (uint64_t)d;
            borrow = (d >> 64) ? 1 : 0;
        }
    }
    uint64_t pio2[4], r[7];
    for (int i = 0; i < 4; i++) pio2[i] = GR_PIO2_256[3 - i];
    mul_limbs(g, 3, pio2, 4, r);                           // remainder = r * 2^-447
    uint64_t m0;
    int sh0;
    double h = nearest(r, 7, -447, &m0, &sh0);
    // tail = r - m0 << sh0, signed
    uint64_t y[7], d[7];
    for (int i = 0; i < 7; i++) y[i] = 0;
    {
        const int li = sh0 >> 6, b = sh0 & 63;
        y[li] = m0 << b;
        if (b && li + 1 < 7) y[li + 1] = m0 >> (64 - b);
    }
    bool tail_neg = false;
    {
        // d = r - y, or y - r when y > r
        int cmp = 0;
        for (int i = 6; i >= 0 && !cmp; i--) cmp = r[i] > y[i] ? 1 : (r[i] < y[i] ? -1 : 0);
        const uint64_t *a = cmp >= 0 ? r : y, *b = cmp >= 0 ? y : r;
        tail_neg = cmp < 0;
        unsigned __int128 borrow = 0;
        for (int i = 0; i < 7; i++) {
            const unsigned __int128 v = (unsigned __int128)a[i] - b[i] - borrow;
            d[i] = (uint64_t)v;
            borrow = (v >> 64) ? 1 : 0;
        }
    }
    uint64_t m1;
    int sh1;
    double tl = nearest(d, 7, -447, &m1, &sh1);
    if (tail_neg) tl = -tl;
    if (flip) {
        h = -h;
        tl = -tl;
    }
    if (neg) {
        y0 = -h;
        y1 = -tl;
        return -(int)q;
    }
    y0 = h;
    y1 = tl;
    return (int)q;
}

static inline int reduce(double x, double &y0, double &y1) {
    const uint64_t u = bits_of(x);
    const bool neg = u >> 63;
    const uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    double k;
    if (ix <= 0x400f6a7a) {                                // |x| ~<= 5pi/4
        if ((ix & 0xfffff) == 0x921fb) return reduce_medium(x, ix, y0, y1);
        k = ix <= 0x4002d97c ? 1.0 : 2.0;
    } else if (ix <= 0x401c463b) {                         // |x| ~<= 9pi/4
        if (ix == 0x4012d97c || ix == 0x401921fb) return reduce_medium(x, ix, y0, y1);
        k = ix <= 0x4015fdbc ? 3.0 : 4.0;
    } else if (ix < 0x413921fb) {                          // |x| ~< 2^20 pi/2
        return reduce_medium(x, ix, y0, y1);
    } else {
        return reduce_large(x, y0, y1);
    }
    if (!neg) {
        const double z = x - k * P1;
        y0 = z - k * P1T;
        y1 = (z - y0) - k * P1T;
        return (int)k;
    }
    const double z = x + k * P1;
    y0 = z + k * P1T;
    y1 = (z - y0) + k * P1T;
    return -(int)k;
}

// libm::cos. Arguments: finite, |x| < 2^60 (the spiral's angles stay below 2^34).
static inline double cos(double x) {
    const uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fb) return ix < 0x3e46a09e ? 1.0 : k_cos(x, 0.0);
    if (ix >= 0x7ff00000) return x - x;
    double y0, y1;
    switch (reduce(x, y0, y1) & 3) {
    case 0: return k_cos(y0, y1);
    case 1: return -k_sin(y0, y1, 1);
    case 2: return -k_cos(y0, y1);
    default: return k_sin(y0, y1, 1);
    }
}

// libm::sin, same domain.
static inline double sin(double x) {
    const uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fb) return ix < 0x3e500000 ? x : k_sin(x, 0.0, 0);
    if (ix >= 0x7ff00000) return x - x;
    double y0, y1;
    switch (reduce(x, y0, y1) & 3) {
    case 0: return k_sin(y0, y1, 1);
    case 1: return k_cos(y0, y1);
    case 2: return -k_sin(y0, y1, 1);
    default: return -k_cos(y0, y1);
    }
}

}  // namespace gr_libm
