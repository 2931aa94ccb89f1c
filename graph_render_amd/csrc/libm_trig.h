// The cosine and sine the reference's golden spiral calls (layout/force/barnes_hut/seed.rs:1-7:
// libm::cos, libm::sin of the Rust libm crate 0.2.16, Cargo.lock:271-273), host side: that
// crate ports musl's math library, i.e. Sun's fdlibm. Arguments within 2^20 pi/2 are reduced
// as rem_pio2 does (the first quadrants directly, then rint(x 2/pi) and up to three rounds of
// pi/2 split into 33-bit pieces); larger ones as its rem_pio2_large does (fdlibm's
// __kernel_rem_pio2 over 24-bit chunks of 2/pi and pi/2 from pi_bits.h). Then the kernels
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

// ---- past 2^20 pi/2: libm's rem_pio2_large (fdlibm __kernel_rem_pio2, double: jk = 4) ----
//
// The reference's bits depend on this routine's own double arithmetic (its remainder's tail is
// not always the exact remainder's rounded tail), so its steps are taken as they are: |x| as
// 24-bit integer chunks times the 24-bit chunks of 2/pi that reach its integer part modulo 8
// and its fraction, the fraction as 24-bit integers (one minus it past one half, more chunks
// while its leading pieces are all zero), those times pi/2 in 24-bit pieces, summed small to
// large into a head and a tail.
struct LargeReduction {
    static constexpr int JK = 4;
    double chunk[3];  // |x| * 2^-e0 in 24-bit integer chunks
    int nchunk, e0;

    // f[k] * x summed over the chunks for fraction piece i (f indexed as the routine does)
    static double dot(const double *x, int jx, const double *f, int i) {
        double acc = 0.0;
        for (int j = 0; j <= jx; j++) acc += x[j] * f[jx + i - j];
        return acc;
    }

    int run(double &head, double &tail) const {
        int32_t iq[20];
        double f[20], fq[20], q[20];
        const int jx = nchunk - 1;
        const int jv = (e0 - 3) / 24 > 0 ? (e0 - 3) / 24 : 0;
        int q0 = e0 - 24 * (jv + 1);
        for (int i = 0; i <= jx + JK; i++) {
            const int j = jv - jx + i;
            f[i] = j < 0 ? 0.0 : (double)GR_TWO_OVER_PI_24[j];
        }
        for (int i = 0; i <= JK; i++) q[i] = dot(chunk, jx, f, i);
        int jz = JK, n = 0, ih = 0;
        double z = 0.0;
        for (bool again = true; again;) {
            again = false;
            // q[jz] .. q[0] as 24-bit integers, least significant first; z the integer part
            z = q[jz];
            for (int i = 0, j = jz; j > 0; i++, j--) {
                const double hi = (double)(int32_t)(0x1p-24 * z);
                iq[i] = (int32_t)(z - 0x1p24 * hi);
                z = q[j - 1] + hi;
            }
            z = __builtin_scalbn(z, q0);
            z -= 8.0 * __builtin_floor(z * 0.125);
            n = (int32_t)z;
            z -= (double)n;
            ih = 0;
            if (q0 > 0) {
                const int32_t carry_n = iq[jz - 1] >> (24 - q0);
                n += carry_n;
                iq[jz - 1] -= carry_n << (24 - q0);
                ih = iq[jz - 1] >> (23 - q0);
            } else if (q0 == 0) {
                ih = iq[jz - 1] >> 23;
            } else if (z >= 0.5) {
                ih = 2;
            }
            if (ih > 0) {
                n += 1;
                bool borrowed = false;
                for (int i = 0; i < jz; i++) {
                    const int32_t v = iq[i];
                    if (!borrowed) {
                        if (v != 0) {
                            borrowed = true;
                            iq[i] = 0x1000000 - v;
                        }
                    } else {
                        iq[i] = 0xffffff - v;
                    }
                }
                if (q0 == 1) iq[jz - 1] &= 0x7fffff;
                if (q0 == 2) iq[jz - 1] &= 0x3fffff;
                if (ih == 2) {
                    z = 1.0 - z;
                    if (borrowed) z -= __builtin_scalbn(1.0, q0);
                }
            }
            if (z == 0.0) {
                int32_t any = 0;
                for (int i = jz - 1; i >= JK; i--) any |= iq[i];
                if (any == 0) {
                    int more = 1;
                    while (iq[JK - more] == 0) more++;
                    for (int i = jz + 1; i <= jz + more; i++) {
                        f[jx + i] = (double)GR_TWO_OVER_PI_24[jv + i];
                        q[i] = dot(chunk, jx, f, i);
                    }
                    jz += more;
                    again = true;
                }
            }
        }
        if (z == 0.0) {
            jz -= 1;
            q0 -= 24;
            while (iq[jz] == 0) {
                jz--;
                q0 -= 24;
            }
        } else {
            z = __builtin_scalbn(z, -q0);
            if (z >= 0x1p24) {
                const double hi = (double)(int32_t)(0x1p-24 * z);
                iq[jz] = (int32_t)(z - 0x1p24 * hi);
                jz += 1;
                q0 += 24;
                iq[jz] = (int32_t)hi;
            } else {
                iq[jz] = (int32_t)z;
            }
        }
        double w = __builtin_scalbn(1.0, q0);
        for (int i = jz; i >= 0; i--) {
            q[i] = w * (double)iq[i];
            w *= 0x1p-24;
        }
        for (int i = jz; i >= 0; i--) {
            double acc = 0.0;
            for (int k = 0; k <= JK && k <= jz - i; k++) acc += GR_PIO2_24[k] * q[i + k];
            fq[jz - i] = acc;
        }
        double sum = 0.0;
        for (int i = jz; i >= 0; i--) sum += fq[i];
        head = ih == 0 ? sum : -sum;
        double rest = fq[0] - sum;
        for (int i = 1; i <= jz; i++) rest += fq[i];
        tail = ih == 0 ? rest : -rest;
        return n & 7;
    }
};

// rem_pio2's large case: the significand of |x| in 24-bit chunks (trailing zero chunks
// dropped), the remainder and the quadrant signed as x
static inline int reduce_large(double x, double &y0, double &y1) {
    const uint64_t u = bits_of(x);
    const uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    const uint64_t scaled = (u & ((uint64_t)-1 >> 12)) | ((uint64_t)(0x3ff + 23) << 52);
    double z;
    memcpy(&z, &scaled, 8);
    LargeReduction r;
    for (int i = 0; i < 2; i++) {
        r.chunk[i] = (double)(int32_t)z;
        z = (z - r.chunk[i]) * 0x1p24;
    }
    r.chunk[2] = z;
    int last = 2;
    while (last > 0 && r.chunk[last] == 0.0) last--;
    r.nchunk = last + 1;
    r.e0 = (int)(ix >> 20) - (0x3ff + 23);
    double h, t;
    const int n = r.run(h, t);
    if (u >> 63) {
        y0 = -h;
        y1 = -t;
        return -n;
    }
    y0 = h;
    y1 = t;
    return n;
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

// libm::cos of a finite x (the spiral's angles stay below 2^34).
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
