/*
 * Test infrastructure (oracle): the cosine and sine of the Rust `libm` crate 0.2.16
 * (Cargo.lock:271-273), which the reference's golden spiral calls
 * (crates/graph-core/src/layout/force/barnes_hut/seed.rs:1-7: libm::cos, libm::sin).
 * That crate ports musl's math library, itself Sun's fdlibm: argument reduction by
 * rem_pio2 (Cody-Waite with pi/2 split into 33-bit pieces, up to three rounds), then the
 * kernels __cos / __sin, degree-12/13 minimax polynomials on [-pi/4, pi/4].
 *
 * Pinned by the reference's own fixture fixtures/gpu/mesh-1k-start.gmfx (the spiral of
 * 1000 points and a 128-point twiddle table, every bit); the reduction constants are checked
 * against pi computed exactly. Arguments past 2^20 pi/2 take that crate's rem_pio2_large
 * (musl __rem_pio2_large, fdlibm __kernel_rem_pio2, double precision: jk = 4), restated from
 * the published algorithm over tables generated from exact pi (gen_ipio2.py); no reference
 * value exists there, so it is pinned by the exact remainder and exact cos/sin instead
 * (tests/test_libm_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "ipio2.h"

static uint64_t bits_of(double x) {
    uint64_t u;
    memcpy(&u, &x, 8);
    return u;
}

/* rem_pio2's constants: 2/pi and pi/2 as three 33-bit pieces with their tails */
static const double TOINT = 6755399441055744.0; /* 1.5 * 2^52 */
static const double INVPIO2 = 6.36619772367581382433e-01;
static const double PIO2_1 = 1.57079632673412561417e+00;
static const double PIO2_1T = 6.07710050650619224932e-11;
static const double PIO2_2 = 6.07710050630396597660e-11;
static const double PIO2_2T = 2.02226624879595063154e-21;
static const double PIO2_3 = 2.02226624871116645580e-21;
static const double PIO2_3T = 8.47842766036889956997e-32;

/* the medium case: n = rint(x 2/pi), y = x - n pi/2 in up to three rounds */
static int reduce_medium(double x, uint32_t ix, double *y0, double *y1) {
    const double fn = x * INVPIO2 + TOINT - TOINT;
    const int n = (int)fn;
    double r = x - fn * PIO2_1;
    double w = fn * PIO2_1T;
    double y = r - w;
    const int ex = (int)(ix >> 20);
    int ey = (int)((bits_of(y) >> 52) & 0x7ff);
    if (ex - ey > 16) {
        double t = r;
        w = fn * PIO2_2;
        r = t - w;
        w = fn * PIO2_2T - ((t - r) - w);
        y = r - w;
        ey = (int)((bits_of(y) >> 52) & 0x7ff);
        if (ex - ey > 49) {
            t = r;
            w = fn * PIO2_3;
            r = t - w;
            w = fn * PIO2_3T - ((t - r) - w);
            y = r - w;
        }
    }
    *y0 = y;
    *y1 = (r - y) - w;
    return n;
}

/* rem_pio2_large: x is |arg| * 2^-e0 cut into nx 24-bit integer chunks; the remainder of
 * |arg| modulo pi/2 as y[0] + y[1] and the quadrant n & 7. The integer part of |arg| 2/pi
 * is formed modulo 8 from the chunks of 2/pi that reach it, the fraction from the next ones
 * in 24-bit integer pieces (more of them while the fraction's leading pieces are all zero),
 * taken as 1 - q past one half; then those pieces times pi/2 in 24-bit pieces, summed small
 * to large. */
static int rem_pio2_large(const double *x, double *y, int e0, int nx) {
    const int jk = 4, jp = 4; /* init_jk for double precision */
    int32_t iq[20];
    double f[20], fq[20], q[20], z, fw;
    const int jx = nx - 1;
    int jv = (e0 - 3) / 24;
    if (jv < 0) jv = 0;
    int q0 = e0 - 24 * (jv + 1);
    /* f[0..jx+jk]: the chunks of 2/pi that meet the argument's chunks */
    for (int i = 0, j = jv - jx; i <= jx + jk; i++, j++) f[i] = j < 0 ? 0.0 : (double)IPIO2[j];
    for (int i = 0; i <= jk; i++) {
        fw = 0.0;
        for (int j = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
        q[i] = fw;
    }
    int jz = jk, n, ih;
    for (;;) {
        /* q[jz..0] into 24-bit integers iq[0..jz-1] (least significant first), z the top */
        int i = 0;
        z = q[jz];
        for (int j = jz; j > 0; i++, j--) {
            fw = (double)(int32_t)(0x1p-24 * z);
            iq[i] = (int32_t)(z - 0x1p24 * fw);
            z = q[j - 1] + fw;
        }
        z = scalbn(z, q0);
        z -= 8.0 * floor(z * 0.125); /* the integer part modulo 8 */
        n = (int32_t)z;
        z -= (double)n;
        ih = 0;
        if (q0 > 0) { /* the last bits of n sit in iq[jz-1] */
            i = iq[jz - 1] >> (24 - q0);
            n += i;
            iq[jz - 1] -= i << (24 - q0);
            ih = iq[jz - 1] >> (23 - q0);
        } else if (q0 == 0) {
            ih = iq[jz - 1] >> 23;
        } else if (z >= 0.5) {
            ih = 2;
        }
        if (ih > 0) { /* fraction past one half: the next quadrant, fraction 1 - q */
            n += 1;
            int carry = 0;
            for (i = 0; i < jz; i++) {
                const int32_t j = iq[i];
                if (carry == 0) {
                    if (j != 0) {
                        carry = 1;
                        iq[i] = 0x1000000 - j;
                    }
                } else {
                    iq[i] = 0xffffff - j;
                }
            }
            if (q0 == 1) iq[jz - 1] &= 0x7fffff;
            if (q0 == 2) iq[jz - 1] &= 0x3fffff;
            if (ih == 2) {
                z = 1.0 - z;
                if (carry != 0) z -= scalbn(1.0, q0);
            }
        }
        if (z != 0.0) break;
        int any = 0;
        for (i = jz - 1; i >= jk; i--) any |= iq[i];
        if (any) break;
        /* every fraction piece beyond jk is zero: take k more chunks of 2/pi */
        int k = 1;
        while (iq[jk - k] == 0) k++;
        for (i = jz + 1; i <= jz + k; i++) {
            f[jx + i] = (double)IPIO2[jv + i];
            fw = 0.0;
            for (int j = 0; j <= jx; j++) fw += x[j] * f[jx + i - j];
            q[i] = fw;
        }
        jz += k;
    }
    /* drop zero pieces, or split z into 24-bit pieces */
    if (z == 0.0) {
        jz -= 1;
        q0 -= 24;
        while (iq[jz] == 0) {
            jz--;
            q0 -= 24;
        }
    } else {
        z = scalbn(z, -q0);
        if (z >= 0x1p24) {
            fw = (double)(int32_t)(0x1p-24 * z);
            iq[jz] = (int32_t)(z - 0x1p24 * fw);
            jz += 1;
            q0 += 24;
            iq[jz] = (int32_t)fw;
        } else {
            iq[jz] = (int32_t)z;
        }
    }
    fw = scalbn(1.0, q0);
    for (int i = jz; i >= 0; i--) {
        q[i] = fw * (double)iq[i];
        fw *= 0x1p-24;
    }
    /* the pieces times pi/2 */
    for (int i = jz; i >= 0; i--) {
        fw = 0.0;
        for (int k = 0; k <= jp && k <= jz - i; k++) fw += PIO2_24[k] * q[i + k];
        fq[jz - i] = fw;
    }
    fw = 0.0;
    for (int i = jz; i >= 0; i--) fw += fq[i];
    y[0] = ih == 0 ? fw : -fw;
    fw = fq[0] - fw;
    for (int i = 1; i <= jz; i++) fw += fq[i];
    y[1] = ih == 0 ? fw : -fw;
    return n & 7;
}

/* rem_pio2's large case: |x| as three 24-bit chunks of its significand (trailing zero
 * chunks dropped), the exponent e0 of the first; the sign applied to the remainder and n */
static int reduce_large(double x, uint32_t ix, double *y0, double *y1) {
    const uint64_t u = bits_of(x);
    uint64_t m = (u & ((uint64_t)-1 >> 12)) | ((uint64_t)(0x3ff + 23) << 52);
    double z;
    memcpy(&z, &m, 8);
    double tx[3], ty[2];
    int i;
    for (i = 0; i < 2; i++) {
        tx[i] = (double)(int32_t)z;
        z = (z - tx[i]) * 0x1p24;
    }
    tx[i] = z;
    while (tx[i] == 0.0) i--;
    const int n = rem_pio2_large(tx, ty, (int)(ix >> 20) - (0x3ff + 23), i + 1);
    if (u >> 63) {
        *y0 = -ty[0];
        *y1 = -ty[1];
        return -n;
    }
    *y0 = ty[0];
    *y1 = ty[1];
    return n;
}

/* rem_pio2: the quadrant n and y0 + y1 = x - n pi/2 (x finite) */
static int reduce(double x, double *y0, double *y1) {
    const uint64_t u = bits_of(x);
    const int neg = (int)(u >> 63);
    const uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    if (ix <= 0x400f6a7a) {                            /* |x| ~<= 5pi/4 */
        if ((ix & 0xfffff) == 0x921fb) return reduce_medium(x, ix, y0, y1);
        const double k = ix <= 0x4002d97c ? 1.0 : 2.0; /* |x| ~<= 3pi/4: one quadrant */
        if (!neg) {
            const double z = x - k * PIO2_1;
            *y0 = z - k * PIO2_1T;
            *y1 = (z - *y0) - k * PIO2_1T;
            return (int)k;
        }
        const double z = x + k * PIO2_1;
        *y0 = z + k * PIO2_1T;
        *y1 = (z - *y0) + k * PIO2_1T;
        return -(int)k;
    }
    if (ix <= 0x401c463b) {                            /* |x| ~<= 9pi/4 */
        double k;
        if (ix <= 0x4015fdbc) {                        /* |x| ~<= 7pi/4 */
            if (ix == 0x4012d97c) return reduce_medium(x, ix, y0, y1);
            k = 3.0;
        } else {
            if (ix == 0x401921fb) return reduce_medium(x, ix, y0, y1);
            k = 4.0;
        }
        if (!neg) {
            const double z = x - k * PIO2_1;
            *y0 = z - k * PIO2_1T;
            *y1 = (z - *y0) - k * PIO2_1T;
            return (int)k;
        }
        const double z = x + k * PIO2_1;
        *y0 = z + k * PIO2_1T;
        *y1 = (z - *y0) + k * PIO2_1T;
        return -(int)k;
    }
    if (ix < 0x413921fb) return reduce_medium(x, ix, y0, y1); /* |x| ~< 2^20 pi/2 */
    return reduce_large(x, ix, y0, y1);
}

/* __cos(x, y) on [-pi/4, pi/4], y the tail of x */
static double kernel_cos(double x, double y) {
    static const double C1 = 4.16666666666666019037e-02, C2 = -1.38888888888741095749e-03,
                        C3 = 2.48015872894767294178e-05, C4 = -2.75573143513906633035e-07,
                        C5 = 2.08757232129817482790e-09, C6 = -1.13596475577881948265e-11;
    const double z = x * x;
    const double w = z * z;
    const double r = z * (C1 + z * (C2 + z * C3)) + w * w * (C4 + z * (C5 + z * C6));
    const double hz = 0.5 * z;
    const double v = 1.0 - hz;
    return v + (((1.0 - v) - hz) + (z * r - x * y));
}

/* __sin(x, y, iy) on [-pi/4, pi/4]; iy 0 means y is zero */
static double kernel_sin(double x, double y, int iy) {
    static const double S1 = -1.66666666666666324348e-01, S2 = 8.33333333332248946124e-03,
                        S3 = -1.98412698298579493134e-04, S4 = 2.75573137070700676789e-06,
                        S5 = -2.50507602534068634195e-08, S6 = 1.58969099521155010221e-10;
    const double z = x * x;
    const double w = z * z;
    const double r = S2 + z * (S3 + z * S4) + z * w * (S5 + z * S6);
    const double v = z * x;
    if (iy == 0) return x + v * (S1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S1);
}

double gro_libm_cos(double x) {
    const uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fb) {                 /* |x| ~< pi/4 */
        if (ix < 0x3e46a09e) return 1.0;    /* |x| < 2^-27 sqrt(2) */
        return kernel_cos(x, 0.0);
    }
    if (ix >= 0x7ff00000) return x - x;     /* inf or NaN */
    double y0, y1;
    const int n = reduce(x, &y0, &y1);
    switch (n & 3) {
    case 0: return kernel_cos(y0, y1);
    case 1: return -kernel_sin(y0, y1, 1);
    case 2: return -kernel_cos(y0, y1);
    default: return kernel_sin(y0, y1, 1);
    }
}

double gro_libm_sin(double x) {
    const uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fb) {                 /* |x| ~< pi/4 */
        if (ix < 0x3e500000) return x;      /* |x| < 2^-26 */
        return kernel_sin(x, 0.0, 0);
    }
    if (ix >= 0x7ff00000) return x - x;
    double y0, y1;
    const int n = reduce(x, &y0, &y1);
    switch (n & 3) {
    case 0: return kernel_sin(y0, y1, 1);
    case 1: return kernel_cos(y0, y1);
    case 2: return -kernel_sin(y0, y1, 1);
    default: return -kernel_cos(y0, y1);
    }
}
