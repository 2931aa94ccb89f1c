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
 * against pi computed exactly. Only the medium reduction (|x| < 2^20 pi/2) is restated:
 * larger arguments return NaN here (the Payne-Hanek path is parity unpinned).
 */
#include <stdint.h>
#include <string.h>

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

/* rem_pio2: the quadrant n and y0 + y1 = x - n pi/2; -1000 for arguments past the medium
 * range */
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
    return -1000;
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
    if (n == -1000) return __builtin_nan("");
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
    if (n == -1000) return __builtin_nan("");
    switch (n & 3) {
    case 0: return kernel_sin(y0, y1, 1);
    case 1: return kernel_cos(y0, y1);
    case 2: return -kernel_sin(y0, y1, 1);
    default: return -kernel_cos(y0, y1);
    }
}
