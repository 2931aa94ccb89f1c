/*
 * CPU oracle for the particle-mesh repulsion: TEST INFRASTRUCTURE ONLY.
 *
 * Two laws share one mesh solve (frame, CIC deposit, FFT convolution, CIC read):
 *   gro_pm_forces  FA2's repulsion, described below;
 *   gro_pm_charge  the reference's own d3 many-body law (particle_mesh/kernel.rs:1-12,
 *                  80-92; charge.rs:29-47): unit charge per node, G(r) = -r/l with l raised
 *                  to sqrt(dmin^2 l) below dmin^2 and zero at l >= dmax^2, reach
 *                  ceil(dmax/h) (frame.rs:70-80), read times charge * alpha. Pinned against
 *                  the reference's fixtures/gpu/mesh-1k-{start,settled}.gmfx (frame and
 *                  charge deltas, tests/test_pm_oracle.py).
 *
 * FA2's repulsion k*m_i*m_j*(p_i - p_j)/|p_i - p_j|^2 evaluated on a mesh, in FP64:
 *   frame   the rung ladder of crates/graph-core/src/layout/force/particle_mesh/frame.rs:50-89
 *           (h = 2^(step/4), cells = floor(span/h) + 3, origin snapped to a multiple of h)
 *           with the FA2 law's infinite reach, reach = cells - 1, so cells + reach <= side
 *           keeps the periodic transform free of wrap-around between occupied cells;
 *   deposit CIC of each node's mass (particle_mesh/deposit.rs weights), node order;
 *   kernel  G(r) = r / |r|^2, G(0) = 0, sampled at cell offsets within reach, wrapped,
 *           scaled by 1/side^2 (particle_mesh/kernel.rs:64-92 with the sign and cutoffs of
 *           the FA2 law: repulsive, no dmin/dmax);
 *   solve   FFT2(density) * FFT2(G) -> inverse FFT2, Ex in the real part, Ey in the
 *           imaginary (the reference's packed-kernel trick);
 *   read    CIC gather of (Ex, Ey) at each node, times k * m_i.
 * Parity status: the reference's PM fixtures (fixtures/gpu, .gmfx files) are for d3's law only;
 * this FA2-law arm is pinned by the FFT against a direct DFT, by the mesh field against the
 * exact pair sum at range, and by the vanishing self-force (tests/test_pm_oracle.py).
 */
#define _USE_MATH_DEFINES
#include <math.h>
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* 2^(q/4), q = 0..3, correctly rounded; h = 2^(step/4) is built from these exactly */
static const double QUARTER[4] = {1.0, 1.189207115002721, 1.4142135623730951,
                                  1.681792830507429};

double gro_pm_h(int step) { return ldexp(QUARTER[step & 3], step >> 2); }

typedef struct {
    int step, cells, reach;
    double h, ox, oy;
} pm_frame;

/* frame.rs:50-89: the smallest rung whose cells and reach (ceil(dmax/h), at most cells - 1;
 * FA2's infinite reach is cells - 1) fit the side. Returns 0 when no frame exists. */
static int place(double lox, double loy, double hix, double hiy, int side, double dmax,
                 pm_frame *f) {
    double span = fmax(hix - lox, hiy - loy);
    if (!isfinite(span)) return 0;
    double st = floor(4.0 * log2(span / (double)side));
    int step = st < -32.0 ? -32 : (int)st;
    for (;; step++) {
        double h = gro_pm_h(step);
        double cells = floor(span / h) + 3.0;
        if (cells > (double)side) continue;
        int c = (int)cells;
        double r = ceil(dmax / h);
        int reach = r < (double)(c - 1) ? (int)r : c - 1;
        if (c + reach > side) continue;
        f->step = step;
        f->h = h;
        f->cells = c;
        f->reach = reach;
        f->ox = floor(lox / h) * h;
        f->oy = floor(loy / h) * h;
        return 1;
    }
}

/* place() for the d3-law mesh of pm_d3_oracle.c: 1 and the frame's fields when one exists */
int gro_pm_place(double lox, double loy, double hix, double hiy, int side, double dmax,
                 int *step, double *h, double *ox, double *oy, int *cells, int *reach) {
    pm_frame f;
    if (!place(lox, loy, hix, hiy, side, dmax, &f)) return 0;
    *step = f.step;
    *h = f.h;
    *ox = f.ox;
    *oy = f.oy;
    *cells = f.cells;
    *reach = f.reach;
    return 1;
}

/* frame.rs:160-187: scaled position, cell, fraction. */
static void stencil(const pm_frame *f, double px, double py, int *cx, int *cy, double *fx,
                    double *fy) {
    double ux = (px - f->ox) / f->h, uy = (py - f->oy) / f->h;
    long ix = ux > 0.0 ? (long)ux : 0, iy = uy > 0.0 ? (long)uy : 0;
    if (ix > f->cells - 2) ix = f->cells - 2;
    if (iy > f->cells - 2) iy = f->cells - 2;
    double gx = ux - (double)ix, gy = uy - (double)iy;
    *fx = gx < 0.0 ? 0.0 : (gx > 1.0 ? 1.0 : gx);
    *fy = gy < 0.0 ? 0.0 : (gy > 1.0 ? 1.0 : gy);
    *cx = (int)ix;
    *cy = (int)iy;
}

/* in-place iterative radix-2 complex FFT of length n (power of two), sign -1 forward */
static void fft1(double *re, double *im, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i];
            re[i] = re[j];
            re[j] = t;
            t = im[i];
            im[i] = im[j];
            im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        double ang = sign * 2.0 * M_PI / (double)len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < len / 2; k++) {
                double wr = cos(ang * k), wi = sin(ang * k);
                double ar = re[i + k], ai = im[i + k];
                double br = re[i + k + len / 2], bi = im[i + k + len / 2];
                double tr = br * wr - bi * wi, ti = br * wi + bi * wr;
                re[i + k] = ar + tr;
                im[i + k] = ai + ti;
                re[i + k + len / 2] = ar - tr;
                im[i + k + len / 2] = ai - ti;
            }
        }
    }
}

void gro_fft2(double *re, double *im, int side, int sign) {
    double *tr = (double *)malloc(sizeof(double) * side), *ti = (double *)malloc(sizeof(double) * side);
    for (int r = 0; r < side; r++) fft1(re + (size_t)r * side, im + (size_t)r * side, side, sign);
    for (int c = 0; c < side; c++) {
        for (int r = 0; r < side; r++) {
            tr[r] = re[(size_t)r * side + c];
            ti[r] = im[(size_t)r * side + c];
        }
        fft1(tr, ti, side, sign);
        for (int r = 0; r < side; r++) {
            re[(size_t)r * side + c] = tr[r];
            im[(size_t)r * side + c] = ti[r];
        }
    }
    free(tr);
    free(ti);
}

/* The mesh field E at every node: deposit `charge` (NULL: unit charges) with CIC weights,
 * convolve with the law (0: FA2's r/|r|^2; 1: d3's -r/l with its cutoffs), read with CIC.
 * Returns 1 if a frame was placed (e is 2n), else 0 (e zeroed). */
static int pm_field(uint32_t n, const double *xy, const double *charge, int side, int law,
                    double dmin2, double dmax2, double *e, double *frame_out) {
    memset(e, 0, sizeof(double) * 2 * n);
    double lox = INFINITY, loy = INFINITY, hix = -INFINITY, hiy = -INFINITY;
    uint32_t nf = 0;
    for (uint32_t i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!isfinite(x) || !isfinite(y)) continue;
        lox = fmin(lox, x);
        loy = fmin(loy, y);
        hix = fmax(hix, x);
        hiy = fmax(hiy, y);
        nf++;
    }
    pm_frame f;
    const double dmax = law == 1 ? sqrt(dmax2) : INFINITY;
    if (n < 2 || nf == 0 || (law == 1 && !(dmax2 > 0.0)) ||
        !place(lox, loy, hix, hiy, side, dmax, &f))
        return 0;
    if (frame_out) {
        frame_out[0] = f.h;
        frame_out[1] = f.ox;
        frame_out[2] = f.oy;
        frame_out[3] = f.step;
        frame_out[4] = f.cells;
        frame_out[5] = f.reach;
    }
    size_t P2 = (size_t)side * side;
    double *dre = (double *)calloc(P2, sizeof(double)), *dim_ = (double *)calloc(P2, sizeof(double));
    double *gre = (double *)calloc(P2, sizeof(double)), *gim = (double *)calloc(P2, sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!isfinite(x) || !isfinite(y)) continue;
        int cx, cy;
        double fx, fy;
        stencil(&f, x, y, &cx, &cy, &fx, &fy);
        double m = charge ? charge[i] : 1.0;
        size_t at = (size_t)cy * side + cx;
        dre[at] += m * ((1.0 - fx) * (1.0 - fy));
        dre[at + 1] += m * (fx * (1.0 - fy));
        dre[at + side] += m * ((1.0 - fx) * fy);
        dre[at + side + 1] += m * (fx * fy);
    }
    double scale = 1.0 / (double)P2;
    for (int dy = -f.reach; dy <= f.reach; dy++) {
        for (int dx = -f.reach; dx <= f.reach; dx++) {
            double rx = dx * f.h, ry = dy * f.h;
            double l = rx * rx + ry * ry;
            size_t at = (size_t)((dy + side) % side) * side + (size_t)((dx + side) % side);
            if (law == 1) {
                if (l == 0.0 || l >= dmax2) continue;
                if (l < dmin2) l = sqrt(dmin2 * l);
                gre[at] = -rx / l * scale;
                gim[at] = -ry / l * scale;
            } else {
                if (l == 0.0) continue;
                gre[at] = rx / l * scale;
                gim[at] = ry / l * scale;
            }
        }
    }
    gro_fft2(dre, dim_, side, -1);
    gro_fft2(gre, gim, side, -1);
    for (size_t q = 0; q < P2; q++) {
        double ar = dre[q], ai = dim_[q];
        dre[q] = ar * gre[q] - ai * gim[q];
        dim_[q] = ar * gim[q] + ai * gre[q];
    }
    gro_fft2(dre, dim_, side, +1);
    for (uint32_t i = 0; i < n; i++) {
        double x = xy[2 * i], y = xy[2 * i + 1];
        if (!isfinite(x) || !isfinite(y)) {
            e[2 * i] = e[2 * i + 1] = NAN;
            continue;
        }
        int cx, cy;
        double fx, fy;
        stencil(&f, x, y, &cx, &cy, &fx, &fy);
        size_t at = (size_t)cy * side + cx;
        double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy};
        size_t cell[4] = {at, at + 1, at + side, at + side + 1};
        double ex = 0.0, ey = 0.0;
        for (int q = 0; q < 4; q++) {
            ex += dre[cell[q]] * w[q];
            ey += dim_[cell[q]] * w[q];
        }
        e[2 * i] = ex;
        e[2 * i + 1] = ey;
    }
    free(dre);
    free(dim_);
    free(gre);
    free(gim);
    return 1;
}

/* FA2: out[i] = k*m_i*E(p_i), mass deposited; frame_out (may be NULL): h, ox, oy, step,
 * cells, reach. Returns 1 if a frame was placed, 0 otherwise (out is zeroed). */
int gro_pm_forces(uint32_t n, const double *xy, const double *mass, double k, int side,
                  double *out, double *frame_out) {
    int placed = pm_field(n, xy, mass, side, 0, 0.0, 0.0, out, frame_out);
    if (!placed) return 0;
    for (uint32_t i = 0; i < n; i++) {
        out[2 * i] *= k * mass[i];
        out[2 * i + 1] *= k * mass[i];
    }
    return 1;
}

/* d3: out[i] = charge*alpha*E(p_i), unit charges deposited (charge.rs:29-47). */
int gro_pm_charge(uint32_t n, const double *xy, int side, double charge, double alpha,
                  double dmin, double dmax, double *out, double *frame_out) {
    int placed = pm_field(n, xy, NULL, side, 1, dmin * dmin, dmax * dmax, out, frame_out);
    if (!placed) return 0;
    const double strength = charge * alpha;
    for (uint32_t i = 0; i < 2 * n; i++) out[i] *= strength;
    return 1;
}
