/*
 * CPU oracle for layout.force.particle_mesh's two mesh passes, d3's law: TEST
 * INFRASTRUCTURE ONLY. Paths relative to crates/graph-core/src/layout/force/particle_mesh/.
 *
 *   charge   charge.rs:29-47 + mesh.rs:151-185,276-288: the frame over the positions'
 *            finite bounds (frame.rs:50-89, gro_pm_place), every node's scaled position in
 *            the grid's slot order (deposit.rs Stencils), a unit CIC deposit summed per cell
 *            in ascending slot (deposit.rs Rows/Deposit), the forward 2D transform over the
 *            live rows (fft/pass.rs), the kernel's spectrum (kernel.rs sample + forward), the
 *            inverse with the spectrum as gain, the CIC read (mesh.rs read) times
 *            charge * alpha; deltas in slot order, merged through the grid's slot.
 *   fft      fft.rs Plan: bit-reversal swaps, the first stage's (u + v, u - v), then stages
 *            half = 2, 4, ... with t = v * w from a table of libm's cos/sin (libm_trig.c).
 *   collide  collide.rs + collide/hash.rs + collide/gather.rs: the projected positions hashed
 *            into next_pow2(2n) buckets of cells `reach` wide (row hash cy * 0x9E3779B97F4A7C15
 *            >> shift, plus cx), a stable counting sort into the grid's order, and each slot's
 *            pushes from the 3x3 cells' buckets in read order, each once.
 * The mesh's grid order carries over between ticks (the charge pass deposits in the order the
 * last collide pass sorted), as the reference's Mesh keeps its Grid.
 *
 * Parity: pinned by the reference's fixtures/gpu/mesh-1k-{start,settled}.gmfx (kernel
 * spectrum, charge and collide deltas from a fresh mesh, bit for bit; tests/test_pm_d3_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fa2_oracle.h"

int gro_pm_place(double lox, double loy, double hix, double hiy, int side, double dmax,
                 int *step, double *h, double *ox, double *oy, int *cells, int *reach);

typedef struct {
    double re, im;
} cplx;

static cplx c_add(cplx a, cplx b) { return (cplx){a.re + b.re, a.im + b.im}; }
static cplx c_sub(cplx a, cplx b) { return (cplx){a.re - b.re, a.im - b.im}; }
static cplx c_mul(cplx a, cplx b) {
    return (cplx){a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

struct gro_pmd3 {
    uint32_t n;
    int side, bits;
    cplx *fwd, *inv;              /* fft.rs Plan twiddle tables */
    cplx *density, *scratch;      /* side * side */
    cplx *kspec;                  /* the kernel's spectrum */
    int k_valid, k_step, k_reach; /* what the spectrum was built for */
    double k_dmin2, k_dmax2;
    uint32_t *order, *slot;       /* the grid: slot -> node, node -> slot */
    double *ax, *ay;              /* scaled positions in slot order */
};

/* fft.rs Plan::new */
static void plan_init(gro_pmd3 *m) {
    const int side = m->side;
    cplx *root = (cplx *)malloc(sizeof(cplx) * (size_t)(side / 2 > 0 ? side / 2 : 1));
    const double tau = 6.283185307179586;
    for (int k = 0; k < side / 2; k++) {
        const double angle = -tau * (double)k / (double)side;
        root[k] = (cplx){gro_libm_cos(angle), gro_libm_sin(angle)};
    }
    m->fwd = (cplx *)malloc(sizeof(cplx) * side);
    m->inv = (cplx *)malloc(sizeof(cplx) * side);
    for (int k = 0; k < side; k++) {
        int half = 1;
        while (half * 2 <= (k > 1 ? k : 1)) half *= 2;
        m->fwd[k] = root[(k % half) * (side / (2 * half))];
        m->inv[k] = (cplx){m->fwd[k].re, -m->fwd[k].im};
    }
    free(root);
    m->bits = 0;
    while ((1 << m->bits) < side) m->bits++;
}

static uint32_t bitrev(uint32_t i, int bits) {
    uint32_t r = 0;
    for (int b = 0; b < bits; b++) r |= ((i >> b) & 1u) << (bits - 1 - b);
    return r;
}

/* fft.rs Plan::line */
static void line(const gro_pmd3 *m, cplx *a, int inverse) {
    const int side = m->side;
    for (int i = 0; i < side; i++) {
        const uint32_t j = bitrev((uint32_t)i, m->bits);
        if ((uint32_t)i < j) {
            const cplx t = a[i];
            a[i] = a[j];
            a[j] = t;
        }
    }
    for (int q = 0; q + 1 < side; q += 2) {
        const cplx u = a[q], v = a[q + 1];
        a[q] = c_add(u, v);
        a[q + 1] = c_sub(u, v);
    }
    const cplx *tw = inverse ? m->inv : m->fwd;
    for (int half = 2; half < side; half *= 2)
        for (int b = 0; b < side; b += 2 * half)
            for (int j = 0; j < half; j++) {
                const cplx u = a[b + j], t = c_mul(a[b + j + half], tw[half + j]);
                a[b + j] = c_add(u, t);
                a[b + j + half] = c_sub(u, t);
            }
}

/* fft/pass.rs Pass: out row r is the transform of src row r (read 0), of src row r times
 * gain row r (read 1) or of src column r (read 2), for r < live; zero past live */
static void pass2d(const gro_pmd3 *m, const cplx *src, int read, const cplx *gain, int inverse,
                   int live, cplx *out) {
    const int side = m->side;
    for (int r = 0; r < side; r++) {
        cplx *dst = out + (size_t)r * side;
        if (r >= live) {
            for (int c = 0; c < side; c++) dst[c] = (cplx){0.0, 0.0};
            continue;
        }
        for (int c = 0; c < side; c++) {
            if (read == 0)
                dst[c] = src[(size_t)r * side + c];
            else if (read == 1)
                dst[c] = c_mul(src[(size_t)r * side + c], gain[(size_t)r * side + c]);
            else
                dst[c] = src[(size_t)c * side + r];
        }
        line(m, dst, inverse);
    }
}

/* pass.rs Fft::forward: rows over `live`, then every column; the result lands in a */
static void forward2d(const gro_pmd3 *m, cplx *a, cplx *b, int live) {
    pass2d(m, a, 0, NULL, 0, live, b);
    pass2d(m, b, 2, NULL, 0, m->side, a);
}

/* pass.rs Fft::inverse: every row times the gain, then the columns over `live` */
static void inverse2d(const gro_pmd3 *m, cplx *a, cplx *b, const cplx *gain, int live) {
    pass2d(m, a, 1, gain, 1, m->side, b);
    pass2d(m, b, 2, NULL, 1, live, a);
}

gro_pmd3 *gro_pmd3_new(uint32_t n, int side) {
    gro_pmd3 *m = (gro_pmd3 *)calloc(1, sizeof(gro_pmd3));
    m->n = n;
    m->side = side;
    plan_init(m);
    const size_t p2 = (size_t)side * side, nn = n ? n : 1;
    m->density = (cplx *)calloc(p2, sizeof(cplx));
    m->scratch = (cplx *)calloc(p2, sizeof(cplx));
    m->kspec = (cplx *)calloc(p2, sizeof(cplx));
    m->order = (uint32_t *)malloc(nn * sizeof(uint32_t));
    m->slot = (uint32_t *)malloc(nn * sizeof(uint32_t));
    m->ax = (double *)calloc(nn, sizeof(double));
    m->ay = (double *)calloc(nn, sizeof(double));
    for (uint32_t i = 0; i < n; i++) m->order[i] = m->slot[i] = i; /* collide.rs Grid::new */
    return m;
}

void gro_pmd3_free(gro_pmd3 *m) {
    if (!m) return;
    free(m->fwd);
    free(m->inv);
    free(m->density);
    free(m->scratch);
    free(m->kspec);
    free(m->order);
    free(m->slot);
    free(m->ax);
    free(m->ay);
    free(m);
}

/* frame.rs bounds: the finite positions' extent; 0 when there is none */
static int finite_bounds(uint32_t n, const double *x, const double *y, double *lo, double *hi) {
    lo[0] = lo[1] = INFINITY;
    hi[0] = hi[1] = -INFINITY;
    for (uint32_t i = 0; i < n; i++) {
        if (!(isfinite(x[i]) && isfinite(y[i]))) continue;
        if (x[i] < lo[0]) lo[0] = x[i];
        if (y[i] < lo[1]) lo[1] = y[i];
        if (x[i] > hi[0]) hi[0] = x[i];
        if (y[i] > hi[1]) hi[1] = y[i];
    }
    return lo[0] <= hi[0];
}

/* frame.rs cell: u as usize (saturating: NaN and negatives to 0), at most cells - 2 */
static size_t frame_cell(double u, int cells) {
    size_t c = u > 0.0 ? (u < 1e18 ? (size_t)u : (size_t)1e18) : 0;
    return c < (size_t)(cells - 2) ? c : (size_t)(cells - 2);
}

/* kernel.rs green */
static cplx green(double rx, double ry, double dmin2, double dmax2, double scale) {
    double l = rx * rx + ry * ry;
    if (l == 0.0 || l >= dmax2) return (cplx){0.0, 0.0};
    if (l < dmin2) l = sqrt(dmin2 * l);
    return (cplx){-rx / l * scale, -ry / l * scale};
}

int gro_pmd3_charge(gro_pmd3 *m, const double *x, const double *y, double dmin, double dmax,
                    double strength, double *out, double *frame_out) {
    const uint32_t n = m->n;
    const int side = m->side;
    const double dmin2 = dmin * dmin, dmax2 = dmax * dmax;
    double lo[2], hi[2];
    int step = 0, cells = 0, reach = 0;
    double h = 0.0, ox = 0.0, oy = 0.0;
    if (frame_out)
        for (int k = 0; k < 7; k++) frame_out[k] = 0.0;
    /* mesh.rs solve: no frame, one node or no reach: the pass adds nothing */
    if (!finite_bounds(n, x, y, lo, hi) || n <= 1 || !(dmax2 > 0.0) ||
        !gro_pm_place(lo[0], lo[1], hi[0], hi[1], side, sqrt(dmax2), &step, &h, &ox, &oy, &cells,
                      &reach))
        return 0;
    if (frame_out) {
        frame_out[0] = step;
        frame_out[1] = h;
        frame_out[2] = ox;
        frame_out[3] = oy;
        frame_out[4] = cells;
        frame_out[5] = reach;
        frame_out[6] = side;
    }
    /* deposit.rs Stencils: scaled positions in slot order (NaN for a non-finite node) */
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t i = m->order[k];
        if (isfinite(x[i]) && isfinite(y[i])) {
            m->ax[k] = (x[i] - ox) / h;
            m->ay[k] = (y[i] - oy) / h;
        } else {
            m->ax[k] = m->ay[k] = NAN;
        }
    }
    /* deposit.rs Deposit: each cell's weights in ascending slot, from zero */
    const size_t p2 = (size_t)side * side;
    for (size_t q = 0; q < p2; q++) m->density[q] = (cplx){0.0, 0.0};
    for (uint32_t k = 0; k < n; k++) {
        if (isnan(m->ax[k])) continue;
        const size_t cx = frame_cell(m->ax[k], cells), cy = frame_cell(m->ay[k], cells);
        double fx = m->ax[k] - (double)cx, fy = m->ay[k] - (double)cy;
        fx = fx < 0.0 ? 0.0 : (fx > 1.0 ? 1.0 : fx);
        fy = fy < 0.0 ? 0.0 : (fy > 1.0 ? 1.0 : fy);
        const size_t at = cy * side + cx;
        m->density[at].re += (1.0 - fx) * (1.0 - fy);
        m->density[at + 1].re += fx * (1.0 - fy);
        m->density[at + side].re += (1.0 - fx) * fy;
        m->density[at + side + 1].re += fx * fy;
    }
    /* kernel.rs refresh: sampled and transformed when the rung, reach or law changed */
    if (!m->k_valid || m->k_step != step || m->k_reach != reach || m->k_dmin2 != dmin2 ||
        m->k_dmax2 != dmax2) {
        const double scale = 1.0 / (double)p2;
        for (size_t q = 0; q < p2; q++) m->kspec[q] = (cplx){0.0, 0.0};
        for (int dy = -reach; dy <= reach; dy++) {
            const size_t row = (size_t)(((dy % side) + side) % side) * side;
            for (int dx = -reach; dx <= reach; dx++)
                m->kspec[row + (size_t)(((dx % side) + side) % side)] =
                    green((double)dx * h, (double)dy * h, dmin2, dmax2, scale);
        }
        forward2d(m, m->kspec, m->scratch, side);
        m->k_valid = 1;
        m->k_step = step;
        m->k_reach = reach;
        m->k_dmin2 = dmin2;
        m->k_dmax2 = dmax2;
    }
    forward2d(m, m->density, m->scratch, cells);
    inverse2d(m, m->density, m->scratch, m->kspec, cells);
    /* charge.rs Interpolate + mesh.rs read, merged through the slots: node order out */
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t i = m->order[k];
        double ex = 0.0, ey = 0.0;
        if (!isnan(m->ax[k])) {
            const size_t cx = frame_cell(m->ax[k], cells), cy = frame_cell(m->ay[k], cells);
            double fx = m->ax[k] - (double)cx, fy = m->ay[k] - (double)cy;
            fx = fx < 0.0 ? 0.0 : (fx > 1.0 ? 1.0 : fx);
            fy = fy < 0.0 ? 0.0 : (fy > 1.0 ? 1.0 : fy);
            const size_t at = cy * side + cx;
            const size_t cell[4] = {at, at + 1, at + side, at + side + 1};
            const double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy,
                                 fx * fy};
            for (int q = 0; q < 4; q++) {
                ex += m->density[cell[q]].re * w[q];
                ey += m->density[cell[q]].im * w[q];
            }
        }
        out[2 * (size_t)i] = ex * strength;
        out[2 * (size_t)i + 1] = ey * strength;
    }
    return 1;
}

void gro_pmd3_spectrum(const gro_pmd3 *m, double *re, double *im) {
    const size_t p2 = (size_t)m->side * m->side;
    for (size_t q = 0; q < p2; q++) {
        re[q] = m->kspec[q].re;
        im[q] = m->kspec[q].im;
    }
}

const uint32_t *gro_pmd3_order(const gro_pmd3 *m) { return m->order; }

/* ---- the hashed grid (collide/hash.rs) ---- */

/* Rust's `f as i64`: truncation, NaN to 0, saturating at the ends */
static int64_t sat_i64(double v) {
    if (isnan(v)) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

typedef struct {
    uint32_t shift;
    uint64_t mask;
    double ox, oy, size;
} hash_t;

static void cell_of(const hash_t *hs, double x, double y, int64_t *cx, int64_t *cy) {
    *cx = sat_i64((x - hs->ox) / hs->size);
    *cy = sat_i64((y - hs->oy) / hs->size);
}

static uint32_t bucket_of(const hash_t *hs, int64_t cx, int64_t cy) {
    const uint64_t row = ((uint64_t)cy * 0x9E3779B97F4A7C15ull) >> hs->shift;
    return (uint32_t)((row + (uint64_t)cx) & hs->mask);
}

int gro_pmd3_collide(gro_pmd3 *m, const double *px, const double *py, double radius,
                     uint32_t seed, uint32_t tick, double *out) {
    const uint32_t n = m->n;
    const double diameter = 2.0 * radius;
    const double d2 = diameter * diameter;
    if (d2 == 0.0) return 0; /* collide.rs apply: no reach, no pass */
    const double reach = sqrt(d2);
    /* Grid::build: origin at the finite bounds' low corner, cells `reach` wide */
    size_t buckets = 4;
    while (buckets < 2 * (size_t)n) buckets <<= 1;
    int lg = 0;
    while (((size_t)1 << lg) < buckets) lg++;
    hash_t hs;
    hs.shift = 64u - (uint32_t)lg;
    hs.mask = (uint64_t)buckets - 1;
    double lo[2], hi[2];
    if (finite_bounds(n, px, py, lo, hi)) {
        hs.ox = lo[0];
        hs.oy = lo[1];
    } else {
        hs.ox = hs.oy = 0.0;
    }
    hs.size = reach;
    uint32_t *start = (uint32_t *)calloc(buckets + 2, sizeof(uint32_t));
    uint32_t *bk = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++) {
        int64_t cx, cy;
        cell_of(&hs, px[i], py[i], &cx, &cy);
        bk[i] = bucket_of(&hs, cx, cy);
        start[bk[i] + 2]++;
    }
    for (size_t b = 1; b < buckets + 2; b++) start[b] += start[b - 1];
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t k = start[bk[i] + 1]++;
        m->order[k] = i;
        m->slot[i] = k;
    }
    /* gather.rs: each slot's pushes, candidates in read order */
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t i = m->order[k];
        const double xk = px[i], yk = py[i];
        int64_t cx, cy;
        cell_of(&hs, xk, yk, &cx, &cy);
        uint32_t run_lo[9], run_hi[9];
        int runs = 0;
        uint64_t firsts[3] = {0, 0, 0};
        for (int r = 0; r < 3; r++) {
            const int64_t dy = r - 1;
            const uint64_t first = bucket_of(&hs, (int64_t)((uint64_t)cx - 1u),
                                             (int64_t)((uint64_t)cy + (uint64_t)dy));
            for (uint64_t bb = first; bb < first + 3; bb++) {
                const uint64_t b = bb & hs.mask;
                int seen = 0;
                for (int e = 0; e < r; e++) seen |= ((b - firsts[e]) & hs.mask) < 3;
                if (seen) continue;
                const uint32_t a = start[b], z = start[b + 1];
                if (a == z) continue;
                if (runs > 0 && run_hi[runs - 1] == a) {
                    run_hi[runs - 1] = z;
                } else {
                    run_lo[runs] = a;
                    run_hi[runs] = z;
                    runs++;
                }
            }
            firsts[r] = first;
        }
        double ox_ = 0.0, oy_ = 0.0;
        for (int r = 0; r < runs; r++)
            for (uint32_t q = run_lo[r]; q < run_hi[r]; q++) {
                if (q == k) continue;
                const uint32_t j = m->order[q];
                double dx = xk - px[j], dy = yk - py[j];
                double l = dx * dx + dy * dy;
                if (!(l < d2)) continue; /* NaN or out of reach */
                if (dx == 0.0) {
                    dx = gro_jiggle(seed, tick, 4, i, j);
                    l += dx * dx;
                }
                if (dy == 0.0) {
                    dy = gro_jiggle(seed, tick, 5, i, j);
                    l += dy * dy;
                }
                const double dist = sqrt(l);
                const double push = (reach - dist) / dist * 0.5;
                ox_ += dx * push;
                oy_ += dy * push;
            }
        out[2 * (size_t)i] = ox_;
        out[2 * (size_t)i + 1] = oy_;
    }
    free(start);
    free(bk);
    return 1;
}

/* One charge pass from a fresh mesh (the identity grid order unless `order` is given) */
int gro_pmd3_charge_once(uint32_t n, const double *xy, int side, const uint32_t *order,
                         double charge, double alpha, double dmin, double dmax, double *out,
                         double *frame_out, double *spec_re, double *spec_im) {
    gro_pmd3 *m = gro_pmd3_new(n, side);
    if (order)
        for (uint32_t k = 0; k < n; k++) {
            m->order[k] = order[k];
            m->slot[order[k]] = k;
        }
    double *x = (double *)malloc((n ? n : 1) * sizeof(double));
    double *y = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        x[i] = xy[2 * (size_t)i];
        y[i] = xy[2 * (size_t)i + 1];
    }
    for (uint32_t i = 0; i < 2 * n; i++) out[i] = 0.0;
    const int r = gro_pmd3_charge(m, x, y, dmin, dmax, charge * alpha, out, frame_out);
    if (r && spec_re && spec_im) gro_pmd3_spectrum(m, spec_re, spec_im);
    free(x);
    free(y);
    gro_pmd3_free(m);
    return r;
}

/* One collide pass over these positions: deltas in node order and the grid order it sorted */
int gro_pmd3_collide_once(uint32_t n, const double *xy, double radius, uint32_t seed,
                          uint32_t tick, double *out, uint32_t *order_out) {
    gro_pmd3 *m = gro_pmd3_new(n, 128);
    double *x = (double *)malloc((n ? n : 1) * sizeof(double));
    double *y = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        x[i] = xy[2 * (size_t)i];
        y[i] = xy[2 * (size_t)i + 1];
    }
    for (uint32_t i = 0; i < 2 * n; i++) out[i] = 0.0;
    const int r = gro_pmd3_collide(m, x, y, radius, seed, tick, out);
    if (order_out) memcpy(order_out, m->order, (size_t)n * sizeof(uint32_t));
    free(x);
    free(y);
    gro_pmd3_free(m);
    return r;
}
