/* Test infrastructure only: a CPU restatement of the reference's layout.forceatlas2.forcesim
 * (SciGraphs' own ForceSim, model FA2, direct repulsion), the checker the device path is
 * compared against. Paths are relative to crates/graph-core/src/ of the reference.
 *
 *   start      rng/pcg64.rs + rng/pcg64/seed_sequence.rs (numpy's default_rng: SeedSequence
 *              hashing, PCG-XSL-RR 128/64, 53-bit doubles), forcesim/seed.rs
 *   constants  forcesim/state.rs:117-132 (renormalize)
 *   forces     forcesim/pair_force.rs (direct, f64 coefficient rows), attraction.rs
 *              (bincount order, f64 accumulators), gravity.rs
 *   step       forcesim/integrate.rs (swing/traction, adaptive speed, capped move,
 *              recentring), state.rs:100-113
 *   answer     forcesim/rescale.rs, forcesim.rs:158-187 (NonFinite refusal, f32 columns)
 *   orders     forcesim/reduce.rs (numpy's pairwise f32/f64 sums, sequential column means,
 *              the f32 three-term dot)
 *
 * Parity is pinned by the reference's own measured bit vectors (forcesim/tests.rs,
 * rng/pcg64/tests.rs, forcesim/seed.rs, forcesim/reduce.rs) and by numpy's default_rng. */
#include "forcesim_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ numpy's default_rng */

/* numpy's SeedSequence (bit_generator.pyx): a pool of four 32-bit words mixed from the
 * entropy words (the seed split into 32-bit words, lowest first), then generate_state. */
#define SS_INIT_A 0x43b0d7e5u
#define SS_MULT_A 0x931e8875u
#define SS_INIT_B 0x8b51f9ddu
#define SS_MULT_B 0x58f38dedu
#define SS_MIX_L 0xca01f9ddu
#define SS_MIX_R 0x4973f715u

static uint32_t ss_hashmix(uint32_t value, uint32_t *hash_const) {
    value ^= *hash_const;
    *hash_const *= SS_MULT_A;
    value *= *hash_const;
    value ^= value >> 16;
    return value;
}

static uint32_t ss_mix(uint32_t x, uint32_t y) {
    uint32_t r = SS_MIX_L * x - SS_MIX_R * y;
    r ^= r >> 16;
    return r;
}

/* generate_state(4, uint64) of SeedSequence(seed): words[0..4] */
static void seed_sequence_words(uint64_t seed, uint64_t words[4]) {
    uint32_t entropy[2];
    int ne = 0;
    uint64_t s = seed;
    do {
        entropy[ne++] = (uint32_t)s;
        s >>= 32;
    } while (s && ne < 2);
    uint32_t pool[4];
    uint32_t hc = SS_INIT_A;
    for (int i = 0; i < 4; i++) pool[i] = ss_hashmix(i < ne ? entropy[i] : 0u, &hc);
    for (int src = 0; src < 4; src++)
        for (int dst = 0; dst < 4; dst++)
            if (src != dst) pool[dst] = ss_mix(pool[dst], ss_hashmix(pool[src], &hc));
    /* no entropy word beyond the pool: a u64 seed is at most two words */
    uint32_t out[8];
    uint32_t hb = SS_INIT_B;
    for (int i = 0; i < 8; i++) {
        uint32_t v = pool[i % 4];
        v ^= hb;
        hb *= SS_MULT_B;
        v *= hb;
        v ^= v >> 16;
        out[i] = v;
    }
    for (int i = 0; i < 4; i++) words[i] = (uint64_t)out[2 * i] | ((uint64_t)out[2 * i + 1] << 32);
}

typedef unsigned __int128 u128;

typedef struct {
    u128 state, inc;
} pcg64_t;

static const u128 PCG_MULT = ((u128)2549297995355413924ull << 64) | (u128)4865540595714422341ull;

static void pcg_step(pcg64_t *r) { r->state = r->state * PCG_MULT + r->inc; }

/* pcg64_set_seed: initstate = w0:w1 (high:low), initseq = w2:w3; state 0, inc = seq<<1|1,
 * step, add initstate, step */
static void pcg64_seed(pcg64_t *r, uint64_t seed) {
    uint64_t w[4];
    seed_sequence_words(seed, w);
    const u128 initstate = ((u128)w[0] << 64) | w[1];
    const u128 initseq = ((u128)w[2] << 64) | w[3];
    r->state = 0;
    r->inc = (initseq << 1) | 1u;
    pcg_step(r);
    r->state += initstate;
    pcg_step(r);
}

static uint64_t pcg64_next(pcg64_t *r) {
    pcg_step(r);
    const uint64_t hi = (uint64_t)(r->state >> 64), lo = (uint64_t)r->state;
    const uint64_t x = hi ^ lo;
    const unsigned rot = (unsigned)(hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

static double pcg64_double(pcg64_t *r) {
    return (double)(pcg64_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

void gro_pcg64_state(uint64_t seed, uint64_t out4[4]) {
    pcg64_t r;
    pcg64_seed(&r, seed);
    out4[0] = (uint64_t)(r.state >> 64);
    out4[1] = (uint64_t)r.state;
    out4[2] = (uint64_t)(r.inc >> 64);
    out4[3] = (uint64_t)r.inc;
}

void gro_pcg64_raw(uint64_t seed, uint32_t count, uint64_t *out) {
    pcg64_t r;
    pcg64_seed(&r, seed);
    for (uint32_t i = 0; i < count; i++) out[i] = pcg64_next(&r);
}

void gro_pcg64_doubles(uint64_t seed, uint32_t count, double *out) {
    pcg64_t r;
    pcg64_seed(&r, seed);
    for (uint32_t i = 0; i < count; i++) out[i] = pcg64_double(&r);
}

/* ------------------------------------------------------------------ numpy's orders */

/* pairwise_sum_FLOAT / _DOUBLE: under 8 sequential; up to 128 eight lanes, a balanced tree
 * over them, then the remainder folded in one value at a time; above, split at half rounded
 * down to a multiple of 8 (reduce.rs:27-91) */
float gro_pairwise_f32(const float *v, size_t n) {
    if (n < 8) {
        float acc = 0.0f;
        for (size_t i = 0; i < n; i++) acc += v[i];
        return acc;
    }
    if (n <= 128) {
        float l[8];
        for (int q = 0; q < 8; q++) l[q] = v[q];
        const size_t stop = n - n % 8;
        for (size_t i = 8; i < stop; i += 8)
            for (int q = 0; q < 8; q++) l[q] += v[i + q];
        float t = ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7]));
        for (size_t i = stop; i < n; i++) t += v[i];
        return t;
    }
    size_t split = n / 2;
    split -= split % 8;
    return gro_pairwise_f32(v, split) + gro_pairwise_f32(v + split, n - split);
}

double gro_pairwise_f64(const double *v, size_t n) {
    if (n < 8) {
        double acc = 0.0;
        for (size_t i = 0; i < n; i++) acc += v[i];
        return acc;
    }
    if (n <= 128) {
        double l[8];
        for (int q = 0; q < 8; q++) l[q] = v[q];
        const size_t stop = n - n % 8;
        for (size_t i = 8; i < stop; i += 8)
            for (int q = 0; q < 8; q++) l[q] += v[i + q];
        double t = ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7]));
        for (size_t i = stop; i < n; i++) t += v[i];
        return t;
    }
    size_t split = n / 2;
    split -= split % 8;
    return gro_pairwise_f64(v, split) + gro_pairwise_f64(v + split, n - split);
}

/* mean(axis=0) of an (n, 3) array: sequential per column, then the division (reduce.rs:93-125) */
static void column_means_f32(const float *p, size_t n, float out[3]) {
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) s[c] += p[3 * i + c];
    const float cnt = (float)n;
    for (int c = 0; c < 3; c++) out[c] = s[c] / cnt;
}

static void column_means_f64(const double *p, size_t n, double out[3]) {
    double s[3] = {0.0, 0.0, 0.0};
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 3; c++) s[c] += p[3 * i + c];
    const double cnt = (double)n;
    for (int c = 0; c < 3; c++) out[c] = s[c] / cnt;
}

/* einsum("ij,ij->i") in f32: ((a0 b0 + a1 b1) + a2 b2) (reduce.rs:127-152) */
float gro_dot3_f32(const float *a, const float *b) {
    const float first = a[0] * b[0];
    const float second = a[1] * b[1];
    return (first + second) + a[2] * b[2];
}

/* ------------------------------------------------------------------ the simulation */

/* The correctly rounded cube root of an integer-valued x in [1, 2^53]: the reference's k
 * takes libm 0.2.16's cbrt (Cargo.lock:271-273), CORE-MATH's correctly rounded cube root;
 * the C library's may be an ulp off (glibc's cbrt(27) is 3.0000000000000004). Start from
 * the C library's answer and move to a neighbour while the exact cube of the midpoint on
 * that side says the root lies beyond it. */
typedef unsigned __int128 u128_t;

/* sign of (S * 2^(e - 1))^3 - x, S < 2^56, x an integer < 2^64, exactly (four 64-bit limbs) */
static int cube_mid_cmp(uint64_t S, int e, uint64_t x) {
    const u128_t s2 = (u128_t)S * S; /* < 2^112 */
    const u128_t p0 = (u128_t)(uint64_t)s2 * S, p1 = (u128_t)(uint64_t)(s2 >> 64) * S;
    uint64_t a[4];
    a[0] = (uint64_t)p0;
    const u128_t mid = p1 + (p0 >> 64);
    a[1] = (uint64_t)mid;
    a[2] = (uint64_t)(mid >> 64);
    a[3] = 0;
    /* compare S^3 with x * 2^k, k = -3 (e - 1) */
    const int k = -3 * (e - 1);
    if (k < 0 || k > 190) return k < 0 ? 1 : -1; /* outside the range this is used on */
    uint64_t b[4] = {0, 0, 0, 0};
    const int w = k / 64, r = k % 64;
    b[w] = x << r;
    if (r && w + 1 < 4) b[w + 1] = x >> (64 - r);
    for (int q = 3; q >= 0; q--) {
        if (a[q] != b[q]) return a[q] > b[q] ? 1 : -1;
    }
    return 0;
}

/* a and b adjacent positive doubles: sign of ((a + b) / 2)^3 - x */
static int mid_cube_vs(double a, double b, uint64_t x) {
    int ea, eb;
    const double fa = frexp(a, &ea), fb = frexp(b, &eb);
    uint64_t A = (uint64_t)ldexp(fa, 53), B = (uint64_t)ldexp(fb, 53);
    ea -= 53;
    eb -= 53;
    const int e = ea < eb ? ea : eb;
    A <<= ea - e;
    B <<= eb - e;
    return cube_mid_cmp(A + B, e, x);
}

double gro_cbrt_cr(double x) {
    double y = cbrt(x);
    if (!(x >= 1.0 && x <= 9007199254740992.0) || x != floor(x)) return y;
    const uint64_t X = (uint64_t)x;
    for (int it = 0; it < 4; it++) {
        const double lo = nextafter(y, 0.0), hi = nextafter(y, INFINITY);
        if (mid_cube_vs(lo, y, X) > 0) {
            y = lo;
        } else if (mid_cube_vs(y, hi, X) < 0) {
            y = hi;
        } else {
            break;
        }
    }
    return y;
}

/* forcesim/state.rs:117-132 */
void gro_forcesim_constants(uint32_t n, const float *mass, double scaling_ratio, double gravity,
                            double out3[3]) {
    const double cn = gro_cbrt_cr((double)(n > 0 ? n : 1));
    const double k = 5.0 / (cn > 1.0 ? cn : 1.0);
    const double mean = n == 0 ? 1.0 : (double)(gro_pairwise_f32(mass, n) / (float)n);
    double m2 = mean * mean, m3 = mean * mean * mean;
    if (m2 < 1e-9) m2 = 1e-9;
    if (m3 < 1e-9) m3 = 1e-9;
    const double rep = scaling_ratio * (k * k / m3);
    const double norm = 0.38 * 2.0 * (5.0 * 5.0) / ((k > 1e-9 ? k : 1e-9) * m2);
    out3[0] = k;
    out3[1] = rep;
    out3[2] = gravity * 0.1 * norm;
}

/* simple_graph (layout/force/mod.rs:70-180): self-loops dropped, an unordered pair kept at
 * its first occurrence as (lo, hi), in first-occurrence order */
static uint32_t simple_edges(uint32_t n, const uint32_t *src, const uint32_t *dst, size_t m,
                             uint32_t *lo, uint32_t *hi) {
    uint64_t *key = (uint64_t *)malloc((m ? m : 1) * sizeof(uint64_t));
    size_t cnt = 0;
    for (size_t e = 0; e < m; e++) {
        uint32_t a = src[e], b = dst[e];
        if (a == b || a >= n || b >= n) continue;
        if (a > b) {
            uint32_t t = a;
            a = b;
            b = t;
        }
        key[cnt++] = ((uint64_t)a << 32) | b;
    }
    /* insertion-free dedup: a stable sort of (key, index) by key, first of each run kept */
    size_t *ord = (size_t *)malloc((cnt ? cnt : 1) * sizeof(size_t));
    for (size_t i = 0; i < cnt; i++) ord[i] = i;
    /* merge sort on ord by key (stable) */
    size_t *tmp = (size_t *)malloc((cnt ? cnt : 1) * sizeof(size_t));
    for (size_t w = 1; w < cnt; w *= 2) {
        for (size_t a0 = 0; a0 < cnt; a0 += 2 * w) {
            size_t a1 = a0 + w < cnt ? a0 + w : cnt, a2 = a0 + 2 * w < cnt ? a0 + 2 * w : cnt;
            size_t i = a0, j = a1, o = a0;
            while (i < a1 && j < a2) tmp[o++] = key[ord[j]] < key[ord[i]] ? ord[j++] : ord[i++];
            while (i < a1) tmp[o++] = ord[i++];
            while (j < a2) tmp[o++] = ord[j++];
        }
        memcpy(ord, tmp, cnt * sizeof(size_t));
    }
    unsigned char *keep = (unsigned char *)calloc(cnt ? cnt : 1, 1);
    for (size_t r = 0; r < cnt; r++)
        if (r == 0 || key[ord[r]] != key[ord[r - 1]]) keep[ord[r]] = 1;
    uint32_t ms = 0;
    for (size_t i = 0; i < cnt; i++)
        if (keep[i]) {
            lo[ms] = (uint32_t)(key[i] >> 32);
            hi[ms] = (uint32_t)key[i];
            ms++;
        }
    free(key);
    free(ord);
    free(tmp);
    free(keep);
    return ms;
}

/* pair_force.rs: every target against every source through an f64 coefficient row */
static void repulsion(uint32_t n, const float *pos, const float *mass, double rep, float soften,
                      float *out, double *row) {
    float *s2 = (float *)malloc((n ? n : 1) * sizeof(float));
    for (uint32_t j = 0; j < n; j++) s2[j] = gro_dot3_f32(pos + 3 * j, pos + 3 * j);
    for (uint32_t i = 0; i < n; i++) {
        const float *ri = pos + 3 * i;
        const float own = gro_dot3_f32(ri, ri);
        for (uint32_t j = 0; j < n; j++) {
            float d2 = own + s2[j];
            d2 -= 2.0f * gro_dot3_f32(ri, pos + 3 * j);
            if (d2 < 0.0f) d2 = 0.0f;
            d2 += soften;
            row[j] = rep * (double)mass[i] * (double)mass[j] / (double)d2;
        }
        row[i] = 0.0;
        const double total = gro_pairwise_f64(row, n);
        for (int c = 0; c < 3; c++) {
            double product = 0.0;
            for (uint32_t j = 0; j < n; j++) product += row[j] * (double)pos[3 * j + c];
            out[3 * i + c] = (float)((double)pos[3 * i + c] * total - product);
        }
    }
    free(s2);
}

/* attraction.rs: f32 pulls, added into f64 accumulators in edge order (sources, then
 * subtracted at the destinations), per axis */
static void attraction(uint32_t n, uint32_t m, const uint32_t *lo, const uint32_t *hi,
                       const float *pos, const float *mass, float *force) {
    float *pull = (float *)malloc((size_t)(m ? m : 1) * 3 * sizeof(float));
    for (uint32_t e = 0; e < m; e++) {
        const uint32_t a = lo[e], b = hi[e];
        const float factor = 1.0f / (0.5f * (mass[a] + mass[b]));
        for (int c = 0; c < 3; c++) pull[3 * e + c] = (pos[3 * b + c] - pos[3 * a + c]) * factor;
    }
    double *acc = (double *)malloc((n ? n : 1) * sizeof(double));
    for (int c = 0; c < 3; c++) {
        for (uint32_t i = 0; i < n; i++) acc[i] = 0.0;
        for (uint32_t e = 0; e < m; e++) acc[lo[e]] += (double)pull[3 * e + c];
        for (uint32_t e = 0; e < m; e++) acc[hi[e]] -= (double)pull[3 * e + c];
        for (uint32_t i = 0; i < n; i++) force[3 * i + c] += (float)acc[i];
    }
    free(pull);
    free(acc);
}

/* gravity.rs */
static void gravity_field(uint32_t n, const float *pos, const float *mass, double grav,
                          float *force) {
    float center[3];
    column_means_f32(pos, n, center);
    for (uint32_t i = 0; i < n; i++) {
        float d[3];
        for (int c = 0; c < 3; c++) d[c] = pos[3 * i + c] - center[c];
        const float dist = sqrtf(gro_dot3_f32(d, d));
        const double scale = -(grav * (double)mass[i] / (double)fmaxf(dist, 1e-9f));
        for (int c = 0; c < 3; c++) force[3 * i + c] += (float)(scale * (double)d[c]);
    }
}

static double clampd(double x, double lo, double hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

int gro_forcesim_run(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m_raw,
                     uint32_t iterations, double scale, double scaling_ratio, double gravity,
                     double jitter_tolerance, uint64_t seed, float *out_xyz, double *consts3) {
    uint32_t *lo = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    uint32_t *hi = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    const uint32_t m = simple_edges(n, source, target, m_raw, lo, hi);
    const size_t nn = n ? n : 1;
    float *mass = (float *)malloc(nn * sizeof(float));
    uint32_t *deg = (uint32_t *)calloc(nn, sizeof(uint32_t));
    for (uint32_t e = 0; e < m; e++) {
        deg[lo[e]]++;
        deg[hi[e]]++;
    }
    for (uint32_t i = 0; i < n; i++) mass[i] = (float)(deg[i] + 1);
    double cst[3];
    gro_forcesim_constants(n, mass, scaling_ratio, gravity, cst);
    if (consts3) memcpy(consts3, cst, sizeof(cst));
    const double k = cst[0], rep = cst[1], grav = cst[2];
    const double softening = 0.01 * k;
    const float soften = (float)(softening * softening);
    float *pos = (float *)malloc(nn * 3 * sizeof(float));
    {
        pcg64_t r;
        pcg64_seed(&r, seed);
        for (size_t i = 0; i < (size_t)n * 3; i++) pos[i] = (float)((pcg64_double(&r) - 0.5) * 5.0);
    }
    float *prev = (float *)calloc(nn * 3, sizeof(float));
    float *force = (float *)malloc(nn * 3 * sizeof(float));
    float *sw = (float *)malloc(nn * sizeof(float));
    float *w = (float *)malloc(nn * sizeof(float));
    double *row = (double *)malloc(nn * sizeof(double));
    double speed = 1.0;
    const uint32_t steps = iterations > 0 ? iterations : 1;
    for (uint32_t it = 0; it < steps; it++) {
        if (n < 2) break;
        float center_was[3];
        column_means_f32(pos, n, center_was);
        repulsion(n, pos, mass, rep, soften, force, row);
        attraction(n, m, lo, hi, pos, mass, force);
        gravity_field(n, pos, mass, grav, force);
        /* integrate.rs: swing, speed, move, recentre */
        for (uint32_t i = 0; i < n; i++) {
            float v[3];
            for (int c = 0; c < 3; c++) v[c] = force[3 * i + c] - prev[3 * i + c];
            sw[i] = sqrtf(gro_dot3_f32(v, v));
        }
        for (uint32_t i = 0; i < n; i++) w[i] = mass[i] * sw[i];
        const float total_swing = gro_pairwise_f32(w, n);
        for (uint32_t i = 0; i < n; i++) {
            float v[3];
            for (int c = 0; c < 3; c++) v[c] = 0.5f * (force[3 * i + c] + prev[3 * i + c]);
            w[i] = mass[i] * sqrtf(gro_dot3_f32(v, v));
        }
        const float total_traction = gro_pairwise_f32(w, n);
        if (total_swing > 0.0f) {
            const double tgt = jitter_tolerance * (double)total_traction / (double)total_swing;
            speed = clampd(tgt, speed * 0.5, speed * 1.5);
        }
        speed = clampd(speed, 1e-4, 10.0);
        const float sp = (float)speed;
        for (uint32_t i = 0; i < n; i++) {
            const float factor = sp / (1.0f + sp * sqrtf(sw[i]));
            float d[3];
            for (int c = 0; c < 3; c++) d[c] = force[3 * i + c] * factor;
            const float norm = sqrtf(gro_dot3_f32(d, d));
            if ((double)norm > k) {
                const double shrink = k / (double)norm;
                for (int c = 0; c < 3; c++) d[c] = (float)((double)d[c] * shrink);
            }
            for (int c = 0; c < 3; c++) pos[3 * i + c] += d[c];
        }
        memcpy(prev, force, (size_t)n * 3 * sizeof(float));
        float center_now[3];
        column_means_f32(pos, n, center_now);
        for (uint32_t i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) pos[3 * i + c] -= center_now[c] - center_was[c];
    }
    /* rescale.rs, then the narrowing and the NonFinite refusal (forcesim.rs:158-187) */
    double *arr = (double *)malloc(nn * 3 * sizeof(double));
    for (size_t i = 0; i < (size_t)n * 3; i++) arr[i] = (double)pos[i];
    if (n > 0) {
        double means[3];
        column_means_f64(arr, n, means);
        for (size_t i = 0; i < (size_t)n * 3; i++) arr[i] -= means[i % 3];
        double span = 0.0;
        for (size_t i = 0; i < (size_t)n * 3; i++) {
            const double a = fabs(arr[i]);
            span = span >= a || isnan(a) ? span : a;
        }
        if (span > 1e-9) {
            const double fit = scale / span;
            for (size_t i = 0; i < (size_t)n * 3; i++) arr[i] *= fit;
        }
    }
    int status = 0;
    for (size_t i = 0; i < (size_t)n * 3; i++)
        if (!isfinite(arr[i])) status = 2;
    for (size_t i = 0; i < (size_t)n * 3; i++) out_xyz[i] = status ? 0.0f : (float)arr[i];
    free(lo);
    free(hi);
    free(mass);
    free(deg);
    free(pos);
    free(prev);
    free(force);
    free(sw);
    free(w);
    free(row);
    free(arr);
    return status;
}
