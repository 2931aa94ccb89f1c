// Particle-mesh FA2 repulsion on gfx950: F_i = k * m_i * E(p_i), E = mass density (*) G,
// G(r) = r / |r|^2, solved on a side x side periodic mesh in FP32.
//
// Reference machinery (paths relative to crates/graph-core/src/layout/force/particle_mesh/):
//   frame    frame.rs:50-89 (rung ladder h = 2^(step/4), cells = floor(span/h) + 3, origin
//            snapped to h); the FA2 law has infinite reach, so reach = cells - 1 and
//            cells + reach <= side keeps occupied cells free of periodic images
//   deposit  deposit.rs (CIC weights); here an exact fixed-point sum, so the density does not
//            depend on the order the atomics land in
//   solve    mesh.rs:151-185 + kernel.rs: packed kernel Gx + i*Gy, one forward transform of
//            the density, multiply, one inverse; Ex = re, Ey = im; 1/side^2 folded into G
//   read     mesh.rs:276-288 (CIC gather)
// The FFT is a hand-written radix-4 (+ one radix-2) Stockham transform per line in LDS; 2D
// transforms are row passes with tiled LDS transposes between them. Only the rows that hold
// density (and the rows that are read back) are transformed along x.
#include <math.h>

#include "fa2_kernels.h"
#include "sort_scan.h"

namespace gr {

struct PmFrame {
    double h, ox, oy;
    int32_t step, cells, reach, valid;
    int32_t kstep, kreach, kvalid, rebuild;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// 2^(q/4) for q = 0..3, correctly rounded: h = 2^(step/4) built exactly from these
__device__ __forceinline__ double pm_h(int step) {
    const double q[4] = {1.0, 1.189207115002721, 1.4142135623730951, 1.681792830507429};
    return ldexp(q[step & 3], step >> 2);
}

__global__ void k_pm_frame(const FaDev d) {
    const GrState *st = d.st;
    const GrGraph *g = d.gr;  // the particle mesh runs on single-graph sessions
    PmFrame *f = (PmFrame *)d.pm_fr;
    f->valid = 0;
    f->rebuild = 0;
    if (st->halt) return;
    const int side = (int)d.pm_side;
    const double span = fmax(g->box_hi[0] - g->box_lo[0], g->box_hi[1] - g->box_lo[1]);
    if (d.n < 2 || g->n_finite == 0 || !isfinite(span)) return;
    if (d.pm_law == 1 && !(d.pm_dmax2 > 0.0)) return;  // mesh.rs:166: no law, no field
    double s0 = floor(4.0 * log2(span / (double)side));
    int step = s0 < -32.0 ? -32 : (int)s0;
    for (int guard = 0; guard < 4096; guard++, step++) {
        const double h = pm_h(step);
        const double cells = floor(span / h) + 3.0;
        if (cells > (double)side) continue;
        const int c = (int)cells;
        int reach = c - 1;
        if (d.pm_law == 1) {  // frame.rs:77-78: the law's own reach
            const double r = ceil(sqrt(d.pm_dmax2) / h);
            reach = r < (double)(c - 1) ? (int)r : c - 1;
        }
        if (c + reach > side) continue;
        f->step = step;
        f->h = h;
        f->cells = c;
        f->reach = reach;
        f->ox = floor(g->box_lo[0] / h) * h;
        f->oy = floor(g->box_lo[1] / h) * h;
        f->valid = 1;
        f->rebuild = !(f->kvalid && f->kstep == step && f->kreach == reach);
        // the spectrum launches that follow read only `rebuild`: the rung they build it for
        // is recorded here
        f->kstep = step;
        f->kreach = reach;
        f->kvalid = 1;
        return;
    }
}

__device__ __forceinline__ void pm_stencil(const PmFrame *f, double px, double py, int &cx,
                                           int &cy, double &fx, double &fy) {
    const double ux = (px - f->ox) / f->h, uy = (py - f->oy) / f->h;
    long ix = ux > 0.0 ? (long)ux : 0, iy = uy > 0.0 ? (long)uy : 0;
    if (ix > f->cells - 2) ix = f->cells - 2;
    if (iy > f->cells - 2) iy = f->cells - 2;
    const double gx = ux - (double)ix, gy = uy - (double)iy;
    fx = gx < 0.0 ? 0.0 : (gx > 1.0 ? 1.0 : gx);
    fy = gy < 0.0 ? 0.0 : (gy > 1.0 ? 1.0 : gy);
    cx = (int)ix;
    cy = (int)iy;
}

// ---------------------------------------------------------------- kernel spectrum

// G sampled at every cell offset within reach, wrapped modulo side, scaled by 1/side^2.
__global__ __launch_bounds__(GR_BLOCK) void k_pm_sample(const FaDev d) {
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid || !f->rebuild) return;
    const uint32_t P = d.pm_side;
    const double scale = 1.0 / ((double)P * (double)P);
    const int reach = f->reach;
    for (uint64_t q = (uint64_t)blockIdx.x * GR_BLOCK + threadIdx.x; q < (uint64_t)P * P;
         q += (uint64_t)gridDim.x * GR_BLOCK) {
        const int r = (int)(q / P), c = (int)(q % P);
        const int dy = r <= reach ? r : r - (int)P;
        const int dx = c <= reach ? c : c - (int)P;
        float2 g = make_float2(0.f, 0.f);
        if (dy >= -reach && dy <= reach && dx >= -reach && dx <= reach) {
            const double rx = dx * f->h, ry = dy * f->h;
            double l = rx * rx + ry * ry;
            if (d.pm_law == 1) {  // kernel.rs:80-92
                if (l != 0.0 && l < d.pm_dmax2) {
                    if (l < d.pm_dmin2) l = sqrt(d.pm_dmin2 * l);
                    g = make_float2((float)(-rx / l * scale), (float)(-ry / l * scale));
                }
            } else if (l != 0.0) {
                g = make_float2((float)(rx / l * scale), (float)(ry / l * scale));
            }
        }
        d.pm_b[q] = g;
    }
}

// ---------------------------------------------------------------- FFT

// exp(-2 pi i t / P) for t < P from the half table tw (t < P/2): w(t) = -w(t - P/2);
// conjugated for the inverse (dir +1)
__device__ __forceinline__ float2 tw_at(const float2 *__restrict__ tw, uint32_t t, uint32_t P,
                                        int dir) {
    const uint32_t h = P >> 1;
    float2 w = t < h ? tw[t] : make_float2(-tw[t - h].x, -tw[t - h].y);
    if (dir > 0) w.y = -w.y;
    return w;
}

// Stockham (self-sorting) FFT over one line of P points held in LDS (x, y: P float2 each):
// radix-4 stages while four sub-transforms fit, then one radix-2 stage when log2 P is odd.
// Stage with sub-transform size ns and radix R: butterfly j takes in[j + r P/R] times the
// twiddle w(r k P/(R ns)), k = j mod ns, and writes out[(j - k) R + k + r ns]. Returns the
// buffer holding the result; dir +1 is the inverse (conjugate twiddles, unnormalised).
__device__ float2 *stockham(float2 *x, float2 *y, uint32_t P, const float2 *__restrict__ tw,
                            int dir) {
    float2 *in = x, *out = y;
    uint32_t ns = 1;
    const uint32_t q = P >> 2;
    for (; ns * 4 <= P; ns <<= 2) {
        const uint32_t tstride = P / (4 * ns);
        for (uint32_t j = threadIdx.x; j < q; j += blockDim.x) {
            const uint32_t k = j & (ns - 1);
            float2 v0 = in[j], v1 = in[j + q], v2 = in[j + 2 * q], v3 = in[j + 3 * q];
            if (ns > 1) {  // block-uniform: the first stage's twiddles are all 1
                v1 = cmul(v1, tw_at(tw, k * tstride, P, dir));
                v2 = cmul(v2, tw_at(tw, 2 * k * tstride, P, dir));
                v3 = cmul(v3, tw_at(tw, 3 * k * tstride, P, dir));
            }
            // 4-point DFT: y1 = a1 - i d, y3 = a1 + i d (forward; signs swap for the inverse)
            const float2 a0 = make_float2(v0.x + v2.x, v0.y + v2.y);
            const float2 a1 = make_float2(v0.x - v2.x, v0.y - v2.y);
            const float2 a2 = make_float2(v1.x + v3.x, v1.y + v3.y);
            const float2 dd = make_float2(v1.x - v3.x, v1.y - v3.y);
            const float2 a3 = dir < 0 ? make_float2(dd.y, -dd.x) : make_float2(-dd.y, dd.x);
            const uint32_t idx = ((j - k) << 2) + k;
            out[idx] = make_float2(a0.x + a2.x, a0.y + a2.y);
            out[idx + ns] = make_float2(a1.x + a3.x, a1.y + a3.y);
            out[idx + 2 * ns] = make_float2(a0.x - a2.x, a0.y - a2.y);
            out[idx + 3 * ns] = make_float2(a1.x - a3.x, a1.y - a3.y);
        }
        __syncthreads();
        float2 *t = in;
        in = out;
        out = t;
    }
    if (ns < P) {  // the radix-2 stage, ns = P / 2
        const uint32_t half = P >> 1;
        for (uint32_t j = threadIdx.x; j < half; j += blockDim.x) {
            const uint32_t k = j & (ns - 1);
            const float2 w = tw_at(tw, k * (P / (2 * ns)), P, dir);
            const float2 v0 = in[j];
            const float2 v1 = cmul(in[j + half], w);
            const uint32_t idx = ((j - k) << 1) + k;
            out[idx] = make_float2(v0.x + v1.x, v0.y + v1.y);
            out[idx + ns] = make_float2(v0.x - v1.x, v0.y - v1.y);
        }
        __syncthreads();
        float2 *t = in;
        in = out;
        out = t;
    }
    return in;
}

// One workgroup per line. mode 0: forward; 1: inverse; 2: forward, times the kernel
// spectrum line, inverse. Lines >= limit are skipped (limit -1 means the frame's cells).
__global__ __launch_bounds__(GR_BLOCK) void k_pm_lines(const FaDev d, float2 *buf, int mode,
                                                       int limit) {
    extern __shared__ float2 lds_f2[];
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid) return;
    if (limit == -2 && !f->rebuild) return;
    const uint32_t P = d.pm_side;
    const uint32_t lim = limit < 0 ? (limit == -1 ? (uint32_t)f->cells : P) : (uint32_t)limit;
    const uint32_t line = blockIdx.x;
    if (line >= lim) return;
    float2 *x = lds_f2, *y = lds_f2 + P;
    float2 *row = buf + (size_t)line * P;
    for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) x[j] = row[j];
    __syncthreads();
    float2 *r = stockham(x, y, P, d.pm_tw, mode == 1 ? 1 : -1);
    if (mode == 2) {
        const float2 *kl = d.pm_k + (size_t)line * P;
        for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) r[j] = cmul(r[j], kl[j]);
        __syncthreads();
        float2 *other = r == x ? y : x;
        r = stockham(r, other, P, d.pm_tw, 1);
    }
    for (uint32_t j = threadIdx.x; j < P; j += blockDim.x) row[j] = r[j];
}

// dst[c][r] = src[r][c] through 32x32 LDS tiles. rows_src: rows of src that hold data
// (others read as zero); rows_dst: rows of dst to write (-1: the frame's cells; 0: all).
__global__ __launch_bounds__(GR_BLOCK) void k_pm_transpose(const FaDev d, const float2 *src,
                                                           float2 *dst, int src_cells,
                                                           int dst_cells, int only_rebuild) {
    __shared__ float2 tile[32][33];
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid) return;
    if (only_rebuild && !f->rebuild) return;
    const uint32_t P = d.pm_side;
    const uint32_t rs = src_cells ? (uint32_t)f->cells : P;  // src rows with data
    const uint32_t rd = dst_cells ? (uint32_t)f->cells : P;  // dst rows needed
    const uint32_t bx = blockIdx.x * 32, by = blockIdx.y * 32;  // src col tile, src row tile
    if (bx >= rd) return;  // dst row = src col
    const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (uint32_t k = ty; k < 32; k += 8) {
        const uint32_t r = by + k, c = bx + tx;
        tile[k][tx] = r < rs ? src[(size_t)r * P + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (uint32_t k = ty; k < 32; k += 8) {
        const uint32_t r = bx + k, c = by + tx;  // dst row r = src col, dst col = src row
        if (r < rd) dst[(size_t)r * P + c] = tile[tx][k];
    }
}

// ---------------------------------------------------------------- deposit / read

// The deposit works tile by tile. Particles are sorted by the PM_TILE x PM_TILE tile of
// cells their stencil's lower corner (cx, cy) falls in; the workgroup that owns a tile sums
// its particles' stencil weights in LDS over the tile plus one column and one row beyond it,
// stores the tile's own cells with plain stores and leaves the 2 * PM_TILE + 1 cells beyond
// it in the tile's halo record; k_pm_to_grid adds the halo records of the tiles left of and
// below a cell when it reads the cell. No global atomics and no clearing pass: every cell of
// the frame's cells x cells region is stored. The weights are the same fixed-point integers
// a per-particle deposit would add, so the sums are exact and independent of order.
constexpr uint32_t PM_TILE = 16;
constexpr uint32_t PM_HALO = 2 * PM_TILE + 1;  // column beyond, row beyond, the corner

__device__ __forceinline__ uint32_t pm_tiles_per_side(const PmFrame *f) {
    return ((uint32_t)f->cells + PM_TILE - 1) / PM_TILE;
}

// the fixed-point CIC weights of particle i (cells (cx, cy), (cx+1, cy), (cx, cy+1),
// (cx+1, cy+1)); false for a non-finite position
template <typename T>
__device__ __forceinline__ bool pm_weights(const FaDev &d, const PmFrame *f, uint32_t i,
                                           int &cx, int &cy, unsigned long long w[4]) {
    const T *pi = (const T *)d.pi;
    const double px = (double)pi[2 * (size_t)i], py = (double)pi[2 * (size_t)i + 1];
    if (!isfinite(px) || !isfinite(py)) return false;
    double fx, fy;
    pm_stencil(f, px, py, cx, cy, fx, fy);
    const double m = (d.pm_law == 1 ? 1.0 : (double)((const T *)d.mass)[i]) * d.pm_fix;
    w[0] = (unsigned long long)llrint(m * ((1.0 - fx) * (1.0 - fy)));
    w[1] = (unsigned long long)llrint(m * (fx * (1.0 - fy)));
    w[2] = (unsigned long long)llrint(m * ((1.0 - fx) * fy));
    w[3] = (unsigned long long)llrint(m * (fx * fy));
    return true;
}

// sort keys: the tile of each particle (tiles row-major, tiles_per_side wide); non-finite
// particles and an invalid frame get `none`, past every tile
template <typename T>
__global__ __launch_bounds__(GR_BLOCK) void k_pm_keys(const FaDev d, uint64_t *keys,
                                                      uint32_t *vals, uint32_t none) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= d.n) return;
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    uint32_t key = none;
    if (f->valid) {
        const T *pi = (const T *)d.pi;
        const double px = (double)pi[2 * (size_t)i], py = (double)pi[2 * (size_t)i + 1];
        if (isfinite(px) && isfinite(py)) {
            int cx, cy;
            double fx, fy;
            pm_stencil(f, px, py, cx, cy, fx, fy);
            key = ((uint32_t)cy / PM_TILE) * pm_tiles_per_side(f) + (uint32_t)cx / PM_TILE;
        }
    }
    keys[i] = key;
    vals[i] = i;
}

// pm_tiles[t] = the first sorted particle of tile t (t = 0 .. tiles, the last one the end of
// the finite particles): position j writes the entries of the tiles in (key[j-1], key[j]],
// so every entry is written exactly once
__global__ __launch_bounds__(GR_BLOCK) void k_pm_tile_starts(const FaDev d,
                                                             const uint64_t *keys) {
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid) return;
    const uint32_t j = blockIdx.x * GR_BLOCK + threadIdx.x;
    const uint32_t n = d.n;
    if (j > n) return;
    const uint32_t nt = pm_tiles_per_side(f);
    const int64_t tiles = (int64_t)nt * nt;
    const int64_t prev = j > 0 ? (int64_t)keys[j - 1] : -1;
    const int64_t cur = j < n ? (int64_t)keys[j] : tiles;
    const int64_t hi = cur < tiles ? cur : tiles;
    for (int64_t t = prev + 1; t <= hi; t++) d.pm_tiles[t] = j;
}

// one workgroup per tile: its particles' weights summed in LDS over (PM_TILE + 1)^2 cells,
// the tile's cells stored, the rest left in its halo record (column PM_TILE rows 0..15, row
// PM_TILE columns 0..15, then the corner)
template <typename T>
__global__ __launch_bounds__(GR_BLOCK) void k_pm_deposit(const FaDev d, const uint32_t *order) {
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid) return;
    const uint32_t nt = pm_tiles_per_side(f);
    const uint32_t t = blockIdx.x;
    if (t >= nt * nt) return;
    constexpr uint32_t W = PM_TILE + 1;
    const uint32_t tx = t % nt, ty = t / nt;
    const int x0 = (int)(tx * PM_TILE), y0 = (int)(ty * PM_TILE);
    __shared__ unsigned long long acc[W * W];
    for (uint32_t q = threadIdx.x; q < W * W; q += GR_BLOCK) acc[q] = 0ull;
    __syncthreads();
    const uint32_t j0 = d.pm_tiles[t], j1 = d.pm_tiles[t + 1];
    for (uint32_t j = j0 + threadIdx.x; j < j1; j += GR_BLOCK) {
        int cx, cy;
        unsigned long long w[4];
        if (!pm_weights<T>(d, f, order[j], cx, cy, w)) continue;
        // the particle's tile is this one: cx - x0 and cy - y0 are in [0, PM_TILE)
        const uint32_t at = (uint32_t)(cy - y0) * W + (uint32_t)(cx - x0);
        atomicAdd(&acc[at], w[0]);  // LDS
        atomicAdd(&acc[at + 1], w[1]);
        atomicAdd(&acc[at + W], w[2]);
        atomicAdd(&acc[at + W + 1], w[3]);
    }
    __syncthreads();
    const int cells = f->cells;
    const uint32_t P = d.pm_side;
    for (uint32_t q = threadIdx.x; q < PM_TILE * PM_TILE; q += GR_BLOCK) {
        const uint32_t lx = q % PM_TILE, ly = q / PM_TILE;
        const int X = x0 + (int)lx, Y = y0 + (int)ly;
        if (X < cells && Y < cells) d.pm_acc[(size_t)Y * P + X] = acc[ly * W + lx];
    }
    unsigned long long *halo = d.pm_halo + (size_t)t * PM_HALO;
    for (uint32_t q = threadIdx.x; q < PM_HALO; q += GR_BLOCK)
        halo[q] = q < PM_TILE       ? acc[q * W + PM_TILE]
                  : q < 2 * PM_TILE ? acc[PM_TILE * W + (q - PM_TILE)]
                                    : acc[PM_TILE * W + PM_TILE];
}

__global__ __launch_bounds__(GR_BLOCK) void k_pm_to_grid(const FaDev d) {
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    if (!f->valid) return;
    const uint32_t P = d.pm_side, c = (uint32_t)f->cells;
    const double inv = 1.0 / d.pm_fix;
    const uint32_t nt = pm_tiles_per_side(f);
    for (uint64_t q = (uint64_t)blockIdx.x * GR_BLOCK + threadIdx.x; q < (uint64_t)c * P;
         q += (uint64_t)gridDim.x * GR_BLOCK) {
        const uint32_t col = (uint32_t)(q % P), row = (uint32_t)(q / P);
        float v = 0.f;
        if (col < c) {
            // the tile's own sum, plus what the tiles left of and below it left over
            unsigned long long a = d.pm_acc[q];
            const uint32_t tx = col / PM_TILE, ty = row / PM_TILE;
            const uint32_t lx = col % PM_TILE, ly = row % PM_TILE;
            if (lx == 0 && tx > 0) a += d.pm_halo[(size_t)(ty * nt + tx - 1) * PM_HALO + ly];
            if (ly == 0 && ty > 0)
                a += d.pm_halo[(size_t)((ty - 1) * nt + tx) * PM_HALO + PM_TILE + lx];
            if (lx == 0 && ly == 0 && tx > 0 && ty > 0)
                a += d.pm_halo[(size_t)((ty - 1) * nt + tx - 1) * PM_HALO + 2 * PM_TILE];
            v = (float)((double)a * inv);
        }
        d.pm_a[q] = make_float2(v, 0.f);
    }
}

template <typename T>
__global__ __launch_bounds__(GR_BLOCK) void k_pm_read(const FaDev d) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= d.n) return;
    const PmFrame *f = (const PmFrame *)d.pm_fr;
    T *rx = (T *)d.rep[0], *ry = (T *)d.rep[1];
    if (!f->valid) {
        rx[i] = (T)0;
        ry[i] = (T)0;
        return;
    }
    const T *pi = (const T *)d.pi;
    const double px = (double)pi[2 * (size_t)i], py = (double)pi[2 * (size_t)i + 1];
    if (!isfinite(px) || !isfinite(py)) {
        rx[i] = (T)NAN;
        ry[i] = (T)NAN;
        return;
    }
    int cx, cy;
    double fx, fy;
    pm_stencil(f, px, py, cx, cy, fx, fy);
    const uint32_t P = d.pm_side;
    const size_t at = (size_t)cy * P + cx;
    const double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy};
    const size_t cell[4] = {at, at + 1, at + P, at + P + 1};
    double ex = 0.0, ey = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const float2 e = d.pm_a[cell[q]];
        ex += (double)e.x * w[q];
        ey += (double)e.y * w[q];
    }
    const double km =
        d.pm_law == 1 ? d.pm_strength : d.scaling_ratio * (double)((const T *)d.mass)[i];
    rx[i] = (T)(km * ex);
    ry[i] = (T)(km * ey);
}

// ---------------------------------------------------------------- driver

void pm_repulsion(FaDev &d) {
    hipStream_t s = d.stream;
    const uint32_t P = d.pm_side;
    const uint32_t tiles = P / 32;
    const size_t lds = 2 * (size_t)P * sizeof(float2);
    const uint32_t fill = 2048;
    hipLaunchKernelGGL(k_pm_frame, dim3(1), dim3(1), 0, s, d);
    // kernel spectrum, only when the rung or reach changed: sample, FFT rows, transpose,
    // FFT rows again, so it sits in the transposed layout the column pass multiplies in
    hipLaunchKernelGGL(k_pm_sample, dim3(fill), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL(k_pm_lines, dim3(P), dim3(GR_BLOCK), lds, s, d, d.pm_b, 0, -2);
    hipLaunchKernelGGL(k_pm_transpose, dim3(tiles, tiles), dim3(GR_BLOCK), 0, s, d,
                       (const float2 *)d.pm_b, d.pm_k, 0, 0, 1);
    hipLaunchKernelGGL(k_pm_lines, dim3(P), dim3(GR_BLOCK), lds, s, d, d.pm_k, 0, -2);
    // density, forward along x (rows with data), transpose, y-lines forward * K inverse,
    // transpose back (rows read), inverse along x, CIC read
    {
        // particles sorted by deposit tile; the keys are written where an odd number of
        // sort passes starts, so the sorted pairs end in (keys, vals)
        const uint32_t side_tiles = P / PM_TILE;
        const uint32_t none = side_tiles * side_tiles;  // past every tile of any frame
        int bits = 1;
        while ((1u << bits) <= none) bits++;
        const bool odd = radix_digits(bits) & 1;
        uint64_t *k0 = odd ? d.keys_alt : d.keys, *k1 = odd ? d.keys : d.keys_alt;
        uint32_t *v0 = odd ? d.vals_alt : d.vals, *v1 = odd ? d.vals : d.vals_alt;
        const uint32_t nb = gr_blocks(d.n, GR_BLOCK);
        if (d.precision == PREC_FP64)
            hipLaunchKernelGGL(k_pm_keys<double>, dim3(nb), dim3(GR_BLOCK), 0, s, d, k0, v0,
                               none);
        else
            hipLaunchKernelGGL(k_pm_keys<float>, dim3(nb), dim3(GR_BLOCK), 0, s, d, k0, v0,
                               none);
        radix_sort_pairs(k0, v0, k1, v1, d.n, bits, d.sort_tmp, s, (int)d.sort_items);
        hipLaunchKernelGGL(k_pm_tile_starts, dim3(gr_blocks((uint64_t)d.n + 1, GR_BLOCK)),
                           dim3(GR_BLOCK), 0, s, d, (const uint64_t *)d.keys);
        if (d.precision == PREC_FP64)
            hipLaunchKernelGGL(k_pm_deposit<double>, dim3(none), dim3(GR_BLOCK), 0, s, d,
                               (const uint32_t *)d.vals);
        else
            hipLaunchKernelGGL(k_pm_deposit<float>, dim3(none), dim3(GR_BLOCK), 0, s, d,
                               (const uint32_t *)d.vals);
    }
    hipLaunchKernelGGL(k_pm_to_grid, dim3(fill), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL(k_pm_lines, dim3(P), dim3(GR_BLOCK), lds, s, d, d.pm_a, 0, -1);
    hipLaunchKernelGGL(k_pm_transpose, dim3(tiles, tiles), dim3(GR_BLOCK), 0, s, d,
                       (const float2 *)d.pm_a, d.pm_b, 1, 0, 0);
    hipLaunchKernelGGL(k_pm_lines, dim3(P), dim3(GR_BLOCK), lds, s, d, d.pm_b, 2, (int)P);
    hipLaunchKernelGGL(k_pm_transpose, dim3(tiles, tiles), dim3(GR_BLOCK), 0, s, d,
                       (const float2 *)d.pm_b, d.pm_a, 0, 1, 0);
    hipLaunchKernelGGL(k_pm_lines, dim3(P), dim3(GR_BLOCK), lds, s, d, d.pm_a, 1, -1);
    if (d.precision == PREC_FP64)
        hipLaunchKernelGGL(k_pm_read<double>, dim3(gr_blocks(d.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                           s, d);
    else
        hipLaunchKernelGGL(k_pm_read<float>, dim3(gr_blocks(d.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                           s, d);
}

size_t pm_frame_bytes() { return sizeof(PmFrame); }

void pm_frame_read(const void *frame, double *out7) {
    const PmFrame *f = (const PmFrame *)frame;
    out7[0] = f->step;
    out7[1] = f->h;
    out7[2] = f->ox;
    out7[3] = f->oy;
    out7[4] = f->cells;
    out7[5] = f->reach;
    out7[6] = f->valid;
}

void pm_twiddles(uint32_t P, float2 *host_out) {
    for (uint32_t t = 0; t < P / 2; t++) {
        const double a = -2.0 * M_PI * (double)t / (double)P;
        host_out[t] = make_float2((float)cos(a), (float)sin(a));
    }
}

}  // namespace gr
