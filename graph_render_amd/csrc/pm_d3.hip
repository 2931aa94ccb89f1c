// layout.force.particle_mesh's two mesh passes on gfx950, FP64 in the reference's operations
// and orders (paths relative to crates/graph-core/src/layout/force/particle_mesh/):
//   charge   frame.rs place over the positions' finite bounds; the CIC deposit summed per
//            cell in ascending grid slot (deposit.rs Rows/Deposit: the slots are bucketed by
//            stencil cell with a stable sort, each cell merges its four stencil buckets);
//            fft.rs's radix-2 line (bit-reversed load, (u + v, u - v), then t = v * w stages
//            on libm's twiddles) one workgroup per line in LDS, as fft/pass.rs runs it: rows
//            over the live rows, then columns, the inverse with the kernel spectrum as gain
//            (kernel.rs, sampled and transformed when the rung or reach changes); the CIC
//            read times charge * alpha.
//   collide  collide.rs + collide/hash.rs + collide/gather.rs: the projected positions
//            hashed into next_pow2(2n) buckets, a stable sort into the grid order (which the
//            next tick's deposit reads), and each slot's pushes from its 3x3 cells' buckets in
//            the reference's read order, one thread per slot.
#include <math.h>

#include "d3.h"
#include "fa2_common.h"
#include "sort_scan.h"

namespace gr {

// ---------------------------------------------------------------- bounds and frame

// frame.rs Blocks: each block's finite extent (x lo, y lo, x hi, y hi), EMPTY for none
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_bounds(const double *x, const double *y,
                                                          uint32_t n, double *part) {
    __shared__ double red[4][GR_BLOCK];
    double lo0 = INFINITY, lo1 = INFINITY, hi0 = -INFINITY, hi1 = -INFINITY;
    for (uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x; i < n; i += gridDim.x * GR_BLOCK) {
        const double px = x[i], py = y[i];
        if (!(isfinite(px) && isfinite(py))) continue;
        lo0 = fmin(lo0, px);
        lo1 = fmin(lo1, py);
        hi0 = fmax(hi0, px);
        hi1 = fmax(hi1, py);
    }
    red[0][threadIdx.x] = lo0;
    red[1][threadIdx.x] = lo1;
    red[2][threadIdx.x] = hi0;
    red[3][threadIdx.x] = hi1;
    __syncthreads();
    for (uint32_t s = GR_BLOCK / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] = fmin(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = fmin(red[1][threadIdx.x], red[1][threadIdx.x + s]);
            red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + s]);
            red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ void fold_bounds(const double *part, uint32_t nparts, double *b) {
    b[0] = b[1] = INFINITY;
    b[2] = b[3] = -INFINITY;
    for (uint32_t q = 0; q < nparts; q++) {
        b[0] = fmin(b[0], part[4 * q]);
        b[1] = fmin(b[1], part[4 * q + 1]);
        b[2] = fmax(b[2], part[4 * q + 2]);
        b[3] = fmax(b[3], part[4 * q + 3]);
    }
}

// 2^(q/4) for q = 0..3, correctly rounded: h = 2^(step/4) built exactly from these
__device__ __forceinline__ double pmd3_h(int step) {
    const double q[4] = {1.0, 1.189207115002721, 1.4142135623730951, 1.681792830507429};
    return ldexp(q[step & 3], step >> 2);
}

// mesh.rs solve's frame (frame.rs place: the smallest rung from floor(4 log2(span / side))
// whose cells and reach ceil(dmax / h) fit the side; origin snapped to h); valid 0 when the
// mesh places none (no finite node, one node, no reach)
__global__ void k_pmd3_frame(const double *part, uint32_t nparts, PmD3Frame *f, uint32_t side,
                             uint32_t n, double dmax2) {
    double b[4];
    fold_bounds(part, nparts, b);
    f->valid = 0;
    if (!(b[0] <= b[2]) || n <= 1 || !(dmax2 > 0.0)) return;
    const double span = fmax(b[2] - b[0], b[3] - b[1]);
    if (!isfinite(span)) return;
    const double dmax = sqrt(dmax2);
    const double s0 = floor(4.0 * log2(span / (double)side));
    int step = s0 < -32.0 ? -32 : (int)s0;
    for (int guard = 0; guard < 4096; guard++, step++) {
        const double h = pmd3_h(step);
        const double cells = floor(span / h) + 3.0;
        if (cells > (double)side) continue;
        const int c = (int)cells;
        const double r = ceil(dmax / h);
        const int reach = r < (double)(c - 1) ? (int)r : c - 1;
        if (c + reach > (int)side) continue;
        f->step = step;
        f->h = h;
        f->cells = c;
        f->reach = reach;
        f->ox = floor(b[0] / h) * h;
        f->oy = floor(b[1] / h) * h;
        f->valid = 1;
        return;
    }
}

// collide.rs Grid::build: the hash origin at the finite bounds' low corner, (0, 0) for none
__global__ void k_pmd3_origin(const double *part, uint32_t nparts, double *origin) {
    double b[4];
    fold_bounds(part, nparts, b);
    const bool found = b[0] <= b[2];
    origin[0] = found ? b[0] : 0.0;
    origin[1] = found ? b[1] : 0.0;
}

// ---------------------------------------------------------------- deposit

__device__ __forceinline__ uint32_t pmd3_cell(double u, int cells) {
    // frame.rs cell: u as usize (NaN and negatives to 0), at most cells - 2
    const double top = (double)(cells - 2);
    return u > 0.0 ? (uint32_t)(u < top ? u : top) : 0u;
}

// deposit.rs Stencils + frame.rs split for every slot: the stencil cell as the sort key
// (side * side for a non-finite node, past every cell) and the fractions
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_stencil(const D3SimDev s, const PmD3Dev m) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= s.n) return;
    const PmD3Frame *f = m.fr;
    const uint32_t i = m.order[k];
    const double px = s.x[i], py = s.y[i];
    uint32_t key = m.side * m.side;
    double fx = 0.0, fy = 0.0;
    if (isfinite(px) && isfinite(py)) {
        const double ux = (px - f->ox) / f->h, uy = (py - f->oy) / f->h;
        const uint32_t cx = pmd3_cell(ux, f->cells), cy = pmd3_cell(uy, f->cells);
        fx = ux - (double)cx;
        fy = uy - (double)cy;
        fx = fx < 0.0 ? 0.0 : (fx > 1.0 ? 1.0 : fx);
        fy = fy < 0.0 ? 0.0 : (fy > 1.0 ? 1.0 : fy);
        key = cy * m.side + cx;
    }
    m.cell[k] = key;
    m.keys[k] = key;
    m.vals[k] = k;
    m.fxy[2 * (size_t)k] = fx;
    m.fxy[2 * (size_t)k + 1] = fy;
}

// start[t] = the first sorted entry with key >= t, for t in [0, nkeys], one thread per key:
// two lanes bracket the workgroup's keys [t0, t0 + GR_BLOCK) among the entries, then each
// thread searches its key within that bracket. Empty stretches of keys (mesh cells outside
// the layout, unused hash buckets) cost no more than full ones.
__device__ __forceinline__ uint32_t lower_bound_u64(const uint64_t *keys, uint32_t lo,
                                                    uint32_t hi, uint64_t want) {
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (keys[mid] < want)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_starts(const uint64_t *keys, uint32_t n,
                                                          uint32_t nkeys, uint32_t *start) {
    __shared__ uint32_t bracket[2];
    const uint64_t t0 = (uint64_t)blockIdx.x * GR_BLOCK;
    if (t0 > nkeys) return;  // block-uniform
    if (threadIdx.x < 2)
        bracket[threadIdx.x] = lower_bound_u64(keys, 0, n, t0 + (threadIdx.x ? GR_BLOCK : 0));
    __syncthreads();
    const uint64_t t = t0 + threadIdx.x;
    if (t > nkeys) return;
    start[t] = lower_bound_u64(keys, bracket[0], bracket[1], t);
}

// deposit.rs Deposit: cell t sums, from zero, the weight each node whose stencil covers it
// gives it, in ascending slot (the merge of its four stencil buckets)
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_deposit(const PmD3Dev m, const uint32_t *slots,
                                                           uint32_t cells) {
    const uint32_t side = m.side;
    const uint32_t t = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (t >= cells * side) return;
    const uint32_t row = t / side, col = t - row * side;
    // buckets: the stencil at t (weight (1-fx)(1-fy)), at t - 1 (fx (1-fy)), at t - side
    // ((1-fx) fy), at t - side - 1 (fx fy)
    uint32_t p[4], e[4];
    const uint32_t src[4] = {t, t - 1u, t - side, t - side - 1u};
    const bool ok[4] = {true, col >= 1, row >= 1, row >= 1 && col >= 1};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        p[q] = ok[q] ? m.cstart[src[q]] : 0u;
        e[q] = ok[q] ? m.cstart[src[q] + 1] : 0u;
    }
    double sum = 0.0;
    for (;;) {
        int best = -1;
        uint32_t bs = 0xFFFFFFFFu;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (p[q] < e[q]) {
                const uint32_t sl = slots[p[q]];
                if (sl < bs) {
                    bs = sl;
                    best = q;
                }
            }
        }
        if (best < 0) break;
        p[best]++;
        const double fx = m.fxy[2 * (size_t)bs], fy = m.fxy[2 * (size_t)bs + 1];
        const double w = best == 0 ? (1.0 - fx) * (1.0 - fy)
                       : best == 1 ? fx * (1.0 - fy)
                       : best == 2 ? (1.0 - fx) * fy
                                   : fx * fy;
        sum += w;
    }
    m.a[t] = make_double2(sum, 0.0);
}

// ---------------------------------------------------------------- transform

__device__ __forceinline__ double2 c_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// fft/pass.rs Pass + fft.rs Plan::line, one workgroup per output row r: the line read (mode
// 0 src row r, 1 src row r times gain row r, 2 src column r) into LDS at bit-reversed
// places (the plan's swaps), the first stage's (u + v, u - v), then each stage's
// t = v * w; rows at or past `live` are zero
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_line(const double2 *src, double2 *dst,
                                                        const double2 *gain, const double2 *tw,
                                                        uint32_t side, uint32_t bits, int mode,
                                                        uint32_t live) {
    __shared__ double2 line[1024];
    const uint32_t r = blockIdx.x;
    double2 *out = dst + (size_t)r * side;
    if (r >= live) {
        for (uint32_t c = threadIdx.x; c < side; c += GR_BLOCK) out[c] = make_double2(0.0, 0.0);
        return;
    }
    for (uint32_t c = threadIdx.x; c < side; c += GR_BLOCK) {
        double2 v;
        if (mode == 0)
            v = src[(size_t)r * side + c];
        else if (mode == 1)
            v = c_mul(src[(size_t)r * side + c], gain[(size_t)r * side + c]);
        else
            v = src[(size_t)c * side + r];
        line[__builtin_bitreverse32(c) >> (32 - bits)] = v;
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; 2 * q + 1 < side; q += GR_BLOCK) {
        const double2 u = line[2 * q], v = line[2 * q + 1];
        line[2 * q] = make_double2(u.x + v.x, u.y + v.y);
        line[2 * q + 1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
    for (uint32_t half = 2; half < side; half *= 2) {
        for (uint32_t b = threadIdx.x; b < side / 2; b += GR_BLOCK) {
            const uint32_t j = b & (half - 1), base = (b - j) * 2;
            const double2 u = line[base + j];
            const double2 t = c_mul(line[base + j + half], tw[half + j]);
            line[base + j] = make_double2(u.x + t.x, u.y + t.y);
            line[base + j + half] = make_double2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
    }
    for (uint32_t c = threadIdx.x; c < side; c += GR_BLOCK) out[c] = line[c];
}

// kernel.rs sample: G at every cell offset within reach, wrapped; where two offsets wrap to
// one cell (reach = side / 2), the one the reference's ascending loops write last
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_sample(const PmD3Dev m, double h, int reach) {
    const uint32_t side = m.side;
    const uint32_t t = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (t >= side * side) return;
    const int row = (int)(t / side), col = (int)(t - (uint32_t)row * side);
    auto offset = [&](int v, int &o) -> bool {
        if (v <= reach) {
            o = v;  // the larger candidate (v itself) is written last
            return true;
        }
        if (v - (int)side >= -reach) {
            o = v - (int)side;
            return true;
        }
        return false;
    };
    int dx, dy;
    double2 g = make_double2(0.0, 0.0);
    if (offset(row, dy) && offset(col, dx)) {
        const double rx = (double)dx * h, ry = (double)dy * h;
        double l = rx * rx + ry * ry;
        if (!(l == 0.0 || l >= m.dmax2)) {
            if (l < m.dmin2) l = sqrt(m.dmin2 * l);
            const double scale = 1.0 / (double)((size_t)side * side);
            g = make_double2(-rx / l * scale, -ry / l * scale);
        }
    }
    m.ks[t] = g;
}

// charge.rs Interpolate + mesh.rs read, merged through the slots: node order[k]'s delta
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_read(const D3SimDev s, const PmD3Dev m,
                                                        double strength) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= s.n) return;
    const uint32_t side = m.side, at = m.cell[k];
    double ex = 0.0, ey = 0.0;
    if (at < side * side) {
        const double fx = m.fxy[2 * (size_t)k], fy = m.fxy[2 * (size_t)k + 1];
        const uint32_t c[4] = {at, at + 1, at + side, at + side + 1};
        const double w[4] = {(1.0 - fx) * (1.0 - fy), fx * (1.0 - fy), (1.0 - fx) * fy, fx * fy};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double2 d = m.a[c[q]];
            ex += d.x * w[q];
            ey += d.y * w[q];
        }
    }
    const uint32_t i = m.order[k];
    s.dl[2 * (size_t)i] = ex * strength;
    s.dl[2 * (size_t)i + 1] = ey * strength;
}

// ---------------------------------------------------------------- hashed-grid collide

// Rust's `f as i64`: truncation, NaN to 0, saturating at the ends
__device__ __forceinline__ int64_t sat_i64(double v) {
    if (isnan(v)) return 0;
    if (v >= 9223372036854775807.0) return INT64_MAX;
    if (v <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)v;
}

struct PmHash {
    uint32_t shift;
    uint64_t mask;
    double ox, oy, size;
    __device__ void cell_of(double x, double y, int64_t &cx, int64_t &cy) const {
        cx = sat_i64((x - ox) / size);
        cy = sat_i64((y - oy) / size);
    }
    __device__ uint32_t bucket_of(int64_t cx, int64_t cy) const {
        const uint64_t row = ((uint64_t)cy * 0x9E3779B97F4A7C15ull) >> shift;
        return (uint32_t)((row + (uint64_t)cx) & mask);
    }
};

__device__ __forceinline__ PmHash pmd3_hash(const PmD3Dev &m, double size) {
    PmHash hs;
    hs.shift = 64u - m.bucket_bits;
    hs.mask = ((uint64_t)1 << m.bucket_bits) - 1;
    hs.ox = m.origin[0];
    hs.oy = m.origin[1];
    hs.size = size;
    return hs;
}

__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_bucket(const D3SimDev s, const PmD3Dev m,
                                                          double size) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= s.n) return;
    const PmHash hs = pmd3_hash(m, size);
    int64_t cx, cy;
    hs.cell_of(s.px[i], s.py[i], cx, cy);
    m.keys[i] = hs.bucket_of(cx, cy);
    m.vals[i] = i;
}

// the grid order from the sorted buckets: order[k], slot[order[k]] = k, and the sorted
// positions
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_order(const D3SimDev s, const PmD3Dev m,
                                                         const uint32_t *sorted) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= s.n) return;
    const uint32_t i = sorted[k];
    m.order[k] = i;
    m.slot[i] = k;
    m.at[k] = make_double2(s.px[i], s.py[i]);
}

// gather.rs: slot k's pushes from the candidates of its 3x3 cells' buckets, in the order
// Grid::reads lists them (three rows of three consecutive buckets, a bucket an earlier row
// already read skipped, touching runs joined), itself left out
__global__ __launch_bounds__(GR_BLOCK) void k_pmd3_gather(const D3SimDev s, const PmD3Dev m,
                                                          double size, uint32_t tick) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= s.n) return;
    const PmHash hs = pmd3_hash(m, size);
    const double d2 = s.reach2, reach = sqrt(d2);
    const double2 me = m.at[k];
    const uint32_t i = m.order[k];
    int64_t cx, cy;
    hs.cell_of(me.x, me.y, cx, cy);
    uint32_t lo[9], hi[9];
    int runs = 0;
    uint64_t firsts[3] = {0, 0, 0};
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const uint64_t first = hs.bucket_of((int64_t)((uint64_t)cx - 1u),
                                            (int64_t)((uint64_t)cy + (uint64_t)(int64_t)(r - 1)));
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const uint64_t b = (first + (uint64_t)q) & hs.mask;
            bool seen = false;
            for (int e = 0; e < r; e++) seen |= ((b - firsts[e]) & hs.mask) < 3;
            if (seen) continue;
            const uint32_t a = m.bstart[b], z = m.bstart[b + 1];
            if (a == z) continue;
            if (runs > 0 && hi[runs - 1] == a) {
                hi[runs - 1] = z;
            } else {
                lo[runs] = a;
                hi[runs] = z;
                runs++;
            }
        }
        firsts[r] = first;
    }
    double ox = 0.0, oy = 0.0;
    for (int r = 0; r < runs; r++)
        for (uint32_t q = lo[r]; q < hi[r]; q++) {
            if (q == k) continue;
            const double2 o = m.at[q];
            double dx = me.x - o.x, dy = me.y - o.y;
            double l = dx * dx + dy * dy;
            if (!(l < d2)) continue;  // NaN or out of reach
            const uint32_t j = m.order[q];
            if (dx == 0.0) {
                dx = d_jiggle_inline(s.seed, tick, 4, i, j);
                l += dx * dx;
            }
            if (dy == 0.0) {
                dy = d_jiggle_inline(s.seed, tick, 5, i, j);
                l += dy * dy;
            }
            const double dist = sqrt(l);
            const double push = (reach - dist) / dist * 0.5;
            ox += dx * push;
            oy += dy * push;
        }
    s.dl[2 * (size_t)i] = ox;
    s.dl[2 * (size_t)i + 1] = oy;
}

// ---------------------------------------------------------------- host launchers

static int key_bits_for(uint64_t maxkey) {
    int b = 1;
    while (b < 63 && ((uint64_t)1 << b) <= maxkey) b++;
    return b;
}

void pmd3_frame(const D3SimDev &s, const PmD3Dev &m) {
    const uint32_t nb = m.nparts;
    hipLaunchKernelGGL(k_pmd3_bounds, dim3(nb), dim3(GR_BLOCK), 0, s.stream, s.x, s.y, s.n,
                       m.part);
    hipLaunchKernelGGL(k_pmd3_frame, dim3(1), dim3(1), 0, s.stream, m.part, nb, m.fr, m.side,
                       s.n, m.dmax2);
}

void pmd3_deposit(const D3SimDev &s, const PmD3Dev &m, uint32_t cells) {
    const uint32_t n = s.n, nkeys = m.side * m.side;
    hipLaunchKernelGGL(k_pmd3_stencil, dim3(gr_blocks(n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, m);
    const int bits = key_bits_for(nkeys);
    radix_sort_pairs(m.keys, m.vals, m.keys_alt, m.vals_alt, n, bits, m.sort_tmp, s.stream);
    const bool odd = radix_digits(bits) & 1;
    const uint64_t *sk = odd ? m.keys_alt : m.keys;
    const uint32_t *sv = odd ? m.vals_alt : m.vals;
    hipLaunchKernelGGL(k_pmd3_starts, dim3(gr_blocks((uint64_t)nkeys + 1, GR_BLOCK)),
                       dim3(GR_BLOCK), 0, s.stream, sk, n, nkeys, m.cstart);
    hipLaunchKernelGGL(k_pmd3_deposit, dim3(gr_blocks((uint64_t)cells * m.side, GR_BLOCK)),
                       dim3(GR_BLOCK), 0, s.stream, m, sv, cells);
}

void pmd3_spectrum(const PmD3Dev &m, double h, int reach, hipStream_t stream) {
    const uint32_t side = m.side;
    hipLaunchKernelGGL(k_pmd3_sample, dim3(gr_blocks((uint64_t)side * side, GR_BLOCK)),
                       dim3(GR_BLOCK), 0, stream, m, h, reach);
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, stream, m.ks, m.b,
                       (const double2 *)nullptr, m.tw_fwd, side, m.bits, 0, side);
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, stream, m.b, m.ks,
                       (const double2 *)nullptr, m.tw_fwd, side, m.bits, 2, side);
}

void pmd3_solve_read(const D3SimDev &s, const PmD3Dev &m, uint32_t cells, double strength) {
    const uint32_t side = m.side;
    // forward: rows over the live rows, then every column; inverse: every row times the
    // spectrum, then the columns over the live rows
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, s.stream, m.a, m.b,
                       (const double2 *)nullptr, m.tw_fwd, side, m.bits, 0, cells);
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, s.stream, m.b, m.a,
                       (const double2 *)nullptr, m.tw_fwd, side, m.bits, 2, side);
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, s.stream, m.a, m.b,
                       (const double2 *)m.ks, m.tw_inv, side, m.bits, 1, side);
    hipLaunchKernelGGL(k_pmd3_line, dim3(side), dim3(GR_BLOCK), 0, s.stream, m.b, m.a,
                       (const double2 *)nullptr, m.tw_inv, side, m.bits, 2, cells);
    hipLaunchKernelGGL(k_pmd3_read, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, m, strength);
}

void pmd3_collide(const D3SimDev &s, const PmD3Dev &m, uint32_t tick) {
    const uint32_t n = s.n, nb = m.nparts;
    const double size = sqrt(s.reach2);
    hipLaunchKernelGGL(k_pmd3_bounds, dim3(nb), dim3(GR_BLOCK), 0, s.stream, s.px, s.py, n,
                       m.part);
    hipLaunchKernelGGL(k_pmd3_origin, dim3(1), dim3(1), 0, s.stream, m.part, nb, m.origin);
    hipLaunchKernelGGL(k_pmd3_bucket, dim3(gr_blocks(n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, m, size);
    const int bits = (int)m.bucket_bits;
    radix_sort_pairs(m.keys, m.vals, m.keys_alt, m.vals_alt, n, bits, m.sort_tmp, s.stream);
    const bool odd = radix_digits(bits) & 1;
    const uint64_t *sk = odd ? m.keys_alt : m.keys;
    const uint32_t *sv = odd ? m.vals_alt : m.vals;
    hipLaunchKernelGGL(k_pmd3_starts, dim3(gr_blocks(((uint64_t)1 << m.bucket_bits) + 1, GR_BLOCK)),
                       dim3(GR_BLOCK), 0, s.stream, sk, n, 1u << m.bucket_bits, m.bstart);
    hipLaunchKernelGGL(k_pmd3_order, dim3(gr_blocks(n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, m, sv);
    pmd3_collide_gather(s, m, tick);
}

void pmd3_collide_gather(const D3SimDev &s, const PmD3Dev &m, uint32_t tick) {
    hipLaunchKernelGGL(k_pmd3_gather, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, m, sqrt(s.reach2), tick);
}

}  // namespace gr
