// d3's many-body force over the quadtree on gfx950: the reference's Barnes-Hut charge pass
// (paths relative to crates/graph-core/src/layout/force/):
//   bodies   barnes_hut/charge.rs:120-140 + charge/threshold.rs: a leaf sits at its chain
//            head, an internal cell at the count-weighted mean of its children's centres in
//            slot order; open = w^2 / theta^2 with the threshold's NaN and overflow cases
//   node ids quadtree/preorder.rs:64-88 records, per cell, the shape slot the insertion
//            (quadtree/build.rs:25-115) pushed it as; the walk keys its jiggle on it
//   walk     barnes_hut/charge.rs:196-294 (node_delta, approx, direct, settle)
//
// The tree is the FA2 build's (the same PR quadtree, bit for bit); what the bulk build does
// not know is the order in which the insertion would have pushed the cells. It follows from
// the arena: the insertion of point p pushes, in this order, the internal cells of the split
// p causes (top down) and then p's own leaf, or only a leaf in an empty slot. So a leaf was
// pushed by its smallest point a (later points are chained into it), and an internal cell
// by the first point of its subtree that is not at the position of the subtree's smallest
// point a (every point before it was chained into a's leaf, which then spanned the cell's
// square). For an internal cell that point is min(its value for the child holding a, the
// smallest point of each other child); for a leaf it is the first point chained there by a
// split that stopped narrowing. The cells one insertion pushes lie on one root-to-leaf path,
// in preorder. A cell's node id is therefore its rank by (pushing point, preorder index).
#include <math.h>

#include "d3.h"
#include "fa2_common.h"
#include "sort_scan.h"

namespace gr {

// charge/threshold.rs:20-41
__device__ __forceinline__ double d3_open(double w, double theta, double theta2) {
    if (theta == 0.0) return INFINITY;
    const double naive = w * w / theta2;
    if (!isnan(naive)) return naive;
    if (!isfinite(theta)) return INFINITY;
    const double r = w / theta;
    return isfinite(r) ? r * r : INFINITY;
}

// One thread per cell: the leaf bodies (centre at the chain head), every cell's count,
// skip, start and threshold, the leaves' smallest and pushing points, and the depth
// difference array (an internal cell covers [k + 1, skip) of its descendants).
__global__ __launch_bounds__(GR_BLOCK) void k_d3_cells(const D3Dev d) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= d.C) return;
    const WalkCell<double, 2> *wc = (const WalkCell<double, 2> *)d.wc;
    const uint32_t skip = wc[k].skip, start = wc[k].start;
    const uint32_t end = wc[k].end < d.nf ? wc[k].end : d.nf;
    const double w = d.cb[(size_t)k * 4 + 2] - d.cb[(size_t)k * 4];
    D3Body b;
    b.open = d3_open(w, d.theta, d.theta2);
    b.count = end > start ? end - start : 0u;
    b.skip = skip;
    b.start = start;
    b.nid = 0;
    if (skip == k + 1) {
        b.comx = start < d.nf ? d.sx[start] : 0.0;
        b.comy = start < d.nf ? d.sy[start] : 0.0;
        uint32_t a = 0xFFFFFFFFu;
        double ax = 0.0, ay = 0.0;
        for (uint32_t q = start; q < end; q++) {
            const uint32_t i = d.order[q];
            if (i < a) {
                a = i;
                ax = d.sx[q];
                ay = d.sy[q];
            }
        }
        uint32_t p = 0xFFFFFFFFu;
        for (uint32_t q = start; q < end; q++) {
            const uint32_t i = d.order[q];
            if (!(d.sx[q] == ax && d.sy[q] == ay) && i < p) p = i;
        }
        d.amin[k] = a;
        d.bmin[k] = p;
    } else {
        b.comx = 0.0;
        b.comy = 0.0;
        atomicAdd(&d.diff[k + 1], 1u);
        atomicAdd(&d.diff[skip <= d.C ? skip : d.C], 0xFFFFFFFFu);  // -1 modulo 2^32
    }
    d.body[k] = b;
}

// Internal cells sort by depth (16-bit keys; leaves carry the largest key and sort last):
// every depth from 0 to the deepest internal cell's holds one, since an internal cell's
// ancestors are internal. Device-scope atomics into a few shared counters serialise across
// the XCDs, so the counting is the radix sort's own.
constexpr uint32_t D3_LEAF_KEY = 0xFFFFu;

// depth[k] = the inclusive prefix of diff (the scan left the exclusive one); the sort keys
__global__ __launch_bounds__(GR_BLOCK) void k_d3_depth(const D3Dev d) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= d.C) return;
    const uint32_t dep = d.depth[k] + d.diff[k];
    d.depth[k] = dep;
    const bool inner = d.body[k].skip != k + 1;
    d.keys[k] = inner ? (dep < D3_LEAF_KEY ? dep : D3_LEAF_KEY - 1u) : D3_LEAF_KEY;
    d.vals[k] = k;
}

// From the sorted keys: lvl[0] the deepest internal depth, lvl[1] the internal cell count,
// lvl[2 + l] the first sorted slot of depth l (lvl cleared before).
__global__ __launch_bounds__(GR_BLOCK) void k_d3_bounds(const D3Dev d, const uint64_t *sk) {
    const uint32_t r = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (r >= d.C) return;
    const uint32_t kr = (uint32_t)(sk[r] & 0xFFFFu);
    const uint32_t kp = r ? (uint32_t)(sk[r - 1] & 0xFFFFu) : 0x10000u;
    if (kr != kp) {
        if (kr != D3_LEAF_KEY) {
            if (kr < d.C) d.lvl[2 + kr] = r;
        } else if (r > 0) {
            d.lvl[1] = r;
            d.lvl[0] = kp;
        }
    }
    if (r == d.C - 1 && kr != D3_LEAF_KEY) {
        d.lvl[1] = d.C;
        d.lvl[0] = kr;
    }
}

// charge/threshold.rs:54-68 (the centre), and the smallest and pushing points, for the
// internal cells of one depth; their children are final (deeper, or leaves).
__global__ __launch_bounds__(GR_BLOCK) void k_d3_level(const D3Dev d, const uint32_t *bucket,
                                                       uint32_t off, uint32_t cnt) {
    const uint32_t x = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (x >= cnt) return;
    const uint32_t k = bucket[off + x];
    if (k >= d.C) return;
    const uint32_t skip = d.body[k].skip < d.C ? d.body[k].skip : d.C;
    double m = 0.0, cx = 0.0, cy = 0.0;
    uint32_t ca[4], cp[4];
    int nk = 0;
    uint32_t c = k + 1;
    // at most four children, every skip points forward: bounded on any arena
    for (int q = 0; q < 4 && c < skip; q++) {
        const D3Body ch = d.body[c];
        const double cm = (double)ch.count;
        m += cm;
        cx += cm * ch.comx;
        cy += cm * ch.comy;
        ca[nk] = d.amin[c];
        cp[nk] = d.bmin[c];
        nk++;
        c = ch.skip > c ? ch.skip : skip;
    }
    uint32_t a = 0xFFFFFFFFu;
    int first = 0;
    for (int q = 0; q < nk; q++)
        if (ca[q] < a) {
            a = ca[q];
            first = q;
        }
    uint32_t p = nk ? cp[first] : 0xFFFFFFFFu;
    for (int q = 0; q < nk; q++)
        if (q != first && ca[q] < p) p = ca[q];
    d.body[k].comx = cx / m;
    d.body[k].comy = cy / m;
    d.amin[k] = a;
    d.bmin[k] = p;
}

// Node ids: a stable sort of the cells (in preorder) by the point whose insertion pushed
// them ranks them by (pushing point, preorder index).
__global__ __launch_bounds__(GR_BLOCK) void k_d3_id_keys(const D3Dev d) {
    const uint32_t k = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (k >= d.C) return;
    d.keys[k] = d.body[k].skip == k + 1 ? d.amin[k] : d.bmin[k];
    d.vals[k] = k;
}

__global__ __launch_bounds__(GR_BLOCK) void k_d3_ids(const D3Dev d, const uint32_t *sorted) {
    const uint32_t r = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (r >= d.C) return;
    const uint32_t k = sorted[r];
    if (k >= d.C) return;
    d.body[k].nid = r;
    if (d.nid_out) d.nid_out[k] = r;
}

// charge.rs:262-294: false past distanceMax, else the gap with an exactly zero axis
// jiggled (keyed on (i, key)) and l raised to distanceMin
__device__ __forceinline__ bool d3_settle(const D3Dev &d, uint32_t i, uint32_t key, double &dx,
                                          double &dy, double &l) {
    if (l >= d.dmax2) return false;
    if (dx == 0.0) {
        dx = d_jiggle_inline(d.seed, d.tick, 2, i, key);
        l += dx * dx;
    }
    if (dy == 0.0) {
        dy = d_jiggle_inline(d.seed, d.tick, 3, i, key);
        l += dy * dy;
    }
    if (l < d.dmin2) l = sqrt(d.dmin2 * l);
    if (l == 0.0) l = d.dmin2 > 2.2250738585072014e-308 ? d.dmin2 : 2.2250738585072014e-308;
    return true;
}

__device__ __forceinline__ uint32_t d3_wave_min(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t o = __shfl_xor(v, m, GR_WAVE);
        v = o < v ? o : v;
    }
    return v;
}

// charge.rs:196-240 as one wave-uniform preorder scan per g consecutive tree positions:
// the wave visits the smallest of its lanes' next cells, so each lane still meets its own
// cells in preorder and adds its terms in the reference's order; a body is read once per
// wave (a scalar load).
constexpr uint32_t D3_WALK = 256;

// The arrays are kernel arguments of their own, marked __restrict__: with the stores to
// `out` known not to touch the bodies, the uniform body reads become scalar loads.
__global__ __launch_bounds__(D3_WALK) void k_d3_walk(const D3Dev d,
                                                     const D3Body *__restrict__ body,
                                                     const uint32_t *__restrict__ order,
                                                     const double *__restrict__ sx,
                                                     const double *__restrict__ sy,
                                                     double *__restrict__ out, uint32_t g) {
    const uint32_t per = (D3_WALK / GR_WAVE) * g;  // g consecutive tree positions per wave
    const uint32_t nb = (d.n + per - 1) / per;
    const uint32_t lb = gr_xcd_block(blockIdx.x, nb);
    if (lb >= nb) return;
    const uint32_t lane = threadIdx.x % GR_WAVE;
    const uint32_t p = lb * per + (threadIdx.x / GR_WAVE) * g + lane;
    // every node walks, the non-finite ones (tree positions nf..n, in no leaf) too
    const bool live = lane < g && p < d.n;
    if (__ballot(live) == 0) return;
    const uint32_t pc = live ? p : d.n - 1;
    const double xi = sx[pc], yi = sy[pc];
    const uint32_t i = order[pc];
    const uint32_t C = d.C;
    uint32_t next = live ? 0u : 0xFFFFFFFFu;
    double ox = 0.0, oy = 0.0;
    uint32_t k = gr_uniform(d3_wave_min(next));
    D3Body bd = body[k < C ? k : 0u];
    while (k < C) {
        const uint32_t k1 = k + 1;
        // most steps descend: the next body is read while this one is worked on
        const D3Body b1 = body[k1 < C ? k1 : k];
        if (next == k) {
            double dx = bd.comx - xi, dy = bd.comy - yi;
            double l = dx * dx + dy * dy;
            double ddx = 0.0, ddy = 0.0;
            bool add = true;
            uint32_t nx = k1;
            if (bd.open >= l) {
                if (bd.skip != k1) {
                    add = false;
                } else {
                    const uint32_t *chain = order + bd.start;
                    if (d3_settle(d, i, chain[0], dx, dy, l)) {
                        const double f = d.charge * d.alpha / l;
                        for (uint32_t q = 0; q < bd.count; q++) {
                            if (chain[q] == i) continue;
                            ddx += dx * f;
                            ddy += dy * f;
                        }
                    }
                }
            } else {
                if (d3_settle(d, i, bd.nid, dx, dy, l)) {
                    const double m = (double)bd.count;
                    const double f = m * d.charge * d.alpha / l;
                    ddx = dx * f;
                    ddy = dy * f;
                }
                nx = bd.skip > k ? bd.skip : k1;
            }
            if (add) {
                ox += ddx;
                oy += ddy;
            }
            next = nx;
        }
        // k + 1 when any lane waits for it, else the smallest wait
        const uint32_t kn = __ballot(next == k1) != 0 ? k1 : gr_uniform(d3_wave_min(next));
        if (kn == k1)
            bd = b1;
        else if (kn < C)
            bd = body[kn];
        k = kn;
    }
    if (live) {
        out[2 * (size_t)i] = ox;
        out[2 * (size_t)i + 1] = oy;
    }
}

size_t d3_tmp_bytes(uint32_t C) {
    const uint32_t c = C ? C : 1;
    const size_t a = scan_tmp_bytes(c + 1), b = radix_tmp_bytes(c);
    return a > b ? a : b;
}

void d3_launch_cells(const D3Dev &d) {
    (void)hipMemsetAsync(d.diff, 0, ((size_t)d.C + 1) * 4, d.stream);
    hipLaunchKernelGGL(k_d3_cells, dim3(gr_blocks(d.C, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream,
                       d);
}

const uint32_t *d3_launch_depth(const D3Dev &d) {
    scan_exclusive(d.diff, d.depth, d.C, nullptr, nullptr, d.tmp, d.stream);
    hipLaunchKernelGGL(k_d3_depth, dim3(gr_blocks(d.C, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream,
                       d);
    radix_sort_pairs(d.keys, d.vals, d.keys_alt, d.vals_alt, d.C, 16, d.tmp, d.stream);
    const bool odd = radix_digits(16) & 1;
    (void)hipMemsetAsync(d.lvl, 0, ((size_t)d.C + 3) * 4, d.stream);
    hipLaunchKernelGGL(k_d3_bounds, dim3(gr_blocks(d.C, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream,
                       d, odd ? d.keys_alt : d.keys);
    return odd ? d.vals_alt : d.vals;
}

void d3_launch_level(const D3Dev &d, const uint32_t *bucket, uint32_t off, uint32_t cnt) {
    hipLaunchKernelGGL(k_d3_level, dim3(gr_blocks(cnt, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream,
                       d, bucket, off, cnt);
}

void d3_launch_ids(const D3Dev &d) {
    uint32_t nbits = 1;
    while (nbits < 32 && (1ull << nbits) < (uint64_t)d.n) nbits++;
    hipLaunchKernelGGL(k_d3_id_keys, dim3(gr_blocks(d.C, GR_BLOCK)), dim3(GR_BLOCK), 0,
                       d.stream, d);
    radix_sort_pairs(d.keys, d.vals, d.keys_alt, d.vals_alt, d.C, (int)nbits, d.tmp, d.stream);
    const uint32_t *sorted = (radix_digits((int)nbits) & 1) ? d.vals_alt : d.vals;
    hipLaunchKernelGGL(k_d3_ids, dim3(gr_blocks(d.C, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream, d,
                       sorted);
}

// Bodies per wave of the two walks: 64 when that gives 4096 waves or more. A smaller tree
// gets narrower waves (32 bodies, the other lanes idle) and twice as many: its scans are
// latency-bound with few waves per SIMD, and a wave's steps fall with the bodies it carries
// (100k nodes: many-body 0.274 -> 0.257 ms, collide 0.266 -> 0.228 ms; 16 bodies gain no
// more; at 1M nodes, where 64-body waves already give 15 per SIMD, 32 bodies are slower: the
// many-body walk 1.40 against 1.16 ms). Each lane's terms do not depend on the lanes beside
// it: the deltas are the same bits.
static uint32_t d3_walk_group(uint32_t n) {
    return (uint64_t)n < (uint64_t)GR_WAVE * 4096 ? GR_WAVE / 2 : GR_WAVE;
}

void d3_launch_walk(const D3Dev &d) {
    const uint32_t g = d3_walk_group(d.n);
    const uint32_t nb = gr_blocks(d.n, (D3_WALK / GR_WAVE) * g);
    hipLaunchKernelGGL(k_d3_walk, dim3(gr_xcd_grid(nb)), dim3(D3_WALK), 0, d.stream, d,
                       (const D3Body *)d.body, d.order, d.sx, d.sy, d.out, g);
}

// ---------------------------------------------------------------- the rest of the tick

// link.rs force and share: the spring of one edge on the displaced positions (x + vx), an
// exactly zero axis jiggled (passes 0 and 1, keyed on the edge's ends), nothing at zero
// length; its two shares (the lower end f * (1 - bias), the higher -f * bias) go to the
// edge's slots in its ends' rows.
__global__ __launch_bounds__(GR_BLOCK) void k_sim_link_force(const D3SimDev s, double alpha,
                                                             uint32_t tick) {
    const uint32_t e = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (e >= s.m) return;
    const uint32_t a = s.lo[e], b = s.hi[e];
    double dx = (s.x[b] + s.vx[b]) - (s.x[a] + s.vx[a]);
    double dy = (s.y[b] + s.vy[b]) - (s.y[a] + s.vy[a]);
    if (dx == 0.0) dx = d_jiggle_inline(s.seed, tick, 0, a, b);
    if (dy == 0.0) dy = d_jiggle_inline(s.seed, tick, 1, a, b);
    const double l = sqrt(dx * dx + dy * dy);
    double fx = 0.0, fy = 0.0;
    if (l != 0.0) {
        const double factor = (l - s.ld[e]) / l * alpha * s.ls[e];
        fx = dx * factor;
        fy = dy * factor;
    }
    const double bias = s.lb[e];
    double2 *sh = (double2 *)s.sh;
    sh[s.slot[2 * (size_t)e]] = make_double2(fx * (1.0 - bias), fy * (1.0 - bias));
    sh[s.slot[2 * (size_t)e + 1]] = make_double2(-fx * bias, -fy * bias);
}

// step.rs LinkPass: each node's shares summed in its row's order (ascending edge), then the
// merge (share + 0). The row is contiguous, so eight slots are loaded ahead of their sums.
__global__ __launch_bounds__(GR_BLOCK) void k_sim_link_gather(const D3SimDev s) {
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (v >= s.n) return;
    const double2 *sh = (const double2 *)s.sh;
    const uint32_t q1 = s.roff[v + 1];
    uint32_t q = s.roff[v];
    double sx = 0.0, sy = 0.0;
    for (; q + 8 <= q1; q += 8) {
        double2 r[8];
#pragma unroll
        for (int j = 0; j < 8; j++) r[j] = sh[q + j];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            sx += r[j].x;
            sy += r[j].y;
        }
    }
    for (; q < q1; q++) {
        const double2 r = sh[q];
        sx += r.x;
        sy += r.y;
    }
    s.vx[v] += sx + 0.0;
    s.vy[v] += sy + 0.0;
}

__global__ __launch_bounds__(GR_BLOCK) void k_sim_merge(const D3SimDev s, const double *dl) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= s.n) return;
    s.vx[i] += dl[2 * (size_t)i] + 0.0;
    s.vy[i] += dl[2 * (size_t)i + 1] + 0.0;
}

// sim.rs center_shift: the two column sums in node order, (sum / n) * strength, into
// s.shift. One chain each in lane 0 of waves 0 (x) and 1 (y); the other waves stage the next
// tile in LDS; each chain reads 32 values of its tile into registers ahead of their adds.
constexpr uint32_t SIM_TILE = 2048;

__device__ __forceinline__ double chain_sum(double acc, const double *v, uint32_t len) {
    uint32_t q = 0;
    for (; q + 32 <= len; q += 32) {
        double r0[16], r1[16];
#pragma unroll
        for (int j = 0; j < 16; j++) r0[j] = v[q + j];
#pragma unroll
        for (int j = 0; j < 16; j++) r1[j] = v[q + 16 + j];
#pragma unroll
        for (int j = 0; j < 16; j++) acc = acc + r0[j];
#pragma unroll
        for (int j = 0; j < 16; j++) acc = acc + r1[j];
    }
    for (; q < len; q++) acc = acc + v[q];
    return acc;
}

__global__ __launch_bounds__(1024) void k_sim_center_sum(const D3SimDev s) {
    __shared__ double tile[2][2][SIM_TILE];
    const uint32_t n = s.n, tid = threadIdx.x;
    double acc = 0.0;
    const uint32_t ntiles = (n + SIM_TILE - 1) / SIM_TILE;
    for (uint32_t q = tid; q < SIM_TILE && q < n; q += blockDim.x) {
        tile[0][0][q] = s.x[q];
        tile[0][1][q] = s.y[q];
    }
    __syncthreads();
    for (uint32_t t = 0; t < ntiles; t++) {
        const uint32_t cur = t & 1, nxt = cur ^ 1;
        const uint32_t base = t * SIM_TILE, len = min(SIM_TILE, n - base);
        if (tid == 0 || tid == GR_WAVE) {
            acc = chain_sum(acc, tile[cur][tid == 0 ? 0 : 1], len);
        } else if (tid >= 2 * GR_WAVE && t + 1 < ntiles) {
            const uint32_t nb = base + SIM_TILE, nl = min(SIM_TILE, n - nb);
            for (uint32_t q = tid - 2 * GR_WAVE; q < nl; q += blockDim.x - 2 * GR_WAVE) {
                tile[nxt][0][q] = s.x[nb + q];
                tile[nxt][1][q] = s.y[nb + q];
            }
        }
        __syncthreads();
    }
    if (tid == 0) s.shift[0] = (acc / (double)n) * s.center_strength;
    if (tid == GR_WAVE) s.shift[1] = (acc / (double)n) * s.center_strength;
}

// sim.rs center, then collide.rs prepare: x -= shift; px = x + vx.
__global__ __launch_bounds__(GR_BLOCK) void k_sim_center_project(const D3SimDev s) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= s.n) return;
    const double x = s.x[i] - s.shift[0], y = s.y[i] - s.shift[1];
    s.x[i] = x;
    s.y[i] = y;
    s.px[i] = x + s.vx[i];
    s.py[i] = y + s.vy[i];
}

__global__ __launch_bounds__(GR_BLOCK) void k_sim_integrate(const D3SimDev s) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= s.n) return;
    const double vx = s.vx[i] * s.decay, vy = s.vy[i] * s.decay;
    s.vx[i] = vx;
    s.x[i] += vx;
    s.vy[i] = vy;
    s.y[i] += vy;
}

// collide.rs node_delta as one wave-uniform preorder scan per g consecutive tree
// positions (as the charge walk): a leaf's points are resolved pair by pair in the leaf's
// order, an internal cell is skipped by the lanes for which its square lies beyond the
// reach on some axis. Every node walks, non-finite ones too.
__global__ __launch_bounds__(D3_WALK) void k_sim_collide(const D3SimDev s, const D3CollideTree t,
                                                         uint32_t tick, uint32_t g) {
    const uint32_t per = (D3_WALK / GR_WAVE) * g;  // g consecutive tree positions per wave
    const uint32_t nb = (t.n + per - 1) / per;
    const uint32_t lb = gr_xcd_block(blockIdx.x, nb);
    if (lb >= nb) return;
    const uint32_t lane = threadIdx.x % GR_WAVE;
    const uint32_t p = lb * per + (threadIdx.x / GR_WAVE) * g + lane;
    const bool live = lane < g && p < t.n;
    if (__ballot(live) == 0) return;
    const uint32_t pc = live ? p : t.n - 1;
    const double xi = t.sx[pc], yi = t.sy[pc];
    const uint32_t i = t.order[pc];
    const double d2 = s.reach2, reach = sqrt(d2);
    const WalkCell<double, 2> *wc = (const WalkCell<double, 2> *)t.wc;
    uint32_t next = live ? 0u : 0xFFFFFFFFu;
    double ox = 0.0, oy = 0.0;
    uint32_t k = gr_uniform(d3_wave_min(next));
    while (k < t.C) {
        const uint32_t skip = wc[k].skip, k1 = k + 1;
        if (next == k) {
            if (skip == k1) {
                for (uint32_t q = wc[k].start; q < wc[k].end; q++) {
                    const uint32_t pj = t.order[q];
                    if (pj == i) continue;
                    double dx = xi - t.sx[q], dy = yi - t.sy[q];
                    double l = dx * dx + dy * dy;
                    if (l >= d2) continue;
                    if (dx == 0.0) {
                        dx = d_jiggle_inline(s.seed, tick, 4, i, pj);
                        l += dx * dx;
                    }
                    if (dy == 0.0) {
                        dy = d_jiggle_inline(s.seed, tick, 5, i, pj);
                        l += dy * dy;
                    }
                    const double dist = sqrt(l);
                    const double push = (reach - dist) / dist * 0.5;
                    ox += dx * push;
                    oy += dy * push;
                }
                next = k1;
            } else {
                const double *b = t.cb + (size_t)k * 4;  // x0, y0, x1, y1
                const bool outside = b[0] > xi + reach || b[2] < xi - reach ||
                                     b[1] > yi + reach || b[3] < yi - reach;
                next = outside ? (skip > k ? skip : k1) : k1;
            }
        }
        k = __ballot(next == k1) != 0 ? k1 : gr_uniform(d3_wave_min(next));
    }
    if (live) {
        s.dl[2 * (size_t)i] = ox;
        s.dl[2 * (size_t)i + 1] = oy;
    }
}

void d3sim_link(const D3SimDev &s, double alpha, uint32_t tick) {
    if (s.m)
        hipLaunchKernelGGL(k_sim_link_force, dim3(gr_blocks(s.m, GR_BLOCK)), dim3(GR_BLOCK), 0,
                           s.stream, s, alpha, tick);
    hipLaunchKernelGGL(k_sim_link_gather, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                       s.stream, s);
}

void d3sim_merge(const D3SimDev &s, const double *deltas) {
    hipLaunchKernelGGL(k_sim_merge, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0, s.stream,
                       s, deltas);
}

void d3sim_center_sum(const D3SimDev &s, hipStream_t stream) {
    if (s.n) hipLaunchKernelGGL(k_sim_center_sum, dim3(1), dim3(1024), 0, stream, s);
}

void d3sim_center_project(const D3SimDev &s) {
    hipLaunchKernelGGL(k_sim_center_project, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                       s.stream, s);
}

void d3sim_collide(const D3SimDev &s, const D3CollideTree &t, uint32_t tick) {
    const uint32_t g = d3_walk_group(t.n);
    const uint32_t nb = gr_blocks(t.n, (D3_WALK / GR_WAVE) * g);
    hipLaunchKernelGGL(k_sim_collide, dim3(gr_xcd_grid(nb)), dim3(D3_WALK), 0, s.stream, s, t,
                       tick, g);
}

void d3sim_integrate(const D3SimDev &s) {
    hipLaunchKernelGGL(k_sim_integrate, dim3(gr_blocks(s.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                       s.stream, s);
}

}  // namespace gr
