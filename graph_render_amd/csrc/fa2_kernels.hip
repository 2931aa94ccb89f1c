// ForceAtlas2 with Barnes-Hut repulsion on gfx950: tree build, centre of mass, the
// wave-coherent stackless walk, CSR attraction, gravity, the adaptive speed and the step.
//
// Reference semantics (paths relative to crates/graph-core/src/):
//   tree          layout/force/quadtree.rs:116-195, quadtree/{bounds,build,preorder}.rs
//                 (3D: layout/force/octree.rs, slot z<<2|y<<1|x)
//   centre        layout/forceatlas2/state/barnes_hut.rs:62-110
//   walk, pairs   layout/forceatlas2/state/barnes_hut.rs:112-184
//   attraction    layout/forceatlas2/state.rs:152-162
//   dense pairs   layout/forceatlas2/state.rs:169-203
//   gravity       layout/forceatlas2/state.rs:208-251
//   speed, step   layout/forceatlas2/state/schedule.rs:16-77
// FP64 kernels keep the reference's operand order and are built without FMA contraction,
// so the tree topology, the centres and the per-body force sums match it bit for bit.
#include <math.h>

#include <type_traits>

#include "fa2_common.h"
#include "fa2_kernels.h"
#include "sort_scan.h"

namespace gr {

// ------------------------------------------------------------------ common

template <int D>
struct KeyFmt {
    // Levels held in one 64-bit key (D bits each); bit 63 marks a non-finite point.
    static constexpr int LEVELS = D == 2 ? 31 : 21;
    static constexpr int USED = D * LEVELS;
};

size_t walkcell_bytes(int precision, uint32_t dim) {
    if (precision == PREC_FP64)
        return dim == 2 ? sizeof(WalkCell<double, 2>) : sizeof(WalkCell<double, 3>);
    return dim == 2 ? sizeof(WalkCell<float, 2>) : sizeof(WalkCell<float, 3>);
}

template <int D>
struct Box {
    double lo[D], hi[D];
};

// bounds.rs:34-50 (octree/bounds.rs narrow): the slot's bit a is set when p[a] >= mid[a].
template <int D>
__device__ __forceinline__ int box_narrow(Box<D> &b, const double *p) {
    int slot = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        double m = (b.lo[a] + b.hi[a]) / 2.0;
        if (p[a] >= m) {
            b.lo[a] = m;
            slot |= 1 << a;
        } else {
            b.hi[a] = m;
        }
    }
    return slot;
}

template <int D>
__device__ __forceinline__ int box_slot(const Box<D> &b, const double *p) {
    int slot = 0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        double m = (b.lo[a] + b.hi[a]) / 2.0;
        if (p[a] >= m) slot |= 1 << a;
    }
    return slot;
}

// bounds.rs:62-86 quadrant / octree/bounds.rs octant.
template <int D>
__device__ __forceinline__ void box_quad(Box<D> &b, int slot) {
#pragma unroll
    for (int a = 0; a < D; a++) {
        double m = (b.lo[a] + b.hi[a]) / 2.0;
        if (slot & (1 << a))
            b.lo[a] = m;
        else
            b.hi[a] = m;
    }
}

template <int D>
__device__ __forceinline__ Box<D> root_box(const GrGraph &g) {
    Box<D> b;
#pragma unroll
    for (int a = 0; a < D; a++) {
        b.lo[a] = g.root_lo[a];
        b.hi[a] = g.root_hi[a];
    }
    return b;
}

// Key layout: [non-finite][graph id][L levels of D slot bits], level 1 highest, in the low
// key_bits bits.
template <int D>
__device__ __forceinline__ int key_slot(uint64_t key, int level, int L) {
    return (int)((key >> (D * (L - level))) & ((1u << D) - 1));
}

// Levels two finite keys share below their root; -1 when they are in different graphs.
template <int D>
__device__ __forceinline__ int key_lcp(uint64_t a, uint64_t b, int L) {
    const uint64_t x = a ^ b;
    if (x >> (D * L)) return -1;
    if (x == 0) return L;
    const int lz = __clzll((long long)x) - (64 - D * L);
    return lz / D;
}

// Same graph and the same first `level` slots (level 0: the same graph).
template <int D>
__device__ __forceinline__ bool key_same(uint64_t a, uint64_t b, int level, int L) {
    return ((a ^ b) >> (D * (L - level))) == 0;
}

// Compensated (Kahan) accumulation, used by the FP32+Kahan precision.
template <typename T, bool KAHAN>
__device__ __forceinline__ void acc_add(T &s, T &c, T v) {
    if (KAHAN) {
        T y = v - c;
        T t = s + y;
        c = (t - s) - y;
        s = t;
    } else {
        s = s + v;
    }
}

constexpr int NB = 10;  // bounds partial: lo[3], hi[3], sum[3], finite count

// One point's contribution to the bounds/centroid partials (quadtree.rs:46-53, the finite
// extent; state.rs:208-220, the centroid over every point).
template <int D>
__device__ __forceinline__ void bounds_take(const double *p, double *lo, double *hi, double *sum,
                                            double &nfin) {
    bool fin = true;
#pragma unroll
    for (int a = 0; a < D; a++) {
        sum[a] = sum[a] + p[a];
        if (!isfinite(p[a])) fin = false;
    }
    if (fin) {
#pragma unroll
        for (int a = 0; a < D; a++) {
            lo[a] = fmin(lo[a], p[a]);
            hi[a] = fmax(hi[a], p[a]);
        }
        nfin += 1.0;
    }
}

template <int D>
__device__ __forceinline__ void bounds_write(double *lo, double *hi, double *sum, double nfin,
                                             double *bpart) {
    __shared__ double lds[4 * 3];
    double vmin[3] = {lo[0], lo[1], D == 3 ? lo[2] : 0.0};
    double vmax[3] = {hi[0], hi[1], D == 3 ? hi[2] : 0.0};
    double vsum[4] = {sum[0], sum[1], D == 3 ? sum[2] : 0.0, nfin};
    block_reduce<3, 1>(vmin, lds);
    block_reduce<3, 2>(vmax, lds);
    __shared__ double lds4[4 * 4];
    block_reduce<4, 0>(vsum, lds4);
    if (threadIdx.x == 0) {
        double *o = bpart + (size_t)blockIdx.x * NB;
        for (int a = 0; a < 3; a++) {
            o[a] = vmin[a];
            o[3 + a] = vmax[a];
            o[6 + a] = vsum[a];
        }
        o[9] = vsum[3];
    }
}

// ------------------------------------------------------------------ bounds / root

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_bounds(const FaDev d) {
    const T *P[3] = {(const T *)d.pos[0], (const T *)d.pos[1],
                     (const T *)(D == 3 ? d.pos[2] : d.pos[0])};
    double lo[D], hi[D], sum[D], nfin = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
        sum[a] = 0.0;
    }
    const uint32_t c0 = d.chunk[blockIdx.x], c1 = d.chunk[blockIdx.x + 1];
    for (uint32_t i = c0 + threadIdx.x; i < c1; i += GR_BLOCK) {
        double p[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            p[a] = (double)P[a][i];
            ((T *)d.pi)[(size_t)i * D + a] = P[a][i];
            ((T *)d.pk)[(size_t)i * 4 + a] = P[a][i];
        }
        ((T *)d.pk)[(size_t)i * 4 + D] = ((const T *)d.mass)[i];
        bounds_take<D>(p, lo, hi, sum, nfin);
    }
    bounds_write<D>(lo, hi, sum, nfin, d.bpart);
}

// quadtree.rs:157-195 (octree.rs:220-254): the root square from the floor of the first
// corner, doubled until it holds the second; a square that stops growing is left as is.
template <int D>
__device__ void cover_root(GrGraph *st, const double *lo, const double *hi) {
    double blo[D], bhi[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        double h = floor(lo[a]) + 1.0;
        blo[a] = h - 1.0;
        bhi[a] = h;
    }
    double z = (bhi[0] - blo[0] != 0.0) ? bhi[0] - blo[0] : 1.0;
    for (;;) {
        bool out = false;
#pragma unroll
        for (int a = 0; a < D; a++)
            if (blo[a] > hi[a] || hi[a] >= bhi[a]) out = true;
        if (!out) break;
        z *= 2.0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            if (hi[a] < blo[a])
                blo[a] = bhi[a] - z;
            else
                bhi[a] = blo[a] + z;
        }
        if (!(isfinite(z) && isfinite(bhi[0] - blo[0]))) break;
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
        st->root_lo[a] = blo[a];
        st->root_hi[a] = bhi[a];
    }
}

// Per graph (one workgroup each): its extent, centroid and root square from the bounds
// partials of its node chunks, reduced in a fixed order.
template <int D>
__global__ __launch_bounds__(GR_BLOCK) void k_head_graph(const FaDev d) {
    if (halted(d)) return;
    const uint32_t b = blockIdx.x;
    GrGraph *g = d.gr + b;
    __shared__ double lds[4 * 4];
    double vmin[3] = {INFINITY, INFINITY, INFINITY};
    double vmax[3] = {-INFINITY, -INFINITY, -INFINITY};
    double vsum[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t c = d.gchunk[b] + threadIdx.x; c < d.gchunk[b + 1]; c += GR_BLOCK) {
        const double *o = d.bpart + (size_t)c * NB;
        for (int a = 0; a < 3; a++) {
            vmin[a] = fmin(vmin[a], o[a]);
            vmax[a] = fmax(vmax[a], o[3 + a]);
            vsum[a] = vsum[a] + o[6 + a];
        }
        vsum[3] = vsum[3] + o[9];
    }
    block_reduce<3, 1>(vmin, lds);
    block_reduce<3, 2>(vmax, lds);
    block_reduce<4, 0>(vsum, lds);
    if (threadIdx.x == 0) {
        const uint32_t nfin = (uint32_t)vsum[3];
        const double nb = (double)(d.gnode[b + 1] - d.gnode[b]);
        g->n_finite = nfin;
        g->has_root = nfin > 0;
        g->cell_lo = g->cell_hi = 0;
        for (int a = 0; a < 3; a++) {
            g->box_lo[a] = vmin[a];
            g->box_hi[a] = vmax[a];
        }
        for (int a = 0; a < D; a++) g->mean[a] = vsum[a] / nb;
        if (nfin > 0) cover_root<D>(g, vmin, vmax);
        if (d.ngraphs == 1) {  // the session totals are this graph's (k_head_global's job)
            GrState *st = d.st;
            st->pool_used = 0;
            st->lcell_used = 0;
            st->span_count = 0;
            st->any_deep = 0;
            st->n_finite = nfin;
            st->live = g->done ? 0u : 1u;
        }
    }
}

// Session-wide: counters of this iteration's build, points in trees, graphs still live.
__global__ __launch_bounds__(GR_BLOCK) void k_head_global(const FaDev d) {
    GrState *st = d.st;
    if (halted(d)) return;
    __shared__ uint32_t s_fin, s_live;
    if (threadIdx.x == 0) {
        s_fin = 0;
        s_live = 0;
    }
    __syncthreads();
    uint32_t fin = 0, live = 0;
    for (uint32_t b = threadIdx.x; b < d.ngraphs; b += GR_BLOCK) {
        fin += d.gr[b].n_finite;
        live += d.gr[b].done ? 0u : 1u;
    }
    atomicAdd(&s_fin, fin);
    atomicAdd(&s_live, live);
    __syncthreads();
    if (threadIdx.x == 0) {
        st->pool_used = 0;
        st->lcell_used = 0;
        st->span_count = 0;
        st->any_deep = 0;
        st->n_finite = s_fin;
        st->live = s_live;
    }
}

// ------------------------------------------------------------------ keys

// Each point replays the reference's own descent (narrow from the root square) for LEVELS
// levels, so its key bits are exactly the slots d3's insertion would take. A level whose
// narrowing no longer shrinks the x-span is where the reference stops splitting and chains
// instead (build.rs:76-85); the key stops there and the level is kept as the point's stall.
// True when every midpoint of a descent from this root to `deepest` levels is exact: a
// square root of power-of-two side w0 whose corners stay below w0 * 2^(50 - deepest) in
// magnitude, so each level's corners are multiples of its width inside 52 bits. Then the
// cell at level l is exactly w0 / 2^l wide, the width the reference's narrowing gives.
template <int D>
__device__ __forceinline__ bool dyadic_root(const Box<D> &b, int deepest) {
    const double w0 = b.hi[0] - b.lo[0];
    int ex = 0;
    if (!(w0 > 0.0) || !isfinite(w0) || frexp(w0, &ex) != 0.5) return false;
    const double lim = ldexp(w0, 50 - deepest);
    bool ok = true;
#pragma unroll
    for (int a = 0; a < D; a++) {
        if (b.hi[a] - b.lo[a] != w0) ok = false;
        if (!(fabs(b.lo[a]) < lim && fabs(b.hi[a]) < lim)) ok = false;
    }
    return ok;
}

// Bit i of x to bit 2i (x < 2^32), and to bit 3i (x < 2^21): the Morton interleave.
__device__ __forceinline__ uint64_t spread2(uint64_t x) {
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
__device__ __forceinline__ uint64_t spread3(uint64_t x) {
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_keys(const FaDev d) {
    if (halted(d)) return;
    uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= d.n) return;
    double p[D];
    bool fin = true;
#pragma unroll
    for (int a = 0; a < D; a++) {
        p[a] = (double)((const T *)d.pos[a])[i];
        if (!isfinite(p[a])) fin = false;
    }
    uint64_t key = 1ull << (d.key_bits - 1);  // non-finite: after every finite key
    uint32_t stall = 0xFF;
    if (fin) {
        const int L = (int)d.levels;
        const uint32_t gb = graph_of(d, i);
        key = (uint64_t)gb << (D * L);
        Box<D> b = root_box<D>(d.gr[gb]);
        if (dyadic_root<D>(b, L)) {
            // Every midpoint of the descent is exact, so no level stalls and the slot bits
            // of level l are bit L - l of each axis' cell index q = max{q : lo + q w_L <= p}:
            // one product gives q to within one, exact grid comparisons settle it.
            const double w0 = b.hi[0] - b.lo[0];
            const double wl = ldexp(w0, -L), inv = ldexp(1.0, L) / w0;
            const int64_t top = ((int64_t)1 << L) - 1;
            uint64_t mort = 0;
#pragma unroll
            for (int a = 0; a < D; a++) {
                int64_t q = (int64_t)floor((p[a] - b.lo[a]) * inv);
                q = q < 0 ? 0 : (q > top ? top : q);
                if (b.lo[a] + (double)q * wl > p[a])
                    q--;
                else if (q < top && b.lo[a] + (double)(q + 1) * wl <= p[a])
                    q++;
                mort |= (D == 2 ? spread2((uint64_t)q) : spread3((uint64_t)q)) << a;
            }
            key |= mort;
        } else {
            for (int l = 1; l <= L; l++) {
                double before = b.hi[0] - b.lo[0];
                int slot = box_narrow<D>(b, p);
                double after = b.hi[0] - b.lo[0];
                if (!(after < before)) {
                    stall = (uint32_t)(l - 1);
                    break;
                }
                key |= (uint64_t)slot << (D * (L - l));
            }
        }
    }
    // Reverse input order: the stable sort then leaves equal keys in descending point
    // index, the order a leaf chains its points in (most recent first, build.rs:112-115).
    uint32_t j = d.n - 1 - i;
    d.keys[j] = key;
    d.vals[j] = i;
    d.stall[i] = (uint8_t)stall;
}

// ------------------------------------------------------------------ groups

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_gather(const FaDev d) {
    const GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (p >= d.n) return;
    const uint32_t nf = sh.n_finite;
    // Non-finite points follow the tree's points in ascending index (preorder.rs:54-59);
    // the sort left them descending, so that tail is read backwards.
    uint32_t src = p < nf ? p : nf + (d.n - 1 - p);
    uint32_t i = d.vals[src];
    uint64_t key = d.keys[src];
    const uint64_t kprev = d.keys[p > 0 ? p - 1 : 0];  // loaded with the others
    d.order[p] = i;
    // position and mass of point i from one 4-slot record (one cache line)
    const T *pk = (const T *)d.pk + (size_t)i * 4;
    T rec[4];
#pragma unroll
    for (int a = 0; a < 4; a++) rec[a] = pk[a];
#pragma unroll
    for (int a = 0; a < D; a++) ((T *)d.sp[a])[p] = rec[a];
    ((T *)d.sm)[p] = rec[D];
    d.flag[p] = (p < nf && (p == 0 || kprev != key)) ? 1u : 0u;
}

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_group_compact(const FaDev d) {
    const GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (p >= sh.n_finite) return;
    if (!d.flag[p]) {
        // a later point of a group: does its position differ from the previous one's?
        uint32_t distinct = 0;
#pragma unroll
        for (int a = 0; a < D; a++)
            if (((const T *)d.sp[a])[p] != ((const T *)d.sp[a])[p - 1]) distinct = 1;
        d.dflag[p] = distinct;
        return;
    }
    d.dflag[p] = 0;
    uint32_t g = d.gid[p];
    d.gstart[g] = p;
    d.gkey[g] = d.keys[p];
}

enum GroupType : uint32_t { G_SIMPLE = 0, G_STALL = 1, G_DEEP = 2 };

// The subtree under a deep group's level-LEVELS cell: the reference's own insertion
// (build.rs:26-115) over the group's points in ascending index, then its preorder flatten
// (preorder.rs:41-99), written to the local-cell pool. Runs on one lane per deep group;
// deep groups (distinct points sharing LEVELS levels) are rare.
template <typename T, int D>
__device__ bool local_build(const FaDev &d, uint32_t s, uint32_t e, Box<D> rootb,
                            uint32_t *out_lc, uint32_t *out_count) {
    GrState *st = d.st;
    constexpr int NS = 1 + (1 << D);
    int32_t *pool = d.pool;
    auto coords = [&](uint32_t i, double *p) {
#pragma unroll
        for (int a = 0; a < D; a++) p[a] = (double)((const T *)d.pos[a])[i];
    };
    auto push = [&](int32_t head) -> int32_t {
        uint32_t at = atomicAdd(&st->pool_used, (uint32_t)NS);
        if (at + NS > d.pool_cap) {
            st->pool_overflow = 1;
            return -1;
        }
        pool[at] = head;
        for (int c = 0; c < (1 << D); c++) pool[at + 1 + c] = -1;
        return (int32_t)at;
    };
    int32_t root = -1;
    uint32_t nshapes = 0;
    bool ok = true;
    for (uint32_t q = e; q-- > s && ok;) {
        uint32_t pt = d.order[q];
        d.chain_next[pt] = -1;
        double p[D];
        coords(pt, p);
        if (root < 0) {
            root = push((int32_t)pt);
            if (root < 0) ok = false;
            nshapes++;
            continue;
        }
        int32_t node = root, parent = -1;
        int pslot = 0;
        Box<D> b = rootb;
        for (;;) {
            if (pool[node] >= 0) {
                // insert_leaf (build.rs:60-96)
                int32_t head = pool[node];
                double h[D];
                coords((uint32_t)head, h);
                bool same = true;
                for (int a = 0; a < D; a++)
                    if (p[a] != h[a]) same = false;
                if (same) {
                    d.chain_next[pt] = head;
                    pool[node] = (int32_t)pt;
                    break;
                }
                int32_t at = parent;
                int slot = parent >= 0 ? pslot : 0;
                for (;;) {
                    double before = b.hi[0] - b.lo[0];
                    Box<D> pre = b;
                    int i = box_narrow<D>(b, p);
                    int j = box_slot<D>(pre, h);
                    double after = b.hi[0] - b.lo[0];
                    if (!(after < before)) {
                        if (at >= 0)
                            pool[at + 1 + slot] = node;
                        else
                            root = node;
                        d.chain_next[pt] = pool[node];
                        pool[node] = (int32_t)pt;
                        break;
                    }
                    int32_t internal = push(-1);
                    if (internal < 0) {
                        ok = false;
                        break;
                    }
                    nshapes++;
                    if (at >= 0)
                        pool[at + 1 + slot] = internal;
                    else
                        root = internal;
                    if (i != j) {
                        pool[internal + 1 + j] = node;
                        int32_t leaf = push((int32_t)pt);
                        if (leaf < 0) {
                            ok = false;
                            break;
                        }
                        nshapes++;
                        pool[internal + 1 + i] = leaf;
                        break;
                    }
                    at = internal;
                    slot = i;
                }
                break;
            }
            int slot = box_narrow<D>(b, p);
            int32_t child = pool[node + 1 + slot];
            if (child >= 0) {
                parent = node;
                pslot = slot;
                node = child;
            } else {
                int32_t leaf = push((int32_t)pt);
                if (leaf < 0) {
                    ok = false;
                    break;
                }
                nshapes++;
                pool[node + 1 + slot] = leaf;
                break;
            }
        }
    }
    if (!ok) return false;
    // flatten into local cells; stack entries (node or -1, cell/parent, slot)
    uint32_t lc = atomicAdd(&st->lcell_used, nshapes);
    uint32_t stk = atomicAdd(&st->pool_used, 3u * (2u * nshapes + 2u));
    if (lc + nshapes > d.lcell_cap || stk + 3u * (2u * nshapes + 2u) > d.pool_cap) {
        st->pool_overflow = 1;
        return false;
    }
    int32_t *stack = pool + stk;
    int top = 0;
    stack[0] = root;
    stack[1] = -1;
    stack[2] = 0;
    top = 1;
    uint32_t ncell = 0, cursor = 0;
    while (top > 0) {
        top--;
        int32_t node = stack[3 * top], a1 = stack[3 * top + 1], a2 = stack[3 * top + 2];
        if (node < 0) {  // close cell a1
            d.lc_int[(size_t)(lc + a1) * 4 + 0] = (int32_t)ncell;
            d.lc_int[(size_t)(lc + a1) * 4 + 2] = (int32_t)(s + cursor);
            continue;
        }
        uint32_t k = ncell++;
        double *bk = d.lc_dbl + (size_t)(lc + k) * 2 * D;
        Box<D> b;
        if (a1 < 0) {
            b = rootb;
        } else {
            const double *bp = d.lc_dbl + (size_t)(lc + a1) * 2 * D;
            for (int a = 0; a < D; a++) {
                b.lo[a] = bp[a];
                b.hi[a] = bp[D + a];
            }
            box_quad<D>(b, a2);
        }
        for (int a = 0; a < D; a++) {
            bk[a] = b.lo[a];
            bk[D + a] = b.hi[a];
        }
        int32_t *ik = d.lc_int + (size_t)(lc + k) * 4;
        ik[1] = (int32_t)(s + cursor);
        ik[3] = a1;
        if (pool[node] >= 0) {
            for (int32_t q = pool[node]; q >= 0; q = d.chain_next[q]) {
                uint32_t at = s + cursor++;
                d.order[at] = (uint32_t)q;
#pragma unroll
                for (int a = 0; a < D; a++) ((T *)d.sp[a])[at] = ((const T *)d.pos[a])[q];
                ((T *)d.sm)[at] = ((const T *)d.mass)[q];
            }
            ik[0] = (int32_t)(k + 1);
            ik[2] = (int32_t)(s + cursor);
        } else {
            stack[3 * top] = -1;
            stack[3 * top + 1] = (int32_t)k;
            stack[3 * top + 2] = 0;
            top++;
            for (int slot = (1 << D) - 1; slot >= 0; slot--) {
                int32_t c = pool[node + 1 + slot];
                if (c >= 0) {
                    stack[3 * top] = c;
                    stack[3 * top + 1] = (int32_t)k;
                    stack[3 * top + 2] = slot;
                    top++;
                }
            }
        }
    }
    // the local subtree's sums, in the reference's aggregate order (reverse preorder, a
    // leaf over its chain, an internal cell over its children in slot order)
    for (uint32_t c = nshapes; c-- > 0;) {
        const int32_t *ik = d.lc_int + (size_t)(lc + c) * 4;
        double sum[1 + D];
        for (int a = 0; a <= D; a++) sum[a] = 0.0;
        if ((uint32_t)ik[0] == c + 1) {
            for (int32_t q = ik[1]; q < ik[2]; q++) {
                double m = (double)((const T *)d.sm)[q];
                sum[0] += m;
                for (int a = 0; a < D; a++) sum[1 + a] += m * (double)((const T *)d.sp[a])[q];
            }
        } else {
            uint32_t ch = c + 1;
            for (int q = 0; q < (1 << D) && ch < (uint32_t)ik[0]; q++) {
                const double *cs = d.lc_sum + (size_t)(lc + ch) * (1 + D);
                for (int a = 0; a <= D; a++) sum[a] += cs[a];
                uint32_t nx = (uint32_t)d.lc_int[(size_t)(lc + ch) * 4];
                ch = nx > ch ? nx : (uint32_t)ik[0];
            }
        }
        double *o = d.lc_sum + (size_t)(lc + c) * (1 + D);
        for (int a = 0; a <= D; a++) o[a] = sum[a];
    }
    *out_lc = lc;
    *out_count = nshapes;
    return true;
}

// Per group: how many cells it owns in preorder. A group owns the internal cells from one
// level below its shared prefix with the previous group down to its own depth, then its
// leaf (or, for a deep group, its local subtree).
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_group_info(const FaDev d) {
    GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    uint32_t g = blockIdx.x * GR_BLOCK + threadIdx.x;
    const uint32_t G = sh.n_groups;
    if (g >= G) return;
    const int L = (int)d.levels;
    // the neighbouring groups' keys and starts, loaded together (clamped, unconditional)
    const uint32_t gp = g > 0 ? g - 1 : g, gn = g + 1 < G ? g + 1 : g;
    const uint64_t key = d.gkey[g], kprev = d.gkey[gp], knext = d.gkey[gn];
    const uint32_t s = d.gstart[g], snext = d.gstart[gn];
    int left = g > 0 ? key_lcp<D>(kprev, key, L) : -1;
    int right = g + 1 < G ? key_lcp<D>(key, knext, L) : -1;
    uint32_t e = g + 1 < G ? snext : sh.n_finite;
    bool distinct = false;
    for (uint32_t q = s + 1; q < e; q++)
        if (d.dflag[q]) {
            distinct = true;
            break;
        }
    uint32_t type, level = 0, ncells;
    if (!distinct) {
        type = G_SIMPLE;
        level = (uint32_t)((left > right ? left : right) + 1);
        ncells = (uint32_t)((right > left ? right - left : 0) + 1);
    } else {
        uint32_t stall = d.stall[d.order[s]];
        if (stall != 0xFF) {
            type = G_STALL;
            level = stall;
            ncells = (uint32_t)((int)stall - left);
        } else {
            type = G_DEEP;
            level = L;
            Box<D> b = root_box<D>(d.gr[graph_of(d, d.order[s])]);
            for (int l = 1; l <= L; l++) box_quad<D>(b, key_slot<D>(key, l, L));
            uint32_t lc = 0, lcount = 0;
            if (!local_build<T, D>(d, s, e, b, &lc, &lcount)) {
                st->halt = 1;
                lcount = 1;
            }
            d.glocal[g] = lc;
            ncells = (uint32_t)((L - 1) - left) + lcount;
            st->any_deep = 1;
        }
    }
    d.gcells[g] = ncells;
    d.ginfo[g] = type | (level << 8);
}

// The last group h in [lo, hi) whose key shares `level` levels with `key`, given that group
// lo does and group hi (when hi < G) does not. Two probes per round, issued together, so a
// round trip narrows the interval threefold; with no bound yet (hi == G) the bracket is
// found first by probing at lo + 4, 16, then 64, 256, ... (two probes per round measured
// faster than one, four or eight at 1M groups).
template <int D>
__device__ uint32_t run_end(const uint64_t *gkey, uint32_t G, uint64_t key, uint32_t lo,
                            uint32_t hi, int level, int L) {
    constexpr int P = 2;
    if (hi >= G) {
        hi = G;
        for (uint32_t shift = 2; hi == G && lo + 1 < G; shift += 4) {
            uint32_t pr[P];
#pragma unroll
            for (int u = 0; u < P; u++) {
                const uint32_t sh = shift + 2 * u < 62 ? shift + 2 * u : 62;
                const uint64_t at = (uint64_t)lo + (1ull << sh);
                pr[u] = at < G ? (uint32_t)at : G - 1;
            }
            bool m[P];
#pragma unroll
            for (int u = 0; u < P; u++) m[u] = key_same<D>(key, gkey[pr[u]], level, L);
#pragma unroll
            for (int u = 0; u < P; u++) {
                if (m[u]) lo = pr[u] > lo ? pr[u] : lo;
                else hi = pr[u] < hi ? pr[u] : hi;
            }
        }
    }
    while (hi - lo > 1) {
        const uint64_t span = hi - lo;
        uint32_t pr[P];
#pragma unroll
        for (int u = 0; u < P; u++) pr[u] = lo + (uint32_t)(span * (uint64_t)(u + 1) / (P + 1));
        bool m[P];
#pragma unroll
        for (int u = 0; u < P; u++) m[u] = key_same<D>(key, gkey[pr[u]], level, L);
        uint32_t nlo = lo, nhi = hi;
#pragma unroll
        for (int u = 0; u < P; u++) {
            if (m[u]) nlo = pr[u] > nlo ? pr[u] : nlo;
            else nhi = pr[u] < nhi ? pr[u] : nhi;
        }
        lo = nlo;
        hi = nhi;
    }
    return lo;
}

template <typename T, int D>
__device__ __forceinline__ void body_final(const FaDev &d, uint32_t k, const double *s) {
    WalkCell<T, D> *wc = (WalkCell<T, D> *)d.wc;
    double *r = d.raw + (size_t)k * (1 + D);
#pragma unroll
    for (int a = 0; a <= D; a++) r[a] = s[a];
    // the record's head (the lo centre in FP32, the centre, the mass) as one store
    constexpr bool LO = !std::is_same<T, double>::value;
    struct __attribute__((packed, aligned(4))) Head {
        T lo[LO ? D : 1];
        T c[D];
        T mass;
    };
    Head h;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const double c = s[1 + a] / s[0];
        h.c[a] = (T)c;
        if constexpr (LO) h.lo[a] = (T)(c - (double)(T)c);
    }
    h.mass = (T)s[0];
    if constexpr (LO) {
        static_assert(sizeof(Head) == sizeof(T) * (2 * D + 1), "head layout");
        *(Head *)&wc[k] = h;
    } else {
        struct __attribute__((packed, aligned(4))) Head64 {
            T c[D];
            T mass;
        };
        Head64 h64;
#pragma unroll
        for (int a = 0; a < D; a++) h64.c[a] = h.c[a];
        h64.mass = h.mass;
        *(Head64 *)&wc[k] = h64;
    }
}

template <typename T, int D>
__device__ __forceinline__ void emit_cell(const FaDev &d, uint32_t k, const Box<D> &b,
                                          uint32_t skip, uint32_t start, uint32_t end,
                                          uint32_t level, double w) {
    WalkCell<T, D> *wc = (WalkCell<T, D> *)d.wc;
    // the record's structural tail (w2, skip, start, end) as one store: four separate
    // stores into 36-byte records cost a write transaction each
    struct __attribute__((packed, aligned(4))) Tail {
        T w2;
        uint32_t skip, start, end;
    };
    static_assert(sizeof(Tail) == sizeof(T) + 12, "tail layout");
    const Tail tail = {open_threshold<T>(w, d.theta2), skip, start, end};
    *(Tail *)&wc[k].w2 = tail;
    d.clevel[k] = (uint8_t)level;
    if (skip == k + 1) {  // a leaf: its positions point back to it (the walk's own-cell test)
        const uint32_t e = end < d.n ? end : d.n;
        for (uint32_t q = start; q < e; q++) d.leaf[q] = k;
    }
    if (d.export_bounds) {
        double *o = d.cbounds + (size_t)k * 2 * D;
#pragma unroll
        for (int a = 0; a < D; a++) {
            o[a] = b.lo[a];
            o[D + a] = b.hi[a];
        }
    }
}

// Writes group g's cells in preorder (preorder.rs:14-99 layout: skip, start, end per cell).
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_emit(const FaDev d) {
    GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    // capacity check of the cells the scan asked for: every workgroup makes it, the first
    // one records the outcome and clears the span-cell level counters of the next pass
    const uint32_t C = sh.need_cells;
    if (blockIdx.x == 0) {
        if (C > d.cap) {
            if (threadIdx.x == 0) st->halt = 1;
        } else {
            if (threadIdx.x == 0) st->n_cells = C;
            if (threadIdx.x < 66) d.span_lcnt[threadIdx.x] = 0;
        }
    }
    if (C > d.cap) return;
    uint32_t g = blockIdx.x * GR_BLOCK + threadIdx.x;
    const uint32_t G = sh.n_groups;
    if (g >= G) return;
    const int L = (int)d.levels;
    const uint32_t nf = sh.n_finite;
    const uint64_t *gkey = d.gkey;
    // this group's and its neighbours' records, loaded together (clamped, unconditional)
    const uint32_t gp = g > 0 ? g - 1 : g, gn = g + 1 < G ? g + 1 : g;
    const uint64_t key = gkey[g], kprev = gkey[gp], knext = gkey[gn];
    const uint32_t info = d.ginfo[g], start = d.gstart[g], snext = d.gstart[gn];
    const uint32_t gbase0 = d.gbase[g], gbase1 = d.gbase[gn];
    int left = g > 0 ? key_lcp<D>(kprev, key, L) : -1;
    int right = g + 1 < G ? key_lcp<D>(key, knext, L) : -1;
    const uint32_t type = info & 0xFF;
    int level = (int)(info >> 8);
    int last_internal = type == G_SIMPLE ? level - 1 : (type == G_STALL ? level - 1 : L - 1);
    uint32_t gend = g + 1 < G ? snext : nf;
    uint32_t gskip = g + 1 < G ? gbase1 : C;
    uint32_t idx = gbase0;
    const uint32_t gb = graph_of(d, d.order[start]);
    Box<D> b = root_box<D>(d.gr[gb]);
    int top = type == G_DEEP ? L - 1 : level;
    // only the widths are needed unless the bounds are exported; on a dyadic root they
    // follow from the level alone and the descent is skipped
    const bool descend = d.export_bounds || !dyadic_root<D>(b, top);
    const double w0 = b.hi[0] - b.lo[0];
    uint32_t run_hi = G;  // groups from here on share no more levels than the last one searched
    for (int l = 0; l <= top; l++) {
        if (l > 0 && descend) box_quad<D>(b, key_slot<D>(key, l, L));
        if (l <= left) continue;
        const double w = descend ? b.hi[0] - b.lo[0] : ldexp(w0, -l);
        uint32_t skip;
        if (l <= last_internal) {
            uint32_t end = gend;
            skip = gskip;
            if (l <= right) {
                // runs nest: this level's run ends no later than the last level's did
                const uint32_t h = run_end<D>(gkey, G, key, g, run_hi, l, L);
                run_hi = h + 1;
                skip = h + 1 < G ? d.gbase[h + 1] : C;
                end = h + 1 < G ? d.gstart[h + 1] : nf;
            }
            emit_cell<T, D>(d, idx, b, skip, start, end, (uint32_t)l, w);
        } else {
            skip = idx + 1;
            emit_cell<T, D>(d, idx, b, skip, start, gend, (uint32_t)l, w);
        }
        if (l == 0) {  // the graph's root: its tree is cells [idx, skip)
            d.gr[gb].cell_lo = idx;
            d.gr[gb].cell_hi = skip;
        }
        idx++;
    }
    if (type == G_DEEP) {
        uint32_t lc = d.glocal[g];
        uint32_t lcount = (g + 1 < G ? d.gbase[g + 1] : C) - idx;
        uint32_t base = idx;
        for (uint32_t c = 0; c < lcount; c++) {
            const int32_t *ik = d.lc_int + (size_t)(lc + c) * 4;
            const double *bk = d.lc_dbl + (size_t)(lc + c) * 2 * D;
            Box<D> lb;
#pragma unroll
            for (int a = 0; a < D; a++) {
                lb.lo[a] = bk[a];
                lb.hi[a] = bk[D + a];
            }
            emit_cell<T, D>(d, base + c, lb, base + (uint32_t)ik[0], (uint32_t)ik[1],
                            (uint32_t)ik[2], 255u, lb.hi[0] - lb.lo[0]);
            body_final<T, D>(d, base + c, d.lc_sum + (size_t)(lc + c) * (1 + D));
        }
    }
}

// ------------------------------------------------------------------ centre of mass


// barnes_hut.rs:62-110: a leaf sums its chain, an internal cell its children's
// unnormalised sums in slot order (k+1, then each child's skip). Sums are FP64 in every
// precision, and each cell's sum is formed exactly as the reference forms it, so the
// order in which cells are visited moves no bit.
template <typename T, int D>
__device__ __forceinline__ void com_cell(const FaDev &d, uint32_t k) {
    const WalkCell<T, D> *wc = (const WalkCell<T, D> *)d.wc;
    const uint32_t skip = wc[k].skip;
    double s[1 + D];
#pragma unroll
    for (int a = 0; a <= D; a++) s[a] = 0.0;
    if (skip == k + 1) {
        const uint32_t qend = wc[k].end < d.n ? wc[k].end : d.n;
        for (uint32_t q = wc[k].start; q < qend; q++) {
            double m = (double)((const T *)d.sm)[q];
            s[0] += m;
#pragma unroll
            for (int a = 0; a < D; a++) s[1 + a] += m * (double)((const T *)d.sp[a])[q];
        }
    } else {
        uint32_t c = k + 1;
        // at most 2^D children, every skip points forward: bounded on any arena
        for (int q = 0; q < (1 << D) && c < skip; q++) {
            const double *r = d.raw + (size_t)c * (1 + D);
#pragma unroll
            for (int a = 0; a <= D; a++) s[a] += r[a];
            uint32_t nc = wc[c].skip;
            c = nc > c ? nc : skip;
        }
    }
    body_final<T, D>(d, k, s);
}

constexpr uint32_t COM_TILE = 4096;
// spanning cells are ancestors of the tile's end: at most one per level
constexpr uint32_t SPAN_SLOTS = 32;

// One workgroup per tile of COM_TILE consecutive cells, its sums held in LDS. Leaves (and
// cells the deep builder already summed) are read in one parallel pass; then every cell
// whose subtree [k, skip) stays inside the tile is summed from its children's LDS sums,
// deepest level first, a workgroup barrier between levels. A cell whose subtree leaves the
// tile is an ancestor of the tile's last cell, at most one per level; it goes to the
// tile's own span slots for k_com_span.
constexpr uint32_t COM_THREADS = 1024;

template <typename T, int D>
__global__ __launch_bounds__(COM_THREADS) void k_com_tiles(const FaDev d) {
    const GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    __shared__ double sums[COM_TILE][1 + D];
    __shared__ uint32_t skips[COM_TILE];
    __shared__ uint16_t bylevel[COM_TILE];  // the tile's internal cells, bucketed by level
    __shared__ uint32_t lcnt[32], loff[33];
    __shared__ uint32_t nspan;
    const uint32_t C = sh.n_cells;
    const uint32_t t0 = blockIdx.x * COM_TILE;
    if (t0 >= C) return;
    const uint32_t t1 = t0 + COM_TILE < C ? t0 + COM_TILE : C;
    const WalkCell<T, D> *wc = (const WalkCell<T, D> *)d.wc;
    if (threadIdx.x < 32) lcnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) nspan = 0;
    __syncthreads();
    uint32_t *slots = d.span + (size_t)blockIdx.x * SPAN_SLOTS;
    constexpr int PER = COM_TILE / COM_THREADS;
    enum : uint8_t { K_NONE, K_DONE, K_LEAF, K_INNER };
    // independent loads first (level, skip, leaf range of every cell this thread owns), so
    // their latencies overlap
    uint32_t lvl[PER], skp[PER], qs[PER], qe[PER];
    uint8_t kind[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t kk = t0 + threadIdx.x + j * COM_THREADS;
        lvl[j] = kk < t1 ? d.clevel[kk] : 254u;
        skp[j] = kk < t1 ? wc[kk].skip : 0u;
        qs[j] = kk < t1 ? wc[kk].start : 0u;
        qe[j] = kk < t1 ? wc[kk].end : 0u;
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t kk = t0 + threadIdx.x + j * COM_THREADS;
        kind[j] = K_NONE;
        if (kk >= t1) continue;
        const uint32_t i = kk - t0;
        const uint32_t l = lvl[j];
        const uint32_t skip = skp[j];
        skips[i] = skip;
        if (l == 255) {  // summed by the deep builder
            const double *r = d.raw + (size_t)kk * (1 + D);
#pragma unroll
            for (int a = 0; a <= D; a++) sums[i][a] = r[a];
            kind[j] = K_DONE;
        } else if (skip > t1) {  // crosses the tile: a span cell
            uint32_t at = atomicAdd(&nspan, 1u);  // LDS
            if (at < SPAN_SLOTS) {
                slots[at] = kk;
                d.span_lev[(size_t)blockIdx.x * SPAN_SLOTS + at] = (uint8_t)l;
                atomicAdd(&d.span_lcnt[l], 1u);
            }
            kind[j] = K_DONE;
        } else if (skip == kk + 1) {
            double s[1 + D];
#pragma unroll
            for (int a = 0; a <= D; a++) s[a] = 0.0;
            const uint32_t qend = qe[j] < d.n ? qe[j] : d.n;
            for (uint32_t q = qs[j]; q < qend; q++) {
                double m = (double)((const T *)d.sm)[q];
                s[0] += m;
#pragma unroll
                for (int a = 0; a < D; a++) s[1 + a] += m * (double)((const T *)d.sp[a])[q];
            }
#pragma unroll
            for (int a = 0; a <= D; a++) sums[i][a] = s[a];
            kind[j] = K_LEAF;
        } else if (l < 32) {
            atomicAdd(&lcnt[l], 1u);
            kind[j] = K_INNER;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d.span_n[blockIdx.x] = nspan < SPAN_SLOTS ? nspan : SPAN_SLOTS;
        uint32_t acc = 0;
        for (int l = 0; l < 32; l++) {
            loff[l] = acc;
            acc += lcnt[l];
        }
        loff[32] = acc;
    }
    __syncthreads();
    if (threadIdx.x < 32) lcnt[threadIdx.x] = loff[threadIdx.x];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; j++)
        if (kind[j] == K_INNER)
            bylevel[atomicAdd(&lcnt[lvl[j]], 1u)] = (uint16_t)(threadIdx.x + j * COM_THREADS);
    __syncthreads();
    // deepest level first: a cell sums its children in slot order (k+1, then each skip)
    for (int l = 31; l >= 0; l--) {
        const uint32_t a = loff[l], b = loff[l + 1];
        if (a == b) continue;  // block-uniform
        for (uint32_t x = a + threadIdx.x; x < b; x += COM_THREADS) {
            const uint32_t i = bylevel[x], k = t0 + i;
            double s[1 + D];
#pragma unroll
            for (int y = 0; y <= D; y++) s[y] = 0.0;
            const uint32_t skip = skips[i];
            uint32_t c = k + 1;
            for (int q = 0; q < (1 << D) && c < skip; q++) {
                const uint32_t ci = c - t0;
#pragma unroll
                for (int y = 0; y <= D; y++) s[y] += sums[ci][y];
                const uint32_t nc = skips[ci];
                c = nc > c ? nc : skip;
            }
#pragma unroll
            for (int y = 0; y <= D; y++) sums[i][y] = s[y];
        }
        __syncthreads();
    }
    // the records of the tile's leaves and internal cells, all at once
#pragma unroll
    for (int j = 0; j < PER; j++) {
        if (kind[j] != K_LEAF && kind[j] != K_INNER) continue;
        const uint32_t i = threadIdx.x + j * COM_THREADS;
        double s[1 + D];
#pragma unroll
        for (int y = 0; y <= D; y++) s[y] = sums[i][y];
        body_final<T, D>(d, t0 + i, s);
    }
}

// The span cells the LDS path of k_com_span holds (2D), and child references: a child that
// is itself a span cell is flagged (named by its cell, then by its slot).
constexpr uint32_t SPAN_LDS = 3072;
constexpr int SPAN_PER = (int)(SPAN_LDS / 1024);
constexpr uint32_t SPAN_REF = 0x80000000u, NO_KID = 0xFFFFFFFFu;

// Level offsets of the span buckets from the per-level counts k_com_tiles took.
__device__ __forceinline__ void span_offsets(const FaDev &d, uint32_t *loff) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int l = 0; l < 33; l++) {
            loff[l] = acc;
            acc += d.span_lcnt[l];
        }
        loff[33] = acc;
    }
    __syncthreads();
}

// One thread per (tile, span slot): puts the span cell in its level's bucket and gathers
// its children in slot order (k + 1, then each child's skip): a child that also crosses a
// tile is referenced, any other child is final and its sums are copied. Spread over many
// workgroups, so the chains of dependent loads of all span cells run side by side.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_span_prep(const FaDev d) {
    const GrState *st = d.st;
    const StateHead sh = state_head(d);
    if (halted(d, sh)) return;
    __shared__ uint32_t loff[34];
    span_offsets(d, loff);
    const uint32_t C = sh.n_cells;
    const uint32_t ntiles = (C + COM_TILE - 1) / COM_TILE;
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= ntiles * SPAN_SLOTS || (i % SPAN_SLOTS) >= d.span_n[i / SPAN_SLOTS]) return;
    constexpr int KIDS = 1 << D;
    const WalkCell<T, D> *wc = (const WalkCell<T, D> *)d.wc;
    const uint32_t k = d.span[i], l = d.span_lev[i];
    const uint32_t at = loff[l] + atomicAdd(&d.span_lcnt[33 + l], 1u);
    d.span_cells[at] = k;
    d.span_slot[k] = at;
    const uint32_t skip = wc[k].skip;
    uint32_t c = k + 1;
#pragma unroll
    for (int x = 0; x < KIDS; x++) {
        uint32_t ref = NO_KID;
        double v[1 + D];
#pragma unroll
        for (int y = 0; y <= D; y++) v[y] = 0.0;
        if (c < skip) {
            const uint32_t cs = wc[c].skip;
            const uint32_t t0c = (c / COM_TILE) * COM_TILE;
            const uint32_t tend = t0c + COM_TILE < C ? t0c + COM_TILE : C;
            // the test k_com_tiles sorts a cell out by
            if (d.clevel[c] != 255 && cs > tend) {
                ref = SPAN_REF | c;
            } else {
                ref = c;
                const double *r = d.raw + (size_t)c * (1 + D);
#pragma unroll
                for (int y = 0; y <= D; y++) v[y] = r[y];
            }
            c = cs > c ? cs : skip;
        }
        d.span_kids[(size_t)at * KIDS + x] = ref;
        // values lane-major ([child][component][span cell]), read back coalesced
#pragma unroll
        for (int y = 0; y <= D; y++)
            d.span_kval[(size_t)(x * (1 + D) + y) * d.span_cap + at] = v[y];
    }
}

// One thread per bucketed span cell: a child that is itself a span cell is named by its
// slot from here on. The references are written lane-major ([child][span cell]) into the
// second half of d.span_kids, where the single-workgroup pass reads them coalesced.
template <int D>
__global__ __launch_bounds__(GR_BLOCK) void k_span_link(const FaDev d) {
    const GrState *st = d.st;
    if (halted(d)) return;
    __shared__ uint32_t loff[34];
    span_offsets(d, loff);
    const uint32_t total = loff[33];
    const uint32_t q = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (q >= total) return;
    constexpr int KIDS = 1 << D;
    uint32_t *link = d.span_kids + (size_t)d.span_cap * KIDS;
#pragma unroll
    for (int x = 0; x < KIDS; x++) {
        uint32_t r = d.span_kids[(size_t)q * KIDS + x];
        if (r != NO_KID && (r & SPAN_REF)) {
            const uint32_t c = r & ~SPAN_REF;
            const uint32_t slot = d.span_slot[c];
            if (slot < total && d.span_cells[slot] == c) {
                r = SPAN_REF | slot;
            } else {  // inconsistent span lists: never expected
                d.st->halt = 1;
                r = NO_KID;
            }
        }
        link[(size_t)x * d.span_cap + q] = r;
    }
}

// The span cells (the ancestors of tile ends), finished by one workgroup level by level,
// deepest first, from the buckets and children k_span_prep laid out. Every child sum is
// added in slot order, exactly as k_com_tiles adds them. In 2D, when they fit, the span
// cells' sums stay in LDS and a level costs only LDS traffic and a barrier; otherwise a
// span child's sum is read back from d.raw, where the level before wrote it.
template <typename T, int D>
__global__ __launch_bounds__(1024) void k_com_span(const FaDev d, uint32_t levelwise) {
    const GrState *st = d.st;
    if (halted(d)) return;
    __shared__ uint32_t loff[34];
    constexpr int KIDS = 1 << D;
    __shared__ double ssum[D == 2 ? SPAN_LDS : 1][1 + D];
    __shared__ uint32_t scells[D == 2 ? SPAN_LDS : 1];
    span_offsets(d, loff);
    const uint32_t total = loff[33];
    const uint32_t *cells = d.span_cells;
    const uint32_t *link = d.span_kids + (size_t)d.span_cap * KIDS;
    if (D == 2 && total <= SPAN_LDS) {
        double pre[SPAN_PER][KIDS][1 + D];
        uint32_t ref[SPAN_PER][KIDS];
#pragma unroll
        for (int j = 0; j < SPAN_PER; j++) {
            const uint32_t q = threadIdx.x + (uint32_t)j * 1024u;
            const bool mine = q < total;
            if (mine) scells[q] = cells[q];
#pragma unroll
            for (int x = 0; x < KIDS; x++) {
                ref[j][x] = mine ? link[(size_t)x * d.span_cap + q] : NO_KID;
#pragma unroll
                for (int y = 0; y <= D; y++)
                    pre[j][x][y] = mine ? d.span_kval[(size_t)(x * (1 + D) + y) * d.span_cap + q]
                                        : 0.0;
            }
        }
        __syncthreads();
        for (int l = 31; l >= 0; l--) {
            const uint32_t a = loff[l], b = loff[l + 1];
            if (a == b) continue;  // block-uniform
#pragma unroll
            for (int j = 0; j < SPAN_PER; j++) {
                const uint32_t q = threadIdx.x + (uint32_t)j * 1024u;
                if (q < a || q >= b) continue;
                double s[1 + D];
#pragma unroll
                for (int y = 0; y <= D; y++) s[y] = 0.0;
#pragma unroll
                for (int x = 0; x < KIDS; x++) {
                    const uint32_t r = ref[j][x];
                    if (r == NO_KID) continue;
                    const bool span = (r & SPAN_REF) != 0;
                    const uint32_t sl = span ? (r & ~SPAN_REF) : 0u;
#pragma unroll
                    for (int y = 0; y <= D; y++) s[y] += span ? ssum[sl][y] : pre[j][x][y];
                }
#pragma unroll
                for (int y = 0; y <= D; y++) ssum[q][y] = s[y];
            }
            __syncthreads();
        }
        // the records of every span cell, all in parallel once the sums are done
#pragma unroll
        for (int j = 0; j < SPAN_PER; j++) {
            const uint32_t q = threadIdx.x + (uint32_t)j * 1024u;
            if (q >= total) continue;
            double s[1 + D];
#pragma unroll
            for (int y = 0; y <= D; y++) s[y] = ssum[q][y];
            body_final<T, D>(d, scells[q], s);
        }
        return;
    }
    if (levelwise) return;  // k_com_span_level takes the cells that overflow LDS
    for (int l = 31; l >= 0; l--) {
        const uint32_t a = loff[l], b = loff[l + 1];
        if (a == b) continue;  // block-uniform
        for (uint32_t q = a + threadIdx.x; q < b; q += blockDim.x) {
            double s[1 + D];
#pragma unroll
            for (int y = 0; y <= D; y++) s[y] = 0.0;
#pragma unroll
            for (int x = 0; x < KIDS; x++) {
                const uint32_t r = link[(size_t)x * d.span_cap + q];
                if (r == NO_KID) continue;
                if (r & SPAN_REF) {
                    const double *v = d.raw + (size_t)cells[r & ~SPAN_REF] * (1 + D);
#pragma unroll
                    for (int y = 0; y <= D; y++) s[y] += v[y];
                } else {
#pragma unroll
                    for (int y = 0; y <= D; y++)
                        s[y] += d.span_kval[(size_t)(x * (1 + D) + y) * d.span_cap + q];
                }
            }
            body_final<T, D>(d, cells[q], s);
        }
        __syncthreads();
    }
}

// One level of the span pass spread over the GPU, for sessions whose span cells overflow
// k_com_span's LDS path (batches): every span cell of level l sums its children, which are
// final cells or span cells of deeper levels finished by the previous launch. Launched once
// per level, deepest first.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_com_span_level(const FaDev d, int l) {
    if (halted(d)) return;
    __shared__ uint32_t loff[34];
    span_offsets(d, loff);
    if (D == 2 && loff[33] <= SPAN_LDS) return;  // k_com_span's LDS path did them all
    constexpr int KIDS = 1 << D;
    const uint32_t a = loff[l], b = loff[l + 1];
    const uint32_t *cells = d.span_cells;
    const uint32_t *link = d.span_kids + (size_t)d.span_cap * KIDS;
    for (uint32_t q = a + blockIdx.x * GR_BLOCK + threadIdx.x; q < b;
         q += gridDim.x * GR_BLOCK) {
        double s[1 + D];
#pragma unroll
        for (int y = 0; y <= D; y++) s[y] = 0.0;
#pragma unroll
        for (int x = 0; x < KIDS; x++) {
            const uint32_t r = link[(size_t)x * d.span_cap + q];
            if (r == NO_KID) continue;
            if (r & SPAN_REF) {
                const double *v = d.raw + (size_t)cells[r & ~SPAN_REF] * (1 + D);
#pragma unroll
                for (int y = 0; y <= D; y++) s[y] += v[y];
            } else {
#pragma unroll
                for (int y = 0; y <= D; y++)
                    s[y] += d.span_kval[(size_t)(x * (1 + D) + y) * d.span_cap + q];
            }
        }
        body_final<T, D>(d, cells[q], s);
    }
}

// ------------------------------------------------------------------ walk

// barnes_hut.rs:112-184 as one wave-uniform preorder scan per 64 consecutive tree
// positions. The wave visits the smallest of its lanes' next cells (each lane's `k` of the
// reference), so each cell is loaded once per wave and each lane still sees its own cells
// in preorder, accepting or opening them exactly as the reference does. A lane that opened
// the cell acts on k + 1; a lane that accepted it waits until the scan reaches its skip.
// BUF: the cell arena spans under 2 GiB, so records are read as buffer loads: a scalar byte
// offset per record and the lane's dword offset, fixed for the whole scan.
// ONE: a single-graph session, where every lane's tree ends at the scan's end.
// COUNT: the measuring variant, which also counts per wave its scan steps, the lanes acting
// on each step, the cells accepted and the leaf pairs taken (d.wstat, four u64 totals).
// WB: threads per workgroup (64, 128 or 256), the schedule's choice; with d.walk_xcd the
// consecutive workgroups of the grid share an XCD. Neither moves a result bit.
template <typename T, int D, bool KAHAN, bool BUF, bool ONE, bool COUNT, int WB>
__global__ __launch_bounds__(WB) void k_walk(const FaDev d) {
    const GrState *st = d.st;
    if (halted(d)) return;
    const uint32_t n = d.n;
    // g consecutive tree positions per wave (64, or fewer on a small graph: more waves)
    const uint32_t g = d.walk_group, per = (WB / GR_WAVE) * g;
    const uint32_t lb = d.walk_xcd ? gr_xcd_block(blockIdx.x, (n + per - 1) / per) : blockIdx.x;
    const uint32_t wl = threadIdx.x % GR_WAVE;
    const uint32_t p = lb * per + (threadIdx.x / GR_WAVE) * g + wl;
    const bool live = wl < g && p < n;
    if (__ballot(live) == 0) return;
    const uint32_t pc = live ? p : n - 1;
    const WalkCell<T, D> *__restrict__ wc = (const WalkCell<T, D> *)d.wc;
    const T *__restrict__ sm = (const T *)d.sm;
    const uint32_t *__restrict__ order = d.order;
    const T *sp[3] = {(const T *)d.sp[0], (const T *)d.sp[1],
                      (const T *)(D == 3 ? d.sp[2] : d.sp[0])};
    T xi[D];
#pragma unroll
    for (int a = 0; a < D; a++) xi[a] = sp[a][pc];
    const T mi = sm[pc];
    const uint32_t i = order[pc];
    const T theta2 = (T)d.theta2, kr = (T)d.scaling_ratio;
    const T mikr = mi * kr;
    // each lane walks its own graph's tree, cells [lo, hi) of the arena
    const GrGraph &gg = d.gr[graph_of(d, i)];
    const uint32_t seed = d.seed, tick = gg.iter_no;
    const uint32_t lo = gg.cell_lo, hi = gg.cell_hi;
    const uint32_t ncells = st->n_cells;
    const uint32_t own = (live && p < st->n_finite) ? d.leaf[p] : 0xFFFFFFFFu;
    T f[D], fc[D];
#pragma unroll
    for (int a = 0; a < D; a++) f[a] = fc[a] = (T)0;
    // A lane acts on the scan's cell k when it opened the cell before (the scan then moves
    // to k + 1, its next cell), or when k is the cell it waits for: its tree's root, or the
    // skip of the last cell it accepted. Waits only move forward with the scan.
    uint32_t wait = (live && lo < hi) ? lo : 0xFFFFFFFFu;
    uint64_t carried = 0;
    uint32_t k = wait;
    // the scan ends at the last cell of the wave's last graph: later trees are others'
    uint32_t end = live ? hi : 0u;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t o = __shfl_xor(k, m, GR_WAVE);
        k = o < k ? o : k;
        const uint32_t e = __shfl_xor(end, m, GR_WAVE);
        end = e > end ? e : end;
    }
    end = gr_uniform(end < ncells ? end : ncells);
    k = gr_uniform(k);
    // The cell record travels as W dwords, lane l holding dword l: vector loads complete in
    // order, so both possible successors (k + 1 and skip) can be fetched while this cell is
    // worked on, and the one taken is waited for alone.
    constexpr uint32_t W = sizeof(WalkCell<T, D>) / 4;
    const uint32_t lane = gr_lane();
    const uint32_t dw = lane < W ? lane : 0u;  // lanes past the record re-read dword 0
    const uint32_t *__restrict__ wcw = (const uint32_t *)d.wc;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)wcw, (short)0, BUF ? (int)((d.cap + 1) * W * 4) : 0, 0x00020000);
    auto record = [&](uint32_t c) -> uint32_t {
        if constexpr (BUF) {
            return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rs, (int)(dw * 4),
                                                                 (int)(c * (W * 4)), 0);
        } else {
            return wcw[(size_t)c * W + dw];
        }
    };
    uint32_t cur = k < end ? record(k) : 0u;
    uint64_t c_steps = 0, c_act = 0, c_acc = 0, c_pairs = 0;
    while (k < end) {
        WalkCell<T, D> cell;
        {
            uint32_t w[W];
#pragma unroll
            for (uint32_t q = 0; q < W; q++)
                w[q] = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)q);
            __builtin_memcpy(&cell, w, sizeof(cell));
        }
        const uint32_t k1 = k + 1;
        // skip always points past k; the max keeps the scan finite on any arena, the min
        // keeps the read inside it (the arena holds one spare record past its capacity, so
        // cell `end` may be read; a successor at or past end is never used)
        const uint32_t ks = min(max(cell.skip, k1), end);
        const uint32_t nx1 = record(k1);
        const uint32_t nxs = record(ks);
        // every lane evaluates the test (no divergent branch); only acting lanes keep it.
        // The tests are kept as wave masks (compare results), combined on the scalar unit.
        // the acting lanes, inside their graph's tree (a lane's tree ends at its hi)
        const uint64_t acting =
            ONE ? carried | wave_cmp_u32<GR_CMP_EQ>(wait, k)
                : (carried | wave_cmp_u32<GR_CMP_EQ>(wait, k)) & wave_cmp_u32<GR_CMP_ULT>(k, hi);
        // the cell holds the lane's own point when its subtree [k, skip) holds that leaf
        // (the same as start <= p < end, as a subtree's positions are its leaves')
        const uint64_t holding = wave_cmp_u32<GR_CMP_ULE>(k, own) &
                                 wave_cmp_u32<GR_CMP_ULT>(own, cell.skip);
        T dx[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            dx[a] = xi[a] - cell.c[a];
            if constexpr (!std::is_same<T, double>::value) dx[a] = dx[a] - cell.clo[a];
        }
        T d2 = mul_add(dx[0], dx[0], dx[1] * dx[1]);
        if (D == 3) d2 = mul_add(dx[D - 1], dx[D - 1], d2);
        // barnes_hut.rs's d2 > 0 is implied: w2 >= 0, so w2 < theta2 * d2 needs d2 > 0
        const uint64_t accepted = acting & ~holding & wave_lt(cell.w2, open_rhs(d2, theta2));
        const bool accept = __builtin_amdgcn_inverse_ballot_w64(accepted);
        const bool act = __builtin_amdgcn_inverse_ballot_w64(acting);
        if constexpr (COUNT) {
            c_steps += 1;
            c_act += (uint64_t)__builtin_popcountll(acting);
            c_acc += (uint64_t)__builtin_popcountll(accepted);
        }
        if (accepted != 0) {
            // the mass is read from the record only here, where a lane takes the force
            constexpr uint32_t MD = WalkCell<T, D>::MASS_DWORD;
            T cmass;
            if constexpr (std::is_same<T, double>::value) {
                const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)MD);
                const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)MD + 1);
                cmass = __builtin_bit_cast(double, lo | (hi << 32));
            } else {
                cmass = __builtin_bit_cast(float, __builtin_amdgcn_readlane((int)cur, (int)MD));
            }
            // only the accepting lanes run the force (an exec mask, not per-lane selects)
            if (accept) {
                // FP64 forms the reference's mi * mc / d2 * kr; FP32 takes (mi * kr) * mc / d2
                const T ff = std::is_same<T, double>::value ? pair_force<T>(mi * cmass, d2, kr)
                                                            : mikr * cmass * fast_rcp(d2);
#pragma unroll
                for (int a = 0; a < D; a++) {
                    if constexpr (KAHAN) {
                        acc_add<T, KAHAN>(f[a], fc[a], dx[a] * ff);
                    } else {
                        f[a] = mul_add(dx[a], ff, f[a]);
                    }
                }
            }
        }
        const bool open = act & !accept;
        // the lanes that opened k act on k + 1 next; an accepting lane acts at skip. When no
        // lane opened k the scan goes to skip itself and the accepting lanes are carried
        // there; only when it descends do they wait for it (a stale wait lies behind the scan
        // and never matches again)
        carried = acting & ~accepted;
        const bool down = carried != 0;
        if (!down) {
            // on to skip, the accepting lanes with it
            carried = accepted;
            k = ks;
            cur = nxs;
            continue;
        }
        wait = accept ? cell.skip : wait;
        if (cell.skip == k1) {
            const bool pairs = open;
            T lf[D], lc[D];
#pragma unroll
            for (int a = 0; a < D; a++) lf[a] = lc[a] = (T)0;
            // the leaf's range is read from the record only here
            constexpr uint32_t SD = WalkCell<T, D>::START_DWORD;
            const uint32_t cstart = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)SD);
            const uint32_t cend = (uint32_t)__builtin_amdgcn_readlane((int)cur, (int)SD + 1);
            const uint32_t qend = cend < n ? cend : n;
            for (uint32_t q = cstart; q < qend; q++) {
                const uint32_t j = order[q];
                T xj[D];
#pragma unroll
                for (int a = 0; a < D; a++) xj[a] = sp[a][q];
                const T mj = sm[q];
                if constexpr (COUNT) c_pairs += (uint64_t)__builtin_popcountll(__ballot(pairs && j != i));
                if (pairs && j != i) {
                    // repel_delta from the lower index, negated for the higher one
                    // (barnes_hut.rs:170-180), the coincidence jiggle included.
                    T dd[D];
                    bool zero = true;
#pragma unroll
                    for (int a = 0; a < D; a++) {
                        dd[a] = i < j ? xi[a] - xj[a] : -(xj[a] - xi[a]);
                        if (dd[a] != (T)0) zero = false;
                    }
                    if (zero) {
#pragma unroll
                        for (int a = 0; a < D; a++) {
                            T jg = (T)d_jiggle(seed, tick, (uint32_t)a, i, j);
                            dd[a] = i < j ? jg : -jg;
                        }
                    }
                    T d2 = dd[0] * dd[0] + dd[1] * dd[1];
                    if (D == 3) d2 = d2 + dd[D - 1] * dd[D - 1];
                    const T ff = pair_force<T>(mi * mj, d2, kr);
#pragma unroll
                    for (int a = 0; a < D; a++) acc_add<T, KAHAN>(lf[a], lc[a], dd[a] * ff);
                }
            }
            if (pairs) {
#pragma unroll
                for (int a = 0; a < D; a++) acc_add<T, KAHAN>(f[a], fc[a], lf[a]);
            }
        }
        // down to k + 1: a lane opened k
        k = k1;
        cur = nx1;
    }
    if constexpr (COUNT) {
        if (gr_lane() == 0) {
            atomicAdd((unsigned long long *)&d.wstat[0], (unsigned long long)c_steps);
            atomicAdd((unsigned long long *)&d.wstat[1], (unsigned long long)c_act);
            atomicAdd((unsigned long long *)&d.wstat[2], (unsigned long long)c_acc);
            atomicAdd((unsigned long long *)&d.wstat[3], (unsigned long long)c_pairs);
        }
    }
    if (live) {
#pragma unroll
        for (int a = 0; a < D; a++) ((T *)d.rep[a])[i] = f[a];
    }
}

// ------------------------------------------------------------------ forces and step

// state.rs:152-162 as a gather: node v sums (p[nbr] - p[v]) over its CSR row, which lists
// its simple edges in ascending id, the order the reference's scatter reaches v in.
template <typename T, int D, bool KAHAN>
__device__ __forceinline__ void attraction_of(const FaDev &d, uint32_t v, const T *pv, T *acc) {
    T c[D];
#pragma unroll
    for (int a = 0; a < D; a++) acc[a] = c[a] = (T)0;
    const uint32_t e0 = d.row_off[v], e1 = d.row_off[v + 1];
    const T *w = (const T *)d.w;
    const T *pi = (const T *)d.pi;
    for (uint32_t e = e0; e < e1; e++) {
        uint32_t j = d.nbr[e];
        T pj[D];
#pragma unroll
        for (int a = 0; a < D; a++) pj[a] = pi[(size_t)j * D + a];
#pragma unroll
        for (int a = 0; a < D; a++) {
            T dd = pj[a] - pv[a];
            if (w) dd = dd * w[e];
            acc_add<T, KAHAN>(acc[a], c[a], dd);
        }
    }
}

template <typename T, int D, bool KAHAN>
__global__ __launch_bounds__(GR_BLOCK) void k_attract_only(const FaDev d) {
    const GrState *st = d.st;
    if (halted(d)) return;
    uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (v >= d.n) return;
    T pv[D], acc[D];
#pragma unroll
    for (int a = 0; a < D; a++) pv[a] = ((const T *)d.pos[a])[v];
    attraction_of<T, D, KAHAN>(d, v, pv, acc);
#pragma unroll
    for (int a = 0; a < D; a++) ((T *)d.u[a])[v] = acc[a];
}

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_zero_u(const FaDev d) {
    uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (v >= d.n) return;
#pragma unroll
    for (int a = 0; a < D; a++) ((T *)d.u[a])[v] = (T)0;
}

// state.rs:169-203 as a gather: node v adds every pair term in ascending partner order, the
// order the reference's i<j double loop reaches v in, starting from v's attraction.
template <typename T, int D, bool KAHAN>
__global__ __launch_bounds__(GR_BLOCK) void k_dense(const FaDev d, uint32_t with_attraction) {
    const GrState *st = d.st;
    if (halted(d)) return;
    __shared__ T tp[D][GR_BLOCK];
    __shared__ T tm[GR_BLOCK];
    const uint32_t n = d.n;
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    const bool live = v < n;
    const uint32_t vc = live ? v : 0;
    T pv[D], acc[D], c[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        pv[a] = ((const T *)d.pos[a])[vc];
        acc[a] = with_attraction ? ((const T *)d.u[a])[vc] : (T)0;
        c[a] = (T)0;
    }
    const T mv = ((const T *)d.mass)[vc];
    const T kr = (T)d.scaling_ratio;
    const uint32_t gv = graph_of(d, vc);
    const uint32_t seed = d.seed, tick = d.gr[gv].iter_no;
    const uint32_t jlo = d.gnode[gv], jhi = d.gnode[gv + 1];
    // a block's nodes can span graphs: it sweeps the union of their node ranges
    __shared__ uint32_t s_lo, s_hi;
    if (threadIdx.x == 0) {
        s_lo = 0xFFFFFFFFu;
        s_hi = 0;
    }
    __syncthreads();
    if (live) {
        atomicMin(&s_lo, jlo);
        atomicMax(&s_hi, jhi);
    }
    __syncthreads();
    const uint32_t blo = s_lo, bhi = s_hi;
    for (uint32_t j0 = blo; j0 < bhi; j0 += GR_BLOCK) {
        __syncthreads();
        uint32_t jj = j0 + threadIdx.x;
        if (jj < bhi) {
#pragma unroll
            for (int a = 0; a < D; a++) tp[a][threadIdx.x] = ((const T *)d.pos[a])[jj];
            tm[threadIdx.x] = ((const T *)d.mass)[jj];
        }
        __syncthreads();
        uint32_t cnt = bhi - j0 < GR_BLOCK ? bhi - j0 : GR_BLOCK;
        if (!live) continue;
        for (uint32_t t = 0; t < cnt; t++) {
            uint32_t j = j0 + t;
            if (j == v || j < jlo || j >= jhi) continue;
            T dd[D];
            bool zero = true;
#pragma unroll
            for (int a = 0; a < D; a++) {
                dd[a] = v < j ? pv[a] - tp[a][t] : -(tp[a][t] - pv[a]);
                if (dd[a] != (T)0) zero = false;
            }
            if (zero) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    T jg = (T)d_jiggle(seed, tick, (uint32_t)a, v, j);
                    dd[a] = v < j ? jg : -jg;
                }
            }
            T d2 = (T)0;
#pragma unroll
            for (int a = 0; a < D; a++) d2 = d2 + dd[a] * dd[a];
            const T ff = mv * tm[t] / d2 * kr;
#pragma unroll
            for (int a = 0; a < D; a++) acc_add<T, KAHAN>(acc[a], c[a], dd[a] * ff);
        }
    }
    if (live) {
#pragma unroll
        for (int a = 0; a < D; a++) ((T *)d.u[a])[v] = acc[a];
    }
}

// state.rs:152-162 for a row longer than LONG_ROW: the workgroup gathers the row's
// neighbour positions (and weights) into LDS 256 at a time, and one lane sums them in
// ascending entry order, the same sequence k_finish uses for a short row, so the sum is
// bit-identical; only the gathers run in parallel instead of one dependent load per entry.
template <typename T, int D, bool KAHAN>
__global__ __launch_bounds__(GR_BLOCK) void k_long_rows(const FaDev d) {
    if (halted(d)) return;
    __shared__ T s_p[D][GR_BLOCK];
    __shared__ T s_w[GR_BLOCK];
    const uint32_t v = d.long_rows[blockIdx.x];
    const uint32_t e0 = d.row_off[v], e1 = d.row_off[v + 1];
    const T *pi = (const T *)d.pi;
    const T *wgt = (const T *)d.w;
    T pv[D], uv[D], cc[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
        pv[a] = pi[(size_t)v * D + a];
        uv[a] = cc[a] = (T)0;
    }
    for (uint32_t base = e0; base < e1; base += GR_BLOCK) {
        const uint32_t e = base + threadIdx.x;
        if (e < e1) {
            const uint32_t j = d.nbr[e];
#pragma unroll
            for (int a = 0; a < D; a++) s_p[a][threadIdx.x] = pi[(size_t)j * D + a];
            if (wgt) s_w[threadIdx.x] = wgt[e];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t cnt = e1 - base < GR_BLOCK ? e1 - base : GR_BLOCK;
            for (uint32_t t = 0; t < cnt; t++) {
#pragma unroll
                for (int a = 0; a < D; a++) {
                    T dd = s_p[a][t] - pv[a];
                    if (wgt) dd = dd * s_w[t];
                    acc_add<T, KAHAN>(uv[a], cc[a], dd);
                }
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < D; a++) ((T *)d.long_acc)[(size_t)blockIdx.x * D + a] = uv[a];
    }
}

// One row's attraction terms from the staged entries [e, lim) (entry q = e - s0), in
// ascending entry order: (p_j - p_v) [* w] added to the row's sum per axis. Four entries'
// LDS reads are issued before their adds, so the reads do not wait one by one.
template <typename T, int D, bool KAHAN, bool W>
__device__ __forceinline__ void row_sum(const T *s_pj, const T *s_w, uint32_t &e, uint32_t lim,
                                        uint32_t s0, const T (&pv)[D], T (&uv)[D],
                                        T (&cc)[D]) {
    constexpr int U = 4;
    for (; e + U <= lim; e += U) {
        T pj[U][D], ww[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint32_t q = e + u - s0;
#pragma unroll
            for (int a = 0; a < D; a++) pj[u][a] = s_pj[(size_t)q * D + a];
            if (W) ww[u] = s_w[q];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
#pragma unroll
            for (int a = 0; a < D; a++) {
                T dd = pj[u][a] - pv[a];
                if (W) dd = dd * ww[u];
                acc_add<T, KAHAN>(uv[a], cc[a], dd);
            }
        }
    }
    for (; e < lim; e++) {
        const uint32_t q = e - s0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            T dd = s_pj[(size_t)q * D + a] - pv[a];
            if (W) dd = dd * s_w[q];
            acc_add<T, KAHAN>(uv[a], cc[a], dd);
        }
    }
}

// Completes u (attraction + repulsion, then gravity, state.rs:130-148 order) and the
// swing/traction partials (schedule.rs:16-30). mode 0: repulsion from the walk (u = attr +
// rep); mode 1: u already holds attraction (+ dense pairs).
//
// Work split: block b owns the rows whose (first entry + row index) falls in its share of
// E + n, so rows and entries are balanced together; blocks are dealt so that one XCD owns
// one contiguous range of rows (graph-local neighbours then share its L2). Each round
// gathers the neighbour positions of up to fin_stage() consecutive CSR entries into LDS,
// one entry per lane with many gathers in flight (coalesced index loads, no divergence);
// a row is then summed by one lane from LDS in ascending entry order, the reference's
// order. Entries of rows longer than LONG_ROW are not staged (k_long_rows sums them).
// The row split (part_rows) is fixed by the CSR and computed once per session; each
// workgroup gets about one round's worth of rows + entries. Dynamic LDS: FIN_LDS bytes.
// waves per SIMD k_finish is built for: six fit its LDS, and the FP32 2D build is held to
// the registers that allow them (it is latency-bound, so occupancy is its speed; the Kahan
// and FP64 builds would spill)
template <typename T, int D, bool KAHAN>
constexpr int fin_waves() {
    return sizeof(T) == 4 && D == 2 && !KAHAN ? 6 : 1;
}
template <typename T, int D, bool KAHAN>
__global__ __launch_bounds__(GR_BLOCK, (fin_waves<T, D, KAHAN>())) void k_finish(
    const FaDev d, uint32_t mode, uint32_t parts) {
    // the row range and (one graph) the centroid are read along with the halt flag, not
    // after it: both reads are in bounds whatever the flag says
    const uint32_t lb = gr_xcd_block(blockIdx.x, gridDim.x);
    const uint32_t r0 = d.part_rows[lb], r1 = d.part_rows[lb + 1];
    double mean1[D];
#pragma unroll
    for (int a = 0; a < D; a++) mean1[a] = d.gidn ? 0.0 : d.gr[0].mean[a];
    if (halted(d)) return;
    const T *pi = (const T *)d.pi;
    const T *wgt = (const T *)d.w;
    const uint32_t STAGE = fin_stage(sizeof(T), D, wgt != nullptr);
    extern __shared__ __align__(16) unsigned char fin_lds[];
    T *s_pj = (T *)fin_lds;     // [STAGE][D]
    T *s_w = s_pj + STAGE * D;  // [STAGE], weighted graphs only
    __shared__ uint32_t s_lr_lo[GR_BLOCK], s_lr_hi[GR_BLOCK], s_wl[GR_BLOCK / GR_WAVE];
    __shared__ uint32_t s_g[2];
    const double grav = d.gravity;
    for (uint32_t base = r0; base < r1; base += GR_BLOCK) {
        const uint32_t rend = base + GR_BLOCK < r1 ? base + GR_BLOCK : r1;
        const uint32_t v = base + threadIdx.x;
        const bool live = v < rend;
        // the row's inputs besides its neighbours, all loaded before the attraction rounds:
        // mode 0 adds the walk's repulsion, mode 1 takes u as it stands
        const bool addrep = mode == 0 && (parts & PART_REPULSION);
        T pv[D], uv[D], cc[D], rv[D];
#pragma unroll
        for (int a = 0; a < D; a++) {
            pv[a] = live ? pi[(size_t)v * D + a] : (T)0;
            const T *other = (const T *)(mode == 0 ? d.rep[a] : d.u[a]);
            rv[a] = live && (addrep || mode != 0) ? other[v] : (T)0;
            uv[a] = cc[a] = (T)0;
        }
        const T mt = live ? ((const T *)d.mass)[v] : (T)0;
        const bool attr = mode == 0 && (parts & PART_ATTRACTION);
        const uint32_t e0 = live ? d.row_off[v] : 0;
        const uint32_t e1 = live ? d.row_off[v + 1] : 0;
        const bool longrow = e1 - e0 > LONG_ROW;  // summed by k_long_rows
        uint32_t e = longrow ? e1 : e0;
        if (attr) {
            // the chunk's entries [g0, g1) and its long rows, ascending, whose entries are
            // left out of the staging; a graph without long rows reads the range directly
            // (alongside the rows' own loads) and needs no exchange between lanes
            uint32_t nl = 0, g0, g1;
            if (d.nlong == 0) {
                g0 = d.row_off[base];
                g1 = d.row_off[rend];
            } else {
                const uint64_t lm = gr_ballot(longrow);
                const uint32_t wid = threadIdx.x / GR_WAVE;
                __syncthreads();  // the previous chunk is done with s_lr / s_wl / s_g
                if (gr_lane() == 0) s_wl[wid] = (uint32_t)__popcll(lm);
                // from the chunk's first and last rows' lanes
                if (live && v == base) s_g[0] = e0;
                if (live && v + 1 == rend) s_g[1] = e1;
                __syncthreads();
                uint32_t off = 0;
#pragma unroll
                for (uint32_t w = 0; w < GR_BLOCK / GR_WAVE; w++) {
                    if (w < wid) off += s_wl[w];
                    nl += s_wl[w];
                }
                if (longrow) {
                    const uint32_t slot = off + (uint32_t)__popcll(lm & gr_lanemask_lt());
                    s_lr_lo[slot] = e0;
                    s_lr_hi[slot] = e1;
                }
                __syncthreads();
                g0 = s_g[0];
                g1 = s_g[1];
            }
            // segments between long rows, each in rounds of STAGE entries, covering
            // [row_off[base], row_off[rend]); every short row lies inside one segment
            uint32_t seg_lo = g0;
            for (uint32_t li = 0; li <= nl; li++) {
                const uint32_t seg_hi = li < nl ? s_lr_lo[li] : g1;
                for (uint32_t s0 = seg_lo; s0 < seg_hi; s0 += STAGE) {
                    const uint32_t s1 = s0 + STAGE < seg_hi ? s0 + STAGE : seg_hi;
                    const uint32_t cnt = s1 - s0;
                    __syncthreads();
                    // G entries per lane per pass: their index loads, then their gathers,
                    // all in flight before the first LDS store
                    constexpr int G = 8;
                    for (uint32_t q0 = threadIdx.x; q0 < cnt; q0 += G * GR_BLOCK) {
                        uint32_t jj[G];
#pragma unroll
                        for (int k = 0; k < G; k++) {
                            const uint32_t q = q0 + k * GR_BLOCK;
                            jj[k] = q < cnt ? d.nbr[s0 + q] : 0u;
                        }
                        T pj[G][D], ww[G];
#pragma unroll
                        for (int k = 0; k < G; k++) {
                            const uint32_t q = q0 + k * GR_BLOCK;
                            if (q < cnt) {
#pragma unroll
                                for (int a = 0; a < D; a++) pj[k][a] = pi[(size_t)jj[k] * D + a];
                                if (wgt) ww[k] = wgt[s0 + q];
                            }
                        }
#pragma unroll
                        for (int k = 0; k < G; k++) {
                            const uint32_t q = q0 + k * GR_BLOCK;
                            if (q < cnt) {
#pragma unroll
                                for (int a = 0; a < D; a++) s_pj[(size_t)q * D + a] = pj[k][a];
                                if (wgt) s_w[q] = ww[k];
                            }
                        }
                    }
                    __syncthreads();
                    const uint32_t lim = e1 < s1 ? e1 : s1;
                    if (wgt)
                        row_sum<T, D, KAHAN, true>(s_pj, s_w, e, lim, s0, pv, uv, cc);
                    else
                        row_sum<T, D, KAHAN, false>(s_pj, s_w, e, lim, s0, pv, uv, cc);
                }
                if (li < nl) seg_lo = s_lr_hi[li];
            }
        }
        if (!live) continue;
        if (attr && longrow) {
            uint32_t lo = 0, hi = d.nlong;  // its slot among the ascending long rows
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2;
                if (d.long_rows[mid] < v)
                    lo = mid + 1;
                else
                    hi = mid;
            }
#pragma unroll
            for (int a = 0; a < D; a++) uv[a] = ((const T *)d.long_acc)[(size_t)lo * D + a];
        }
        if (mode == 0) {
            if (addrep) {
#pragma unroll
                for (int a = 0; a < D; a++) uv[a] = uv[a] + rv[a];
            }
        } else {
#pragma unroll
            for (int a = 0; a < D; a++) uv[a] = rv[a];
        }
        const double mv = (double)mt;
        if (parts & PART_GRAVITY) {
            double cen[D], nn = 0.0;
#pragma unroll
            for (int a = 0; a < D; a++) {
                const double ma = d.gidn ? d.gr[d.gidn[v]].mean[a] : mean1[a];
                cen[a] = (double)pv[a] - ma;
                nn += cen[a] * cen[a];
            }
            double norm = sqrt(nn);
            double g = grav * mv;
#pragma unroll
            for (int a = 0; a < D; a++) {
                double unit = norm > 0.0 ? cen[a] / norm : 0.0;
                uv[a] = (T)((double)uv[a] - g * unit);
            }
        }
        double nbk = 0.0, nfw = 0.0;
#pragma unroll
        for (int a = 0; a < D; a++) {
            ((T *)d.u[a])[v] = uv[a];
            double back = (double)pv[a] - (double)uv[a];
            double fwd = (double)pv[a] + (double)uv[a];
            nbk += back * back;
            nfw += fwd * fwd;
        }
        d.nodeval[2 * (size_t)v] = mv * sqrt(nbk);
        d.nodeval[2 * (size_t)v + 1] = 0.5 * mv * sqrt(nfw);
    }
}

// ------------------------------------------------------------------ sequential sums

// The reference's sums in its own order (state.rs:208-220, schedule.rs:16-30 and :62-77):
// each one a chain from 0.0 adding the nodes' terms one after another in ascending node
// order. Lane 0 runs the chains (K side by side) over SEQ_TILE nodes' terms at a time staged
// in LDS, while the block's other waves stage the next tile (double-buffered), and reads its
// tile eight nodes ahead of its adds, so the loads hide under the add chains. The result is
// valid in thread 0. Launch with GR_BLOCK threads.
constexpr uint32_t SEQ_TILE = 2048;

template <int K, typename Term>
__device__ void seq_sum_chain(uint32_t v0, uint32_t v1, double (&s)[K], Term term) {
    __shared__ double tile[2][K * SEQ_TILE];
#pragma unroll
    for (int k = 0; k < K; k++) s[k] = 0.0;
    const uint32_t nt = (v1 - v0 + SEQ_TILE - 1) / SEQ_TILE;
    auto fill = [&](uint32_t t) {
        const uint32_t base = v0 + t * SEQ_TILE;
        const uint32_t cnt = v1 - base < SEQ_TILE ? v1 - base : SEQ_TILE;
        double *tb = tile[t & 1];
        for (uint32_t q = threadIdx.x - GR_WAVE; q < cnt; q += GR_BLOCK - GR_WAVE) {
#pragma unroll
            for (int k = 0; k < K; k++) tb[K * q + k] = term(base + q, k);
        }
    };
    if (threadIdx.x >= GR_WAVE && nt > 0) fill(0);
    __syncthreads();
    for (uint32_t t = 0; t < nt; t++) {
        if (threadIdx.x >= GR_WAVE) {
            if (t + 1 < nt) fill(t + 1);
        } else if (threadIdx.x == 0) {
            const uint32_t base = v0 + t * SEQ_TILE;
            const uint32_t cnt = v1 - base < SEQ_TILE ? v1 - base : SEQ_TILE;
            const double *tb = tile[t & 1];
            constexpr uint32_t B = 8;
            uint32_t q = 0;
            if (cnt >= B) {
                double cur[B][K];
#pragma unroll
                for (uint32_t b = 0; b < B; b++)
#pragma unroll
                    for (int k = 0; k < K; k++) cur[b][k] = tb[K * b + k];
                for (; q + 2 * B <= cnt; q += B) {
                    double nxt[B][K];
#pragma unroll
                    for (uint32_t b = 0; b < B; b++)
#pragma unroll
                        for (int k = 0; k < K; k++) nxt[b][k] = tb[K * (q + B + b) + k];
#pragma unroll
                    for (uint32_t b = 0; b < B; b++)
#pragma unroll
                        for (int k = 0; k < K; k++) s[k] = s[k] + cur[b][k];
#pragma unroll
                    for (uint32_t b = 0; b < B; b++)
#pragma unroll
                        for (int k = 0; k < K; k++) cur[b][k] = nxt[b][k];
                }
#pragma unroll
                for (uint32_t b = 0; b < B; b++)
#pragma unroll
                    for (int k = 0; k < K; k++) s[k] = s[k] + cur[b][k];
                q += B;
            }
            for (; q < cnt; q++) {
#pragma unroll
                for (int k = 0; k < K; k++) s[k] = s[k] + tb[K * q + k];
            }
        }
        __syncthreads();
    }
}

// K interleaved doubles per node at src[K * v + k]: sums k = 0..K-1 over v in [v0, v1)
template <int K>
__device__ void seq_sum_nodes(const double *src, uint32_t v0, uint32_t v1, double (&s)[K]) {
    seq_sum_chain<K>(v0, v1, s, [&](uint32_t v, int k) { return src[(size_t)K * v + k]; });
}

// The centroid as state.rs:208-220 forms it (one block per graph), over the planar
// positions; it replaces the block-reduced one k_head_graph wrote.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_seq_mean(const FaDev d) {
    if (halted(d)) return;
    const uint32_t b = blockIdx.x;
    const uint32_t v0 = d.gnode[b], v1 = d.gnode[b + 1];
    double s[D];
    seq_sum_chain<D>(v0, v1, s,
                     [&](uint32_t v, int a) { return (double)((const T *)d.pos[a])[v]; });
    if (threadIdx.x == 0) {
        const double nb = (double)(v1 - v0);
#pragma unroll
        for (int a = 0; a < D; a++) d.gr[b].mean[a] = s[a] / nb;
    }
}

// swing and traction partial sums per node chunk (chunks never cross graphs), fixed order
__global__ __launch_bounds__(GR_BLOCK) void k_stats(const FaDev d) {
    if (halted(d)) return;
    const uint32_t c0 = d.chunk[blockIdx.x], c1 = d.chunk[blockIdx.x + 1];
    double v2[2] = {0.0, 0.0};
    // lane t adds nodes c0 + t, + 256, ... in ascending order, eight loads in flight
    constexpr int SI = 8;
    const double2 *nv = (const double2 *)d.nodeval;
    for (uint32_t vb = c0; vb < c1; vb += SI * GR_BLOCK) {
        double2 q[SI];
#pragma unroll
        for (int k = 0; k < SI; k++) {
            const uint32_t v = vb + k * GR_BLOCK + threadIdx.x;
            q[k] = nv[v < c1 ? v : c0];
        }
#pragma unroll
        for (int k = 0; k < SI; k++) {
            if (vb + k * GR_BLOCK + threadIdx.x < c1) {
                v2[0] = v2[0] + q[k].x;
                v2[1] = v2[1] + q[k].y;
            }
        }
    }
    __shared__ double lds[4 * 2];
    block_reduce<2, 0>(v2, lds);
    if (threadIdx.x == 0) {
        d.spart[2 * blockIdx.x] = v2[0];
        d.spart[2 * blockIdx.x + 1] = v2[1];
    }
}

// schedule.rs:33-60 on the device, after the swing/traction sums are reduced.
__global__ __launch_bounds__(GR_BLOCK) void k_estimate(const FaDev d) {
    if (halted(d)) return;
    GrGraph *st = d.gr + blockIdx.x;
    if (st->done) return;  // block-uniform
    __shared__ double lds[4 * 2];
    double v[2] = {0.0, 0.0};
    if (d.seq_sums) {  // block-uniform
        seq_sum_nodes<2>(d.nodeval, d.gnode[blockIdx.x], d.gnode[blockIdx.x + 1], v);
    } else {
        for (uint32_t c = d.gchunk[blockIdx.x] + threadIdx.x; c < d.gchunk[blockIdx.x + 1];
             c += GR_BLOCK) {
            v[0] = v[0] + d.spart[2 * c];
            v[1] = v[1] + d.spart[2 * c + 1];
        }
        block_reduce<2, 0>(v, lds);
    }
    if (threadIdx.x != 0) return;
    st->swing += v[0];
    st->traction += v[1];
    const double swing = st->swing, traction = st->traction;
    const double n = (double)(d.gnode[blockIdx.x + 1] - d.gnode[blockIdx.x]);
    const double jt = d.jitter_tolerance;
    const double opt_jitter = 0.05 * sqrt(n);
    const double min_jitter = sqrt(opt_jitter);
    const double min_speed_efficiency = 0.05;
    const double other = fmin(10.0, opt_jitter * traction / (n * n));
    double jitter = jt * fmax(min_jitter, other);
    double se = st->speed_efficiency;
    if (swing / traction > 2.0) {
        if (se > min_speed_efficiency) se *= 0.5;
        jitter = fmax(jitter, jt);
    }
    const double target_speed = swing == 0.0 ? INFINITY : jitter * se * traction / swing;
    if (swing > jitter * traction) {
        if (se > min_speed_efficiency) se *= 0.7;
    } else if (st->speed < 1000.0) {
        se *= 1.3;
    }
    st->speed_efficiency = se;
    st->speed += fmin(target_speed - st->speed, 0.5 * st->speed);
    st->iter_no += 1;
}

// schedule.rs:62-77, fused with the next iteration's bounds/centroid partials.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_apply(const FaDev d) {
    if (halted(d)) return;
    const GrGraph &gg = d.gr[d.chunk_g[blockIdx.x]];
    const double speed = gg.speed;
    const bool frozen = gg.done != 0;  // an exited graph keeps its positions
    double moved = 0.0;
    double lo[D], hi[D], sum[D], nfin = 0.0;
#pragma unroll
    for (int a = 0; a < D; a++) {
        lo[a] = INFINITY;
        hi[a] = -INFINITY;
        sum[a] = 0.0;
    }
    const uint32_t c0 = d.chunk[blockIdx.x], c1 = d.chunk[blockIdx.x + 1];
    // AI nodes per lane (v = c0 + lane, + 256, ...): all their loads first, in flight
    // together, then their steps in that same ascending order
    constexpr int AI = sizeof(T) == 4 ? 8 : 4;
    for (uint32_t vb = c0; vb < c1; vb += AI * GR_BLOCK) {
        T pv[AI][D], uu[AI][D], mm[AI];
#pragma unroll
        for (int k = 0; k < AI; k++) {
            const uint32_t v = vb + k * GR_BLOCK + threadIdx.x;
            const uint32_t vi = v < c1 ? v : c0;  // an idle lane re-reads a valid node
#pragma unroll
            for (int a = 0; a < D; a++) pv[k][a] = ((const T *)d.pos[a])[vi];
            if (!frozen) {
#pragma unroll
                for (int a = 0; a < D; a++) uu[k][a] = ((const T *)d.u[a])[vi];
                mm[k] = ((const T *)d.mass)[vi];
            }
        }
#pragma unroll
        for (int k = 0; k < AI; k++) {
            const uint32_t v = vb + k * GR_BLOCK + threadIdx.x;
            if (v >= c1) continue;
            double p[D];
            if (frozen) {
#pragma unroll
                for (int a = 0; a < D; a++) p[a] = (double)pv[k][a];
                bounds_take<D>(p, lo, hi, sum, nfin);
                continue;
            }
            double uv[D], nn = 0.0;
#pragma unroll
            for (int a = 0; a < D; a++) {
                uv[a] = (double)uu[k][a];
                nn += uv[a] * uv[a];
            }
            const double norm = sqrt(nn);
            const double factor = speed / (1.0 + sqrt(speed * (double)mm[k] * norm));
            double step = 0.0;
#pragma unroll
            for (int a = 0; a < D; a++) {
                const double dd = uv[a] * factor;
                const T np = (T)((double)pv[k][a] + dd);
                ((T *)d.pos[a])[v] = np;
                ((T *)d.pi)[(size_t)v * D + a] = np;
                ((T *)d.pk)[(size_t)v * 4 + a] = np;
                p[a] = (double)np;
                step += fabs(dd);
            }
            moved += step;
            // the node's step for k_tail's sequential sum (its swing term is spent)
            if (d.seq_sums) d.nodeval[2 * (size_t)v] = step;
            bounds_take<D>(p, lo, hi, sum, nfin);
        }
    }
    bounds_write<D>(lo, hi, sum, nfin, d.bpart);
    __shared__ double lds[4];
    double mv[1] = {moved};
    block_reduce<1, 0>(mv, lds);
    if (threadIdx.x == 0) d.mpart[blockIdx.x] = mv[0];
}

// state.rs:122-128: the iteration counts, and one that moved less than 1e-10 ends the run.
__global__ __launch_bounds__(GR_BLOCK) void k_tail(const FaDev d) {
    if (halted(d)) return;
    GrGraph *st = d.gr + blockIdx.x;
    if (!st->done) {  // block-uniform
        __shared__ double lds[4];
        __shared__ double s_blk;
        double v[1] = {0.0};
        for (uint32_t c = d.gchunk[blockIdx.x] + threadIdx.x; c < d.gchunk[blockIdx.x + 1];
             c += GR_BLOCK)
            v[0] = v[0] + d.mpart[c];
        block_reduce<1, 0>(v, lds);
        if (d.seq_sums) {  // block-uniform
            // The reference adds the nodes' steps one after another and only compares the
            // total with 1e-10. Every order of adding n non-negative terms lies within
            // (n - 1) u S of their exact sum S (u = 2^-53), so the two orders' totals differ by
            // under 2.5 n u times the block-order one: unless 1e-10 lies that close to it, the
            // block-order total already decides the comparison the sequential one would, and
            // only then is the sequential sum formed (each node's step, left in nodeval by
            // k_apply). A NaN total decides "not done" either way.
            if (threadIdx.x == 0) s_blk = v[0];
            __syncthreads();
            const double blk = s_blk;
            const double n_g = (double)(d.gnode[blockIdx.x + 1] - d.gnode[blockIdx.x]);
            const double margin = 2.5 * n_g * 0x1p-53 * blk;
            if (blk >= 1e-10 - margin && blk <= 1e-10 + margin) {
                const double *nv = d.nodeval;
                v[0] = 0.0;
                seq_sum_chain<1>(d.gnode[blockIdx.x], d.gnode[blockIdx.x + 1], v,
                                 [&](uint32_t q, int) { return nv[2 * (size_t)q]; });
            } else {
                v[0] = blk;
            }
        }
        if (threadIdx.x == 0) {
            st->moved = v[0];
            st->iters_done += 1;
            if (v[0] < 1e-10) st->done = 1;
        }
    }
    // one graph: the session's bookkeeping (k_tail_global's) from the same thread
    if (d.ngraphs == 1 && threadIdx.x == 0) {
        GrState *gs = d.st;
        const uint32_t live = st->done ? 0u : 1u;
        if (gs->live > 0) gs->iters_done += 1;
        gs->live = live;
        if (live == 0) gs->done = 1;
    }
}

// The session's iteration count and whether every graph has exited.
__global__ __launch_bounds__(GR_BLOCK) void k_tail_global(const FaDev d) {
    GrState *st = d.st;
    if (halted(d)) return;
    __shared__ uint32_t s_live;
    if (threadIdx.x == 0) s_live = 0;
    __syncthreads();
    uint32_t live = 0;
    for (uint32_t b = threadIdx.x; b < d.ngraphs; b += GR_BLOCK) live += d.gr[b].done ? 0u : 1u;
    atomicAdd(&s_live, live);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (st->live > 0) st->iters_done += 1;
        st->live = s_live;
        if (s_live == 0) st->done = 1;
    }
}

__global__ void k_clear_halt(const FaDev d) {
    d.st->halt = 0;
    d.st->pool_overflow = 0;
}


// ------------------------------------------------------------------ launchers

template <typename T, int D>
static void head_impl(FaDev &d) {
    hipLaunchKernelGGL((k_head_graph<D>), dim3(d.ngraphs), dim3(GR_BLOCK), 0, d.stream, d);
    if (d.seq_sums)
        hipLaunchKernelGGL((k_seq_mean<T, D>), dim3(d.ngraphs), dim3(GR_BLOCK), 0, d.stream, d);
    if (d.ngraphs > 1)
        hipLaunchKernelGGL(k_head_global, dim3(1), dim3(GR_BLOCK), 0, d.stream, d);
}

template <typename T, int D>
static void tree_impl(FaDev &d, bool com = true) {
    const uint32_t n = d.n;
    const uint32_t nbl = gr_blocks(n, GR_BLOCK);
    hipStream_t s = d.stream;
    head_impl<T, D>(d);
    // keys are written where an odd number of sort passes starts, so the sorted pairs
    // always end in (d.keys, d.vals)
    FaDev dk = d;
    if (radix_digits((int)d.key_bits) & 1) {
        dk.keys = d.keys_alt;
        dk.vals = d.vals_alt;
    }
    hipLaunchKernelGGL((k_keys<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, dk);
    radix_sort_pairs(dk.keys, dk.vals, dk.keys == d.keys ? d.keys_alt : d.keys,
                     dk.vals == d.vals ? d.vals_alt : d.vals, n, (int)d.key_bits, d.sort_tmp, s,
                     (int)d.sort_items);
    hipLaunchKernelGGL((k_gather<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
    scan_exclusive(d.flag, d.gid, n, nullptr, &d.st->n_groups, d.scan_tmp, s);
    hipLaunchKernelGGL((k_group_compact<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL((k_group_info<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
    scan_exclusive(d.gcells, d.gbase, n, &d.st->n_groups, &d.st->need_cells, d.scan_tmp, s);
    hipLaunchKernelGGL((k_emit<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
    if (!com) return;  // the arena alone
    const uint32_t ntiles = gr_blocks(d.cap, COM_TILE);
    hipLaunchKernelGGL((k_com_tiles<T, D>), dim3(ntiles), dim3(COM_THREADS), 0, s, d);
    const uint32_t span_blocks = gr_blocks((uint64_t)ntiles * SPAN_SLOTS, GR_BLOCK);
    hipLaunchKernelGGL((k_span_prep<T, D>), dim3(span_blocks), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL((k_span_link<D>), dim3(span_blocks), dim3(GR_BLOCK), 0, s, d);
    // a batch's span cells overflow the one-workgroup LDS pass: it then runs level by level
    // over the whole GPU
    const bool levelwise = d.ngraphs > 1;
    hipLaunchKernelGGL((k_com_span<T, D>), dim3(1), dim3(1024), 0, s, d, (uint32_t)levelwise);
    if (levelwise) {
        for (int l = d.levels < 31 ? (int)d.levels : 31; l >= 0; l--)
            hipLaunchKernelGGL((k_com_span_level<T, D>), dim3(gr_xcd_grid(64)), dim3(GR_BLOCK), 0,
                               s, d, l);
    }
}

// Tree positions per walk wave: 64 from 16,384 bodies up; a smaller graph's walk, a few
// dozen latency-bound waves, gets narrower waves (at least 8 bodies, the other lanes idle)
// and more of them. A lane's terms do not depend on the lanes beside it: the same bits.
static uint32_t walk_group(uint32_t n) {
    uint32_t g = GR_WAVE;
    while (g > 8 && (uint64_t)n < (uint64_t)g * 256) g >>= 1;
    return g;
}

template <typename T, int D, bool KAHAN, bool COUNT, int WB>
static void walk_launch(FaDev &d) {
    d.walk_group = walk_group(d.n);
    const uint32_t nb = gr_blocks(d.n, (WB / GR_WAVE) * d.walk_group);
    const dim3 grid(d.walk_xcd ? gr_xcd_grid(nb) : nb);
    const bool buf = (uint64_t)(d.cap + 1) * sizeof(WalkCell<T, D>) < (1ull << 31);
    const bool one = d.gidn == nullptr;
    if (buf && one)
        hipLaunchKernelGGL((k_walk<T, D, KAHAN, true, true, COUNT, WB>), grid, dim3(WB), 0,
                           d.stream, d);
    else if (buf)
        hipLaunchKernelGGL((k_walk<T, D, KAHAN, true, false, COUNT, WB>), grid, dim3(WB), 0,
                           d.stream, d);
    else if (!COUNT)  // the measuring variant needs buffer loads (an arena under 2 GiB)
        hipLaunchKernelGGL((k_walk<T, D, KAHAN, false, false, false, WB>), grid, dim3(WB), 0,
                           d.stream, d);
}

template <typename T, int D, bool KAHAN>
static void walk_impl(FaDev &d, bool count = false) {
    if (count) {
        walk_launch<T, D, KAHAN, true, 256>(d);
    } else if (d.walk_block == 64) {
        walk_launch<T, D, KAHAN, false, 64>(d);
    } else if (d.walk_block == 128) {
        walk_launch<T, D, KAHAN, false, 128>(d);
    } else {
        walk_launch<T, D, KAHAN, false, 256>(d);
    }
}

void pm_repulsion(FaDev &d);  // pm.hip

template <typename T, int D, bool KAHAN>
static void forces_impl(FaDev &d, uint32_t parts) {
    const uint32_t nbl = gr_blocks(d.n, GR_BLOCK);
    hipStream_t s = d.stream;
    uint32_t mode = 0;
    if (d.repulsion == REP_TREE) {
        tree_impl<T, D>(d);
        if (parts & PART_REPULSION) walk_impl<T, D, KAHAN>(d);
    } else if (d.repulsion == REP_DENSE) {
        head_impl<T, D>(d);
        if (parts & PART_ATTRACTION)
            hipLaunchKernelGGL((k_attract_only<T, D, KAHAN>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
        if (parts & PART_REPULSION) {
            hipLaunchKernelGGL((k_dense<T, D, KAHAN>), dim3(nbl), dim3(GR_BLOCK), 0, s, d,
                               (uint32_t)((parts & PART_ATTRACTION) ? 1 : 0));
        } else if (!(parts & PART_ATTRACTION)) {
            hipLaunchKernelGGL((k_zero_u<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d);
        }
        mode = 1;
    } else if (d.repulsion == REP_PM) {
        head_impl<T, D>(d);
        if (parts & PART_REPULSION) pm_repulsion(d);
    } else {
        head_impl<T, D>(d);
        parts &= ~(uint32_t)PART_REPULSION;
    }
    if (mode == 0 && (parts & PART_ATTRACTION) && d.nlong > 0)
        hipLaunchKernelGGL((k_long_rows<T, D, KAHAN>), dim3(d.nlong), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL((k_finish<T, D, KAHAN>), dim3(d.nparts), dim3(GR_BLOCK),
                       FIN_LDS, s, d, mode, parts);
    hipLaunchKernelGGL(k_stats, dim3(d.nchunks), dim3(GR_BLOCK), 0, s, d);
}

template <typename T, int D, bool KAHAN>
static void iteration_impl(FaDev &d) {
    hipStream_t s = d.stream;
    forces_impl<T, D, KAHAN>(d, PART_ATTRACTION | PART_REPULSION | PART_GRAVITY);
    hipLaunchKernelGGL(k_estimate, dim3(d.ngraphs), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL((k_apply<T, D>), dim3(d.nchunks), dim3(GR_BLOCK), 0, s, d);
    hipLaunchKernelGGL(k_tail, dim3(d.ngraphs), dim3(GR_BLOCK), 0, s, d);
    if (d.ngraphs > 1) hipLaunchKernelGGL(k_tail_global, dim3(1), dim3(GR_BLOCK), 0, s, d);
}

#define GR_DISPATCH(FN, ...)                                                                   \
    do {                                                                                       \
        if (d.precision == PREC_FP64) {                                                        \
            if (d.dim == 2)                                                                    \
                FN<double, 2, false>(__VA_ARGS__);                                             \
            else                                                                               \
                FN<double, 3, false>(__VA_ARGS__);                                             \
        } else if (d.precision == PREC_FP32) {                                                 \
            if (d.dim == 2)                                                                    \
                FN<float, 2, false>(__VA_ARGS__);                                              \
            else                                                                               \
                FN<float, 3, false>(__VA_ARGS__);                                              \
        } else {                                                                               \
            if (d.dim == 2)                                                                    \
                FN<float, 2, true>(__VA_ARGS__);                                               \
            else                                                                               \
                FN<float, 3, true>(__VA_ARGS__);                                               \
        }                                                                                      \
    } while (0)

template <typename T, int D, bool KAHAN>
static void bounds_impl(FaDev &d) {
    hipLaunchKernelGGL((k_bounds<T, D>), dim3(d.nchunks), dim3(GR_BLOCK), 0, d.stream, d);
}

template <typename T, int D, bool KAHAN>
static void tree_only_impl(FaDev &d, bool com) {
    tree_impl<T, D>(d, com);
}

void launch_bounds(FaDev &d) { GR_DISPATCH(bounds_impl, d); }
void launch_iteration(FaDev &d) { GR_DISPATCH(iteration_impl, d); }
void launch_forces(FaDev &dd, uint32_t parts) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(forces_impl, d, parts);
}
void launch_tree(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(tree_only_impl, d, true);
}
void launch_tree_arena(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(tree_only_impl, d, false);
}
void launch_walk(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(walk_impl, d);
}
template <typename T, int D, bool KAHAN>
static void finish_impl(FaDev &d) {
    if (d.nlong > 0)
        hipLaunchKernelGGL((k_long_rows<T, D, KAHAN>), dim3(d.nlong), dim3(GR_BLOCK), 0, d.stream,
                           d);
    hipLaunchKernelGGL((k_finish<T, D, KAHAN>), dim3(d.nparts), dim3(GR_BLOCK), FIN_LDS, d.stream,
                       d, 0u, (uint32_t)(PART_ATTRACTION | PART_REPULSION | PART_GRAVITY));
}
void launch_finish(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(finish_impl, d);
}
void launch_walk_count(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    GR_DISPATCH(walk_impl, d, true);
}
void launch_clear_halt(FaDev &d) {
    hipLaunchKernelGGL(k_clear_halt, dim3(1), dim3(1), 0, d.stream, d);
}
void launch_pm_charge(FaDev &dd) {
    FaDev d = dd;
    d.ignore_done = 1;
    bounds_impl<double, 2, false>(d);
    head_impl<double, 2>(d);
    pm_repulsion(d);
}

}  // namespace gr
