// layout.forceatlas2.forcesim on gfx950: SciGraphs' own ForceSim (model FA2, direct
// repulsion) as the reference ports it (paths relative to crates/graph-core/src/):
//   stage      layout/force/forcesim.rs:158-187 (Stage::run, NonFinite, f32 columns)
//   start      layout/force/forcesim/seed.rs + rng/pcg64.rs (numpy's default_rng)
//   constants  layout/force/forcesim/state.rs:117-132
//   forces     layout/force/forcesim/{pair_force,attraction,gravity}.rs
//   step       layout/force/forcesim/integrate.rs, state.rs:100-113
//   answer     layout/force/forcesim/rescale.rs
//   orders     layout/force/forcesim/reduce.rs
//
// Every arithmetic step is the reference's, in its dtype and its order: f32 positions and
// forces, f64 coefficient rows, numpy's pairwise sums (eight lanes per block of at most 128,
// halves split at a multiple of eight), sequential column means and bincount-ordered f64
// accumulators. The layout is a few hundred nodes by design (the reference's DIRECT_MAX 400,
// its scale ceiling 2000), so one workgroup lays out one graph from start to rescale in a
// single launch: a thread per force row, the sequential sums on single threads, barriers
// between the phases. A batch of graphs is one launch with a workgroup per graph.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/graph_render_fa2.h"
#include "gr_device.h"

extern "C" gr_status gr_internal_fail(gr_status code, const char *name, const char *rule);
void gr_internal_simple_edges(uint32_t n, const uint32_t *src, const uint32_t *dst, size_t m,
                              std::vector<uint32_t> &lo, std::vector<uint32_t> &hi);

namespace gr {
namespace {

// ---------------------------------------------------------------- host: numpy's default_rng

// SeedSequence(seed).generate_state(4, uint64) (numpy bit_generator.pyx): the seed's 32-bit
// words hashed into a pool of four, every pool word mixed into every other, then eight
// output words drawn from the pool.
struct SeedSeq {
    static uint32_t hashmix(uint32_t v, uint32_t &h) {
        v ^= h;
        h *= 0x931e8875u;
        v *= h;
        return v ^ (v >> 16);
    }
    static uint32_t mix(uint32_t x, uint32_t y) {
        const uint32_t r = 0xca01f9ddu * x - 0x4973f715u * y;
        return r ^ (r >> 16);
    }
    static void words(uint64_t seed, uint64_t w[4]) {
        const uint32_t ent[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
        const int ne = (seed >> 32) ? 2 : 1;
        uint32_t pool[4], h = 0x43b0d7e5u;
        for (int i = 0; i < 4; i++) pool[i] = hashmix(i < ne ? ent[i] : 0u, h);
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++)
                if (a != b) pool[b] = mix(pool[b], hashmix(pool[a], h));
        uint32_t hb = 0x8b51f9ddu, o[8];
        for (int i = 0; i < 8; i++) {
            uint32_t v = pool[i & 3] ^ hb;
            hb *= 0x58f38dedu;
            v *= hb;
            o[i] = v ^ (v >> 16);
        }
        for (int i = 0; i < 4; i++) w[i] = ((uint64_t)o[2 * i + 1] << 32) | o[2 * i];
    }
};

// PCG-XSL-RR 128/64 seeded as numpy's pcg64_set_seed (rng/pcg64.rs:44-110): the 128-bit
// LCG steps, then the output is the xor of its halves rotated by the top six bits.
struct Pcg64 {
    unsigned __int128 state = 0, inc = 0;
    static constexpr unsigned __int128 MULT =
        ((unsigned __int128)2549297995355413924ull << 64) | 4865540595714422341ull;
    explicit Pcg64(uint64_t seed) {
        uint64_t w[4];
        SeedSeq::words(seed, w);
        const unsigned __int128 init = ((unsigned __int128)w[0] << 64) | w[1];
        const unsigned __int128 seq = ((unsigned __int128)w[2] << 64) | w[3];
        inc = (seq << 1) | 1u;
        step();
        state += init;
        step();
    }
    void step() { state = state * MULT + inc; }
    uint64_t next() {
        step();
        const uint64_t hi = (uint64_t)(state >> 64), x = hi ^ (uint64_t)state;
        const unsigned r = (unsigned)(hi >> 58);
        return r ? (x >> r) | (x << (64 - r)) : x;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// ---------------------------------------------------------------- host: constants

// The double nearest the cube root of an integer x in [1, 2^32] (libm's cbrt, which the
// reference's k takes, is correctly rounded; the C library's need not be): from the C
// library's answer, step to a neighbour while the exact cube of the midpoint between them
// lies on the neighbour's side of x.
int mid_cube_sign(double a, double b, uint64_t x) {
    // a < b adjacent and positive: a = A 2^e, b = B 2^e with integers A, B below 2^54
    int ea = 0, eb = 0;
    const double fa = std::frexp(a, &ea), fb = std::frexp(b, &eb);
    const int e = std::min(ea, eb) - 53;
    const uint64_t A = (uint64_t)std::ldexp(fa, ea - e), B = (uint64_t)std::ldexp(fb, eb - e);
    // ((A + B) 2^(e - 1))^3 against x: S^3 against x 2^k, k = 3 (1 - e)
    const unsigned __int128 S = (unsigned __int128)A + B;  // below 2^55
    const int k = 3 * (1 - e);
    if (k < 0 || k > 190) return k < 0 ? 1 : -1;
    const unsigned __int128 s2 = S * S;                    // below 2^110
    const unsigned __int128 lo = (s2 & ~0ull) * S;         // below 2^119
    const unsigned __int128 hi = (s2 >> 64) * S + (lo >> 64);
    const uint64_t cube[4] = {(uint64_t)lo, (uint64_t)hi, (uint64_t)(hi >> 64), 0};
    uint64_t rhs[4] = {0, 0, 0, 0};
    rhs[k / 64] = x << (k % 64);
    if (k % 64 && k / 64 + 1 < 4) rhs[k / 64 + 1] = x >> (64 - k % 64);
    for (int q = 3; q >= 0; q--)
        if (cube[q] != rhs[q]) return cube[q] > rhs[q] ? 1 : -1;
    return 0;
}

double cube_root_nearest(uint32_t n) {
    double y = std::cbrt((double)n);
    for (int it = 0; it < 4; it++) {
        const double down = std::nextafter(y, 0.0), up = std::nextafter(y, INFINITY);
        if (mid_cube_sign(down, y, n) > 0)
            y = down;
        else if (mid_cube_sign(y, up, n) < 0)
            y = up;
        else
            break;
    }
    return y;
}

// numpy's pairwise sum as a program over the array's positions: LEAF(lo, len) pushes the sum
// of a block of at most 128, ADD adds the top two. Leaves come in ascending order.
constexpr uint32_t FS_ADD = 0xFFFFFFFFu;
void pairwise_program(uint32_t lo, uint32_t len, std::vector<uint32_t> &out) {
    if (len <= 128) {
        out.push_back((lo << 8) | len);
        return;
    }
    uint32_t split = len / 2;
    split -= split % 8;
    pairwise_program(lo, split, out);
    pairwise_program(lo + split, len - split, out);
    out.push_back(FS_ADD);
}

float pairwise_host(const float *v, uint32_t n) {  // mass.mean()'s sum
    std::vector<uint32_t> prog;
    pairwise_program(0, n, prog);
    std::vector<float> st;
    for (uint32_t t : prog) {
        if (t == FS_ADD) {
            const float r = st.back();
            st.pop_back();
            st.back() = st.back() + r;
            continue;
        }
        const uint32_t lo = t >> 8, len = t & 0xFF;
        float s = 0.0f;
        if (len < 8) {
            for (uint32_t q = 0; q < len; q++) s += v[lo + q];
        } else {
            float l[8];
            for (int q = 0; q < 8; q++) l[q] = v[lo + q];
            const uint32_t stop = len - len % 8;
            for (uint32_t g = 8; g < stop; g += 8)
                for (int q = 0; q < 8; q++) l[q] += v[lo + g + q];
            s = ((l[0] + l[1]) + (l[2] + l[3])) + ((l[4] + l[5]) + (l[6] + l[7]));
            for (uint32_t q = stop; q < len; q++) s += v[lo + q];
        }
        st.push_back(s);
    }
    return st.empty() ? 0.0f : st[0];
}

// ---------------------------------------------------------------- device

struct FsGraph {
    uint32_t n, node0, tok0, ntok;
    uint32_t iters, gid, pad[2];  // gid: the graph's index in the caller's batch
    double k, rep, grav, jt, scale;
    float soften, pad2;
};

struct FsDev {
    const FsGraph *g;
    float *pos, *prev, *force, *s2, *sw, *wsw, *wtr;
    const float *mass;
    const uint32_t *aoff, *adj, *elo, *ehi, *tok;
    float *out;
    int32_t *status;
    uint32_t lds_n;  // node records the LDS variant holds (the launch's largest graph)
};

constexpr uint32_t FS_BLOCK = 256;
constexpr int FS_STACK = 40;

// One block of numpy's pairwise sum, reading the values in ascending position.
template <typename T, typename F>
__device__ __forceinline__ T fs_leaf(uint32_t lo, uint32_t len, F &val) {
    if (len < 8) {
        T t = 0;
        for (uint32_t q = 0; q < len; q++) t += val(lo + q);
        return t;
    }
    T l0 = val(lo);
    T l1 = val(lo + 1);
    T l2 = val(lo + 2);
    T l3 = val(lo + 3);
    T l4 = val(lo + 4);
    T l5 = val(lo + 5);
    T l6 = val(lo + 6);
    T l7 = val(lo + 7);
    const uint32_t stop = len - len % 8;
    for (uint32_t g = 8; g < stop; g += 8) {
        l0 += val(lo + g);
        l1 += val(lo + g + 1);
        l2 += val(lo + g + 2);
        l3 += val(lo + g + 3);
        l4 += val(lo + g + 4);
        l5 += val(lo + g + 5);
        l6 += val(lo + g + 6);
        l7 += val(lo + g + 7);
    }
    T t = ((l0 + l1) + (l2 + l3)) + ((l4 + l5) + (l6 + l7));
    for (uint32_t q = stop; q < len; q++) t += val(lo + q);
    return t;
}

template <typename T, typename F>
__device__ T fs_pairwise(const uint32_t *tok, uint32_t ntok, F &val) {
    T st[FS_STACK];
    int sp = 0;
    for (uint32_t q = 0; q < ntok; q++) {
        const uint32_t t = tok[q];
        if (t == FS_ADD) {
            if (sp >= 2) {
                st[sp - 2] = st[sp - 2] + st[sp - 1];
                sp--;
            }
        } else if (sp < FS_STACK) {
            st[sp++] = fs_leaf<T>(t >> 8, t & 0xFFu, val);
        }
    }
    return sp ? st[0] : T(0);
}

__device__ __forceinline__ float fs_dot3(float a0, float a1, float a2, float b0, float b1,
                                         float b2) {
    const float first = a0 * b0;
    const float second = a1 * b1;
    return (first + second) + a2 * b2;
}

// Column means of an (n, 3) f32 array, sequential per column: threads 0-2 take a column each.
__device__ __forceinline__ void fs_column_means(const float *p, uint32_t n, float *out3) {
    if (threadIdx.x < 3) {
        const uint32_t c = threadIdx.x;
        float s = 0.0f;
        for (uint32_t i = 0; i < n; i++) s += p[3 * i + c];
        out3[c] = s / (float)n;
    }
}

// LDS: every graph of the launch has at most FS_LDS_NODES nodes, and each iteration stages
// its positions with their own dots (x, y, z, |p|^2) and the masses in LDS, where the rows
// read them as broadcasts; otherwise the rows read global memory.
constexpr uint32_t FS_LDS_NODES = 2048;

template <bool LDS>
__global__ __launch_bounds__(FS_BLOCK) void k_fs_layout(const FsDev d) {
    extern __shared__ float4 fs_lds[];
    float4 *Q = fs_lds;                          // LDS: n records (x, y, z, s2)
    float *MS = (float *)(fs_lds + d.lds_n);     // LDS: n masses
    const FsGraph G = d.g[blockIdx.x];
    const uint32_t n = G.n, T = FS_BLOCK, tid = threadIdx.x;
    float *P = d.pos + 3 * (size_t)G.node0;
    float *PV = d.prev + 3 * (size_t)G.node0;
    float *F = d.force + 3 * (size_t)G.node0;
    float *S2 = d.s2 + G.node0;
    float *SW = d.sw + G.node0;
    float *WSW = d.wsw + G.node0;
    float *WTR = d.wtr + G.node0;
    const float *M = d.mass + G.node0;
    const uint32_t *tok = d.tok + G.tok0;
    __shared__ float cw[3], cn[3];
    __shared__ float tot[2];
    __shared__ double speed_s;
    __shared__ double mean64[3];
    __shared__ double spanw[FS_BLOCK / GR_WAVE];
    if (tid == 0) speed_s = 1.0;
    if constexpr (LDS)
        for (uint32_t j = tid; j < n; j += T) MS[j] = M[j];
    const uint32_t steps = G.iters > 0 ? G.iters : 1;
    for (uint32_t it = 0; it < steps && n >= 2; it++) {
        // center_was (before the field), and every row's own dot
        fs_column_means(P, n, cw);
        for (uint32_t j = tid; j < n; j += T) {
            const float x = P[3 * j], y = P[3 * j + 1], z = P[3 * j + 2];
            const float s2 = fs_dot3(x, y, z, x, y, z);
            S2[j] = s2;
            if constexpr (LDS) Q[j] = make_float4(x, y, z, s2);
        }
        __syncthreads();
        // the field, a row per thread: repulsion, then attraction, then gravity
        for (uint32_t i = tid; i < n; i += T) {
            const float xi = P[3 * i], yi = P[3 * i + 1], zi = P[3 * i + 2];
            const float own = S2[i];
            const double mi = (double)M[i];
            double px = 0.0, py = 0.0, pz = 0.0;
            auto coef = [&](uint32_t j) -> double {
                float xj, yj, zj, sj, mj;
                if constexpr (LDS) {
                    const float4 q = Q[j];
                    xj = q.x;
                    yj = q.y;
                    zj = q.z;
                    sj = q.w;
                    mj = MS[j];
                } else {
                    xj = P[3 * j];
                    yj = P[3 * j + 1];
                    zj = P[3 * j + 2];
                    sj = S2[j];
                    mj = M[j];
                }
                float d2 = own + sj;
                d2 = d2 - 2.0f * fs_dot3(xi, yi, zi, xj, yj, zj);
                if (d2 < 0.0f) d2 = 0.0f;
                d2 = d2 + G.soften;
                double c = G.rep * mi * (double)mj / (double)d2;
                if (j == i) c = 0.0;  // the diagonal, zeroed after the division
                px += c * (double)xj;
                py += c * (double)yj;
                pz += c * (double)zj;
                return c;
            };
            const double total = fs_pairwise<double>(tok, G.ntok, coef);
            float fx = (float)((double)xi * total - px);
            float fy = (float)((double)yi * total - py);
            float fz = (float)((double)zi * total - pz);
            // bincount order: the node's edges as source, then as destination, ascending
            double ax = 0.0, ay = 0.0, az = 0.0;
            for (uint32_t q = d.aoff[G.node0 + i]; q < d.aoff[G.node0 + i + 1]; q++) {
                const uint32_t ent = d.adj[q], e = ent & 0x7FFFFFFFu;
                const uint32_t a = d.elo[e], b = d.ehi[e];
                const float factor = 1.0f / (0.5f * (M[a] + M[b]));
                const float qx = (P[3 * b] - P[3 * a]) * factor;
                const float qy = (P[3 * b + 1] - P[3 * a + 1]) * factor;
                const float qz = (P[3 * b + 2] - P[3 * a + 2]) * factor;
                if (ent >> 31) {
                    ax -= (double)qx;
                    ay -= (double)qy;
                    az -= (double)qz;
                } else {
                    ax += (double)qx;
                    ay += (double)qy;
                    az += (double)qz;
                }
            }
            fx += (float)ax;
            fy += (float)ay;
            fz += (float)az;
            const float dx = xi - cw[0], dy = yi - cw[1], dz = zi - cw[2];
            const float dist = sqrtf(fs_dot3(dx, dy, dz, dx, dy, dz));
            const double scale = -(G.grav * mi / (double)fmaxf(dist, 1e-9f));
            fx += (float)(scale * (double)dx);
            fy += (float)(scale * (double)dy);
            fz += (float)(scale * (double)dz);
            F[3 * i] = fx;
            F[3 * i + 1] = fy;
            F[3 * i + 2] = fz;
        }
        __syncthreads();
        // swing and traction per node, weighted by mass
        for (uint32_t i = tid; i < n; i += T) {
            const float f0 = F[3 * i], f1 = F[3 * i + 1], f2 = F[3 * i + 2];
            const float p0 = PV[3 * i], p1 = PV[3 * i + 1], p2 = PV[3 * i + 2];
            const float s0 = f0 - p0, s1 = f1 - p1, s2 = f2 - p2;
            const float sw = sqrtf(fs_dot3(s0, s1, s2, s0, s1, s2));
            const float t0 = 0.5f * (f0 + p0), t1 = 0.5f * (f1 + p1), t2 = 0.5f * (f2 + p2);
            SW[i] = sw;
            WSW[i] = M[i] * sw;
            WTR[i] = M[i] * sqrtf(fs_dot3(t0, t1, t2, t0, t1, t2));
        }
        __syncthreads();
        // the two totals in numpy's pairwise order, one thread each (two waves)
        if (tid == 0 || tid == GR_WAVE) {
            const float *v = tid == 0 ? WSW : WTR;
            auto at = [&](uint32_t j) -> float { return v[j]; };
            tot[tid == 0 ? 0 : 1] = fs_pairwise<float>(tok, G.ntok, at);
        }
        __syncthreads();
        if (tid == 0) {
            double sp = speed_s;
            if (tot[0] > 0.0f) {
                const double target = G.jt * (double)tot[1] / (double)tot[0];
                const double lo = sp * 0.5, hi = sp * 1.5;
                sp = target < lo ? lo : target > hi ? hi : target;
            }
            sp = sp < 1e-4 ? 1e-4 : sp > 10.0 ? 10.0 : sp;
            speed_s = sp;
        }
        __syncthreads();
        // the capped move; the field becomes the previous one
        const float spf = (float)speed_s;
        for (uint32_t i = tid; i < n; i += T) {
            const float factor = spf / (1.0f + spf * sqrtf(SW[i]));
            float d0 = F[3 * i] * factor, d1 = F[3 * i + 1] * factor, d2 = F[3 * i + 2] * factor;
            const float norm = sqrtf(fs_dot3(d0, d1, d2, d0, d1, d2));
            if ((double)norm > G.k) {
                const double shrink = G.k / (double)norm;
                d0 = (float)((double)d0 * shrink);
                d1 = (float)((double)d1 * shrink);
                d2 = (float)((double)d2 * shrink);
            }
            P[3 * i] += d0;
            P[3 * i + 1] += d1;
            P[3 * i + 2] += d2;
            PV[3 * i] = F[3 * i];
            PV[3 * i + 1] = F[3 * i + 1];
            PV[3 * i + 2] = F[3 * i + 2];
        }
        __syncthreads();
        // recentre on the mean the step began at
        fs_column_means(P, n, cn);
        __syncthreads();
        for (uint32_t i = tid; i < n; i += T) {
            P[3 * i] -= cn[0] - cw[0];
            P[3 * i + 1] -= cn[1] - cw[1];
            P[3 * i + 2] -= cn[2] - cw[2];
        }
        __syncthreads();
    }
    // rescale.rs: f64, centred by sequential column means, the widest |coordinate| to scale
    if (tid < 3) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; i++) s += (double)P[3 * i + tid];
        mean64[tid] = s / (double)n;
    }
    __syncthreads();
    double span = 0.0;
    for (uint32_t q = tid; q < 3 * n; q += T) {
        const double a = fabs((double)P[q] - mean64[q % 3]);
        span = span >= a || isnan(a) ? span : a;
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const double o = __shfl_xor(span, m, GR_WAVE);
        span = span >= o || isnan(o) ? span : o;
    }
    if (gr_lane() == 0) spanw[tid / GR_WAVE] = span;
    __syncthreads();
    span = 0.0;
    for (uint32_t w = 0; w < T / GR_WAVE; w++) span = span >= spanw[w] || isnan(spanw[w]) ? span : spanw[w];
    const bool fit = span > 1e-9;
    const double f = G.scale / span;
    bool bad = false;
    for (uint32_t q = tid; q < 3 * n; q += T) {
        double v = (double)P[q] - mean64[q % 3];
        if (fit) v *= f;
        bad |= !isfinite(v);
        d.out[3 * (size_t)G.node0 + q] = (float)v;
    }
    if (__syncthreads_or(bad) && tid == 0) d.status[G.gid] = 2;
}

// ---------------------------------------------------------------- host driver

struct FsBatch {
    std::vector<FsGraph> graphs;
    std::vector<float> start, mass;
    std::vector<uint32_t> aoff, adj, elo, ehi, tok;
    uint32_t N = 0;
};

gr_status fs_prepare(uint32_t ngraphs, const uint32_t *node_counts, const size_t *edge_counts,
                     const uint32_t *source, const uint32_t *target,
                     const gr_forcesim_params *params, FsBatch &b) {
    uint64_t N = 0;
    for (uint32_t g = 0; g < ngraphs; g++) N += node_counts[g];
    if (N >= 0x7FFFFFFFull) return gr_internal_fail(GR_ERR_CAPACITY, "nodes", "nodes fit 31 bits");
    b.N = (uint32_t)N;
    b.graphs.resize(ngraphs);
    b.start.resize((size_t)N * 3);
    b.mass.resize(N);
    b.aoff.assign((size_t)N + 1, 0);
    size_t e0 = 0;
    uint32_t node0 = 0;
    std::vector<std::vector<uint32_t>> adj_of;
    for (uint32_t g = 0; g < ngraphs; g++) {
        const uint32_t n = node_counts[g];
        const size_t m = edge_counts[g];
        for (size_t e = 0; e < m; e++)
            if (source[e0 + e] >= n || target[e0 + e] >= n)
                return gr_internal_fail(GR_ERR_PARAM, "edges", "edge endpoints below the node count");
        std::vector<uint32_t> lo, hi;
        gr_internal_simple_edges(n, source + e0, target + e0, m, lo, hi);
        e0 += m;
        const uint32_t ebase = (uint32_t)b.elo.size();
        if ((uint64_t)ebase + lo.size() >= 0x7FFFFFFFull)
            return gr_internal_fail(GR_ERR_CAPACITY, "edges", "simple edges fit 31 bits");
        std::vector<uint32_t> deg(n, 0);
        for (size_t e = 0; e < lo.size(); e++) {
            deg[lo[e]]++;
            deg[hi[e]]++;
            b.elo.push_back(lo[e]);  // graph-local: the kernel indexes its graph's arrays
            b.ehi.push_back(hi[e]);
        }
        for (uint32_t v = 0; v < n; v++) b.mass[node0 + v] = (float)(deg[v] + 1);
        // each node's edges: as source (ascending), then as destination (ascending, sign bit)
        std::vector<uint32_t> cnt(n + 1, 0);
        for (size_t e = 0; e < lo.size(); e++) {
            cnt[lo[e] + 1]++;
            cnt[hi[e] + 1]++;
        }
        for (uint32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
        const uint32_t abase = (uint32_t)b.adj.size();
        b.adj.resize(abase + cnt[n]);
        std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
        for (size_t e = 0; e < lo.size(); e++) b.adj[abase + fill[lo[e]]++] = ebase + (uint32_t)e;
        for (size_t e = 0; e < lo.size(); e++)
            b.adj[abase + fill[hi[e]]++] = (ebase + (uint32_t)e) | 0x80000000u;
        for (uint32_t v = 0; v <= n; v++) b.aoff[node0 + v] = abase + cnt[v];
        // constants (state.rs:117-132)
        const gr_forcesim_params &p = params[g];
        const double cr = cube_root_nearest(n > 0 ? n : 1);
        const double k = 5.0 / (cr > 1.0 ? cr : 1.0);
        const double mean =
            n == 0 ? 1.0 : (double)(pairwise_host(b.mass.data() + node0, n) / (float)n);
        const double m2 = std::max(mean * mean, 1e-9), m3 = std::max(mean * mean * mean, 1e-9);
        FsGraph &G = b.graphs[g];
        std::memset(&G, 0, sizeof(G));
        G.n = n;
        G.node0 = node0;
        G.iters = p.iterations;
        G.k = k;
        G.rep = p.scaling_ratio * (k * k / m3);
        G.grav = p.gravity * 0.1 * (0.38 * 2.0 * (5.0 * 5.0) / (std::max(k, 1e-9) * m2));
        G.jt = p.jitter_tolerance;
        G.scale = p.scale;
        const double soft = 0.01 * k;
        G.soften = (float)(soft * soft);
        G.tok0 = (uint32_t)b.tok.size();
        pairwise_program(0, n, b.tok);
        G.ntok = (uint32_t)b.tok.size() - G.tok0;
        // the seeded start (seed.rs): row-major draws, (u - 0.5) * 5 narrowed to f32
        Pcg64 rng(p.seed);
        for (size_t q = 0; q < (size_t)n * 3; q++)
            b.start[(size_t)node0 * 3 + q] = (float)((rng.uniform() - 0.5) * 5.0);
        node0 += n;
    }
    if (b.adj.empty()) b.adj.push_back(0);
    if (b.elo.empty()) {
        b.elo.push_back(0);
        b.ehi.push_back(0);
    }
    return GR_OK;
}

struct FsDevice {
    std::vector<void *> allocs;
    FsDev d{};
    FsGraph *graphs = nullptr;
    float *start = nullptr;
    hipStream_t stream = nullptr;
    ~FsDevice() {
        for (void *p : allocs) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
    template <typename P>
    hipError_t alloc(P **p, size_t bytes) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 8);
        if (e == hipSuccess) allocs.push_back(q);
        *p = (P *)q;
        return e;
    }
    template <typename P>
    hipError_t upload(P **p, const std::vector<P> &v) {
        hipError_t e = alloc(p, v.size() * sizeof(P));
        if (e != hipSuccess || v.empty()) return e;
        return gr_copy_wait((void *)*p, v.data(), v.size() * sizeof(P), hipMemcpyHostToDevice,
                            stream);
    }
};

#define FS_TRY(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, #x, hipGetErrorString(e_)); \
    } while (0)

gr_status fs_upload(const FsBatch &b, FsDevice &dv) {
    const size_t N = b.N ? b.N : 1, G = b.graphs.size() ? b.graphs.size() : 1;
    FS_TRY(hipStreamCreateWithFlags(&dv.stream, hipStreamNonBlocking));
    FsDev &d = dv.d;
    // Workgroup b lays out the b-th record: records in decreasing order of work (n^2 pairs per
    // iteration times iterations), so the launch deals the largest graphs out first, one per
    // CU round after round, rather than leaving some CUs several of them. Each record holds
    // its graph's offsets and index, so the order moves no result.
    std::vector<FsGraph> order = b.graphs;
    for (size_t g = 0; g < order.size(); g++) order[g].gid = (uint32_t)g;
    std::stable_sort(order.begin(), order.end(), [](const FsGraph &a, const FsGraph &c) {
        return (uint64_t)a.n * a.n * a.iters > (uint64_t)c.n * c.n * c.iters;
    });
    FS_TRY(dv.upload(&dv.graphs, order));
    FS_TRY(dv.upload(&dv.start, b.start));
    float *mass = nullptr;
    uint32_t *aoff = nullptr, *adj = nullptr, *elo = nullptr, *ehi = nullptr, *tok = nullptr;
    FS_TRY(dv.upload(&mass, b.mass));
    FS_TRY(dv.upload(&aoff, b.aoff));
    FS_TRY(dv.upload(&adj, b.adj));
    FS_TRY(dv.upload(&elo, b.elo));
    FS_TRY(dv.upload(&ehi, b.ehi));
    FS_TRY(dv.upload(&tok, b.tok));
    d.g = dv.graphs;
    d.mass = mass;
    d.aoff = aoff;
    d.adj = adj;
    d.elo = elo;
    d.ehi = ehi;
    d.tok = tok;
    FS_TRY(dv.alloc(&d.pos, N * 12));
    FS_TRY(dv.alloc(&d.prev, N * 12));
    FS_TRY(dv.alloc(&d.force, N * 12));
    FS_TRY(dv.alloc(&d.s2, N * 4));
    FS_TRY(dv.alloc(&d.sw, N * 4));
    FS_TRY(dv.alloc(&d.wsw, N * 4));
    FS_TRY(dv.alloc(&d.wtr, N * 4));
    FS_TRY(dv.alloc(&d.out, N * 12));
    FS_TRY(dv.alloc(&d.status, G * 4));
    return GR_OK;
}

// One whole layout of every graph from the resident start positions.
hipError_t fs_launch(const FsBatch &b, FsDevice &dv) {
    const size_t N = b.N;
    hipError_t e;
    if ((e = hipMemcpyAsync(dv.d.pos, dv.start, N * 12, hipMemcpyDeviceToDevice, dv.stream)) ||
        (e = hipMemsetAsync(dv.d.prev, 0, N * 12, dv.stream)) ||
        (e = hipMemsetAsync(dv.d.status, 0, b.graphs.size() * 4, dv.stream)))
        return e;
    if (b.graphs.empty()) return hipSuccess;
    uint32_t biggest = 0;
    for (const FsGraph &g : b.graphs) biggest = std::max(biggest, g.n);
    if (biggest <= FS_LDS_NODES) {
        FsDev d = dv.d;
        d.lds_n = std::max(biggest, 1u);
        hipLaunchKernelGGL(k_fs_layout<true>, dim3((uint32_t)b.graphs.size()), dim3(FS_BLOCK),
                           d.lds_n * (sizeof(float4) + sizeof(float)), dv.stream, d);
    } else
        hipLaunchKernelGGL(k_fs_layout<false>, dim3((uint32_t)b.graphs.size()), dim3(FS_BLOCK),
                           0, dv.stream, dv.d);
    return hipGetLastError();
}

gr_status fs_run(uint32_t ngraphs, const uint32_t *node_counts, const size_t *edge_counts,
                 const uint32_t *source, const uint32_t *target, const gr_forcesim_params *params,
                 float *out_x, float *out_y, float *out_z, int32_t *status, uint32_t reps,
                 double *ms_out) {
    FsBatch b;
    gr_status r = fs_prepare(ngraphs, node_counts, edge_counts, source, target, params, b);
    if (r) return r;
    FsDevice dv;
    if ((r = fs_upload(b, dv))) return r;
    FS_TRY(fs_launch(b, dv));
    FS_TRY(hipStreamSynchronize(dv.stream));
    if (reps && ms_out) {
        hipEvent_t e0, e1;
        FS_TRY(hipEventCreate(&e0));
        FS_TRY(hipEventCreate(&e1));
        double t[3];
        for (int q = 0; q < 3; q++) {
            FS_TRY(hipEventRecord(e0, dv.stream));
            for (uint32_t k = 0; k < reps; k++) FS_TRY(fs_launch(b, dv));
            FS_TRY(hipEventRecord(e1, dv.stream));
            FS_TRY(hipEventSynchronize(e1));
            float ms = 0.f;
            FS_TRY(hipEventElapsedTime(&ms, e0, e1));
            t[q] = (double)ms / reps;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        std::sort(t, t + 3);
        *ms_out = t[1];
        FS_TRY(hipGetLastError());
    }
    std::vector<float> xyz((size_t)b.N * 3);
    std::vector<int32_t> st(b.graphs.size());
    if (b.N)
        FS_TRY(gr_copy_wait(xyz.data(), dv.d.out, xyz.size() * 4, hipMemcpyDeviceToHost,
                            dv.stream));
    if (!st.empty())
        FS_TRY(gr_copy_wait(st.data(), dv.d.status, st.size() * 4, hipMemcpyDeviceToHost,
                            dv.stream));
    for (size_t i = 0; i < b.N; i++) {
        if (out_x) out_x[i] = xyz[3 * i];
        if (out_y) out_y[i] = xyz[3 * i + 1];
        if (out_z) out_z[i] = xyz[3 * i + 2];
    }
    if (status) std::memcpy(status, st.data(), st.size() * 4);
    return GR_OK;
}

}  // namespace
}  // namespace gr

extern "C" {

void gr_forcesim_params_default(gr_forcesim_params *out) {
    if (!out) return;
    out->iterations = 50;
    out->scale = 5.0;
    out->scaling_ratio = 2.0;
    out->gravity = 1.0;
    out->jitter_tolerance = 1.0;
    out->seed = 1767573729ull;
}

gr_status gr_forcesim_run(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m,
                          const gr_forcesim_params *params, float *out_x, float *out_y,
                          float *out_z) {
    auto zero = [&]() {
        for (float *o : {out_x, out_y, out_z})
            if (o) std::memset(o, 0, (size_t)n * 4);
    };
    if (n > 0 && (!out_x || !out_y || !out_z)) {
        return gr_internal_fail(GR_ERR_PARAM, "out", "three output columns of n floats");
    }
    if (m > 0 && (!source || !target)) {
        zero();
        return gr_internal_fail(GR_ERR_PARAM, "edges", "source and target columns");
    }
    gr_forcesim_params p;
    gr_forcesim_params_default(&p);
    if (params) p = *params;
    int32_t status = 0;
    gr_status r = gr::fs_run(1, &n, &m, source, target, &p, out_x, out_y, out_z, &status, 0,
                             nullptr);
    if (r) {
        zero();
        return r;
    }
    if (status) {
        zero();
        return gr_internal_fail(GR_ERR_NON_FINITE, "node.x", "finite coordinates");
    }
    return gr_internal_fail(GR_OK, nullptr, nullptr);
}

gr_status gr_forcesim_batch(uint32_t ngraphs, const uint32_t *node_counts,
                            const size_t *edge_counts, const uint32_t *source,
                            const uint32_t *target, const gr_forcesim_params *params,
                            float *out_x, float *out_y, float *out_z, int32_t *status) {
    if (ngraphs > 0 && (!node_counts || !edge_counts || !status))
        return gr_internal_fail(GR_ERR_PARAM, "graphs", "node counts, edge counts and statuses");
    gr_forcesim_params p;
    gr_forcesim_params_default(&p);
    std::vector<gr_forcesim_params> all(params ? 0 : (ngraphs ? ngraphs : 1), p);
    gr_status r = gr::fs_run(ngraphs, node_counts, edge_counts, source, target,
                             params ? params : all.data(), out_x, out_y, out_z, status, 0,
                             nullptr);
    return r ? r : gr_internal_fail(GR_OK, nullptr, nullptr);
}

gr_status gr_forcesim_batch_time(uint32_t ngraphs, const uint32_t *node_counts,
                                 const size_t *edge_counts, const uint32_t *source,
                                 const uint32_t *target, const gr_forcesim_params *params,
                                 float *out_x, float *out_y, float *out_z, int32_t *status,
                                 uint32_t reps, double *ms_out) {
    if (!ms_out || reps == 0) return gr_internal_fail(GR_ERR_PARAM, "reps", "reps >= 1 and ms_out");
    if (ngraphs > 0 && (!node_counts || !edge_counts || !status))
        return gr_internal_fail(GR_ERR_PARAM, "graphs", "node counts, edge counts and statuses");
    gr_forcesim_params p;
    gr_forcesim_params_default(&p);
    std::vector<gr_forcesim_params> all(params ? 0 : (ngraphs ? ngraphs : 1), p);
    gr_status r = gr::fs_run(ngraphs, node_counts, edge_counts, source, target,
                             params ? params : all.data(), out_x, out_y, out_z, status, reps,
                             ms_out);
    return r ? r : gr_internal_fail(GR_OK, nullptr, nullptr);
}

}  // extern "C"
