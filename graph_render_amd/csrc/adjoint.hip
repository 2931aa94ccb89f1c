// Reverse mode of unrolled FA2 iterations on gfx950: the exact transpose of one iteration
// (layout/forceatlas2/state.rs:130-148) at its recorded state, applied backward.
//
// Forward iteration t (one graph, positions p, force u, cumulative swing S and traction T,
// speed s, efficiency e):
//   u   = attraction(p; w) + k * repulsion(p) - g * m * unit(p - mean(p))
//   S' = S + sum m|p - u|,  T' = T + sum m|p + u|/2           (schedule.rs:16-30)
//   s'  = estimate(S', T', s, e)                                 (schedule.rs:33-60)
//   p'  = p + u * s' / (1 + sqrt(s' m |u|))                      (schedule.rs:62-77)
// The efficiency e only ever changes by constant factors on discrete branches, so it
// carries no derivative; the tree's shape and each body's accept/open decisions are
// likewise piecewise constant. Everything else is differentiated exactly:
//   k_adj_apply    ub = (d step / d u)^T pb, and the step's share of d loss / d s'
//   k_adj_scalar1  d s' / d (S', T', s) on the branch the forward took
//   k_adj_node     swing/traction and gravity terms of pb and ub, gravity's d/d g
//   k_adj_scalar2  the centroid's adjoint
//   k_adj_attr     pb += sum_j w_ij (ub_j - ub_i) (the Laplacian is symmetric), as a CSR
//                  gather; k_adj_wbar: d/d w_e = (p_hi - p_lo).(ub_lo - ub_hi)
//   k_adj_dense    pb_i += sum_j k m_i m_j J(r_ij) (ub_i - ub_j), J(r) = (I - 2 r r^T/|r|^2)/|r|^2
//   k_adj_walk     the Barnes-Hut transpose: the same wave-uniform walk as the forward;
//                  an accepted cell C adds k m_i M_C J ub_i to pb_i and -k m_i J ub_i to
//                  C's accumulator (the cell centre moves with its bodies); a leaf's exact
//                  pairs add both terms, the partner's through the leaf (accept/open
//                  decisions are not symmetric: i may open j's leaf while j accepts a cell
//                  holding i)
//   k_adj_push     pb_j += m_j * (sum of the accumulators of the cells holding j)
// Scaling ratio and gravity adjoints are sums of u's parts against ub.
#include <math.h>

#include <type_traits>

#include "fa2_common.h"
#include "fa2_kernels.h"

namespace gr {

namespace {

// J(r) v with inv = 1/|r|^2: v/|r|^2 - 2 r (r.v)/|r|^4
template <typename T, int D>
__device__ __forceinline__ void jmul(const T *r, const T *v, T inv, T *out) {
    T rv = (T)0;
#pragma unroll
    for (int a = 0; a < D; a++) rv = rv + r[a] * v[a];
    const T c = (T)2 * rv * inv * inv;
#pragma unroll
    for (int a = 0; a < D; a++) out[a] = v[a] * inv - r[a] * c;
}

template <typename T>
__device__ __forceinline__ T recip(T x) {
    return (T)1 / x;
}

}  // namespace

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_load(const FaDev d, const T *planar) {
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (v >= d.n) return;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const T x = planar[(size_t)a * d.n + v];
        ((T *)d.pos[a])[v] = x;
        ((T *)d.pi)[(size_t)v * D + a] = x;
        ((T *)d.pk)[(size_t)v * 4 + a] = x;
    }
}

// ---- exact accumulation: a term v = m * 2^e (m a 53-bit integer) lands as up to three
// 32-bit chunks in the digit words of its place in a 256-bit fixed-point number (lowest bit
// 2^ADJ_BASE). Integer additions commute, so the sum is the same whatever order the waves'
// atomics land in; terms below 2^ADJ_BASE are truncated (toward zero), which is as exact a
// rule as any. Each word takes chunks below 2^32: 2^31 terms cannot overflow it.
__device__ __forceinline__ void exact_add(unsigned long long *acc, double v) {
    if (v == 0.0) return;
    if (!isfinite(v)) {
        atomicAdd(&acc[ADJ_LIMBS - 1], 1ull);
        return;
    }
    int e = 0;
    const double fr = frexp(v, &e);               // v = fr * 2^e, 0.5 <= |fr| < 1
    long long m = (long long)ldexp(fr, 53);       // exact: |m| < 2^53
    int pos = e - 53 - ADJ_BASE;                  // the place of m's lowest bit
    if (pos < 0) {
        if (pos <= -64) return;
        m = m >= 0 ? (long long)((unsigned long long)m >> -pos) : -(long long)((unsigned long long)(-m) >> -pos);
        pos = 0;
        if (m == 0) return;
    }
    const bool neg = m < 0;
    const unsigned long long mag = (unsigned long long)(neg ? -m : m);
    const int limb = pos >> 5, off = pos & 31;
    const unsigned long long lo = mag << off;                        // bits [0, 64)
    const unsigned long long hi = off ? mag >> (64 - off) : 0ull;    // bits [64, 85)
    const unsigned long long c[3] = {lo & 0xFFFFFFFFull, lo >> 32, hi};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        if (c[q] == 0) continue;
        if (limb + q >= ADJ_LIMBS - 1) {  // past the window's top: the sum reads as NaN
            atomicAdd(&acc[ADJ_LIMBS - 1], 1ull);
            continue;
        }
        atomicAdd(&acc[limb + q], neg ? (unsigned long long)(-(long long)c[q]) : c[q]);
    }
}

// The accumulator's value: carries propagated through the digit words, then the number
// read from its top word down (one rounding per word, the same for every run).
__device__ __forceinline__ double exact_value(const unsigned long long *acc) {
    if (acc[ADJ_LIMBS - 1] != 0) return NAN;
    long long carry = 0;
    unsigned long long dig[ADJ_LIMBS - 1];
#pragma unroll
    for (int q = 0; q < ADJ_LIMBS - 1; q++) {
        const long long w = (long long)acc[q] + carry;
        dig[q] = (unsigned long long)w & 0xFFFFFFFFull;
        carry = w >> 32;  // arithmetic shift: floor(w / 2^32)
    }
    double r = (double)carry;
#pragma unroll
    for (int q = ADJ_LIMBS - 2; q >= 0; q--) r = r * 4294967296.0 + (double)dig[q];
    return ldexp(r, ADJ_BASE);
}

template <int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_clear(const FaDev d, const AdjDev a) {
    const size_t q = (size_t)blockIdx.x * GR_BLOCK + threadIdx.x;
    if (q < (size_t)d.st->n_cells * D * ADJ_LIMBS) a.cacc[q] = 0ull;
    if (q < (size_t)d.n * D * ADJ_LIMBS) a.pscat[q] = 0ull;
}

// schedule.rs:62-77 reversed: p' = p + phi(u, s') u with phi = s'/(1 + q), q = sqrt(s' m |u|).
//   d(phi u)/du  = phi I - phi q / (2 (1 + q) |u|^2) u u^T
//   d(phi u)/ds' = u (2 + q) / (2 (1 + q)^2)
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_apply(const FaDev d, const AdjDev a) {
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    const double s1 = a.g1->speed;
    double ds = 0.0;
    if (v < d.n) {
        double u[D], pb[D], nu2 = 0.0, up = 0.0;
#pragma unroll
        for (int x = 0; x < D; x++) {
            u[x] = (double)((const T *)d.u[x])[v];
            pb[x] = a.pb[x][v];
            nu2 += u[x] * u[x];
            up += u[x] * pb[x];
        }
        const double nu = sqrt(nu2);
        const double m = (double)((const T *)d.mass)[v];
        const double q = sqrt(s1 * m * nu);
        const double phi = s1 / (1.0 + q);
        const double coef = nu2 > 0.0 ? -phi * q / (2.0 * (1.0 + q) * nu2) * up : 0.0;
#pragma unroll
        for (int x = 0; x < D; x++) a.ub[x][v] = phi * pb[x] + coef * u[x];
        ds = up * (2.0 + q) / (2.0 * (1.0 + q) * (1.0 + q));
        if (!isfinite(ds)) ds = 0.0;
    }
    __shared__ double lds[4];
    double r[1] = {ds};
    block_reduce<1, 0>(r, lds);
    if (threadIdx.x == 0) a.part[(size_t)blockIdx.x * ADJ_NPART] = r[0];
}

// schedule.rs:33-60 reversed on the branch the forward took. Inputs: the cumulative totals
// S', T' after the step (g1), the speed and efficiency before it (g0).
__global__ void k_adj_scalar1(const FaDev d, const AdjDev a) {
    if (threadIdx.x != 0) return;
    double ds = 0.0;
    for (uint32_t b = 0; b < a.nblk; b++) ds += a.part[(size_t)b * ADJ_NPART];
    AdjScal *sc = a.sc;
    const double sb = sc->sbar + ds;  // d loss / d s'
    const double S = a.g1->swing, Tt = a.g1->traction;
    const double s0 = a.g0->speed;
    const double n = (double)d.n;
    const double jt = d.jitter_tolerance;
    const double opt_jitter = 0.05 * sqrt(n);
    const double min_jitter = sqrt(opt_jitter);
    const double raw = opt_jitter * Tt / (n * n);
    const double other = fmin(10.0, raw);
    double jitter = jt * fmax(min_jitter, other);
    double djdT = (raw < 10.0 && other > min_jitter) ? jt * opt_jitter / (n * n) : 0.0;
    double se = a.g0->speed_efficiency;
    if (S / Tt > 2.0) {
        if (se > 0.05) se *= 0.5;
        if (jt > jitter) {
            jitter = jt;
            djdT = 0.0;
        }
    }
    const double target = S == 0.0 ? INFINITY : jitter * se * Tt / S;
    double dS = 0.0, dT = 0.0, d0 = 0.0;
    if (target - s0 < 0.5 * s0) {  // s' = target
        dS = -target / S;
        dT = jitter * se / S + se * Tt / S * djdT;
    } else {  // s' = 1.5 s
        d0 = 1.5;
    }
    sc->Sbar += sb * dS;
    sc->Tbar += sb * dT;
    sc->sbar = sb * d0;
}

// Swing/traction (S' = S + sum m|p-u|, T' = T + sum m|p+u|/2) and gravity
// (u -= g m (p - mean)/|p - mean|, state.rs:208-251) terms of pb and ub.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_node(const FaDev d, const AdjDev a) {
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    double cs[D], gb = 0.0;
#pragma unroll
    for (int x = 0; x < D; x++) cs[x] = 0.0;
    if (v < d.n) {
        const double Sb = a.sc->Sbar, Tb = a.sc->Tbar;
        const GrGraph &gg = d.gr[0];
        const double m = (double)((const T *)d.mass)[v];
        double p[D], u[D], ub[D], pb[D], nb2 = 0.0, nf2 = 0.0;
#pragma unroll
        for (int x = 0; x < D; x++) {
            p[x] = (double)((const T *)d.pi)[(size_t)v * D + x];
            u[x] = (double)((const T *)d.u[x])[v];
            ub[x] = a.ub[x][v];
            pb[x] = a.pb[x][v];
            nb2 += (p[x] - u[x]) * (p[x] - u[x]);
            nf2 += (p[x] + u[x]) * (p[x] + u[x]);
        }
        const double gs = nb2 > 0.0 ? Sb * m / sqrt(nb2) : 0.0;
        const double gt = nf2 > 0.0 ? Tb * 0.5 * m / sqrt(nf2) : 0.0;
#pragma unroll
        for (int x = 0; x < D; x++) {
            pb[x] += gs * (p[x] - u[x]) + gt * (p[x] + u[x]);
            ub[x] += -gs * (p[x] - u[x]) + gt * (p[x] + u[x]);
        }
        double c[D], nc2 = 0.0;
#pragma unroll
        for (int x = 0; x < D; x++) {
            c[x] = p[x] - gg.mean[x];
            nc2 += c[x] * c[x];
        }
        const double nc = sqrt(nc2);
        if (nc > 0.0) {
            double cu = 0.0;
#pragma unroll
            for (int x = 0; x < D; x++) cu += c[x] / nc * ub[x];
            const double g = d.gravity * m;
#pragma unroll
            for (int x = 0; x < D; x++) {
                const double cb = -g * (ub[x] - c[x] / nc * cu) / nc;
                pb[x] += cb;
                cs[x] = cb;
            }
            gb = -m * cu;
        }
#pragma unroll
        for (int x = 0; x < D; x++) {
            a.ub[x][v] = ub[x];
            a.pb[x][v] = pb[x];
        }
    }
    __shared__ double lds[4 * (D + 1)];
    double r[D + 1];
#pragma unroll
    for (int x = 0; x < D; x++) r[x] = cs[x];
    r[D] = gb;
    block_reduce<D + 1, 0>(r, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int x = 0; x <= D; x++) a.part[(size_t)blockIdx.x * ADJ_NPART + 1 + x] = r[x];
    }
}

// mean = sum p / n: every position gets -(sum of the gravity terms' centre adjoints)/n.
template <int D>
__global__ void k_adj_scalar2(const FaDev d, const AdjDev a) {
    if (threadIdx.x != 0) return;
    double cs[D + 1];
#pragma unroll
    for (int x = 0; x <= D; x++) cs[x] = 0.0;
    for (uint32_t b = 0; b < a.nblk; b++)
#pragma unroll
        for (int x = 0; x <= D; x++) cs[x] += a.part[(size_t)b * ADJ_NPART + 1 + x];
#pragma unroll
    for (int x = 0; x < D; x++) a.sc->mubar[x] = -cs[x];
    a.sc->gbar += cs[D];
}

// state.rs:152-162 reversed as a gather over v's CSR row, plus the centroid's share.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_attr(const FaDev d, const AdjDev a) {
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (v >= d.n) return;
    const double invn = 1.0 / (double)d.n;
    double acc[D], ubv[D];
#pragma unroll
    for (int x = 0; x < D; x++) {
        acc[x] = a.sc->mubar[x] * invn;
        ubv[x] = a.ub[x][v];
    }
    const T *w = (const T *)d.w;
    for (uint32_t e = d.row_off[v]; e < d.row_off[v + 1]; e++) {
        const uint32_t j = d.nbr[e];
        const double we = w ? (double)w[e] : 1.0;
#pragma unroll
        for (int x = 0; x < D; x++) acc[x] += we * (a.ub[x][j] - ubv[x]);
    }
#pragma unroll
    for (int x = 0; x < D; x++) a.pb[x][v] += acc[x];
}

template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_wbar(const FaDev d, const AdjDev a) {
    const uint32_t e = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (e >= a.m) return;
    const uint32_t lo = a.elo[e], hi = a.ehi[e];
    const T *pi = (const T *)d.pi;
    double s = 0.0;
#pragma unroll
    for (int x = 0; x < D; x++)
        s += ((double)pi[(size_t)hi * D + x] - (double)pi[(size_t)lo * D + x]) *
             (a.ub[x][lo] - a.ub[x][hi]);
    a.wb[e] += s;
}

// state.rs:169-203 reversed: every pair of the graph, tiles of partners staged in LDS. A
// coincident pair repels along a constant jiggle, which has no position derivative.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_dense(const FaDev d, const AdjDev a) {
    __shared__ T tp[D][GR_BLOCK];
    __shared__ T tu[D][GR_BLOCK];
    __shared__ T tm[GR_BLOCK];
    const uint32_t n = d.n;
    const uint32_t v = blockIdx.x * GR_BLOCK + threadIdx.x;
    const bool live = v < n;
    const uint32_t vc = live ? v : 0;
    T pv[D], uv[D], acc[D];
#pragma unroll
    for (int x = 0; x < D; x++) {
        pv[x] = ((const T *)d.pos[x])[vc];
        uv[x] = (T)a.ub[x][vc];
        acc[x] = (T)0;
    }
    const T mv = ((const T *)d.mass)[vc];
    const T kr = (T)d.scaling_ratio;
    const uint32_t seed = d.seed, tick = d.gr[0].iter_no;
    T kb = (T)0;
    for (uint32_t j0 = 0; j0 < n; j0 += GR_BLOCK) {
        __syncthreads();
        const uint32_t jj = j0 + threadIdx.x;
        if (jj < n) {
#pragma unroll
            for (int x = 0; x < D; x++) {
                tp[x][threadIdx.x] = ((const T *)d.pos[x])[jj];
                tu[x][threadIdx.x] = (T)a.ub[x][jj];
            }
            tm[threadIdx.x] = ((const T *)d.mass)[jj];
        }
        __syncthreads();
        const uint32_t cnt = n - j0 < GR_BLOCK ? n - j0 : GR_BLOCK;
        if (!live) continue;
        for (uint32_t t = 0; t < cnt; t++) {
            const uint32_t j = j0 + t;
            if (j == v) continue;
            T r[D];
            bool zero = true;
#pragma unroll
            for (int x = 0; x < D; x++) {
                r[x] = pv[x] - tp[x][t];
                if (r[x] != (T)0) zero = false;
            }
            if (zero) {
#pragma unroll
                for (int x = 0; x < D; x++) {
                    const T jg = (T)d_jiggle(seed, tick, (uint32_t)x, v, j);
                    r[x] = v < j ? jg : -jg;
                }
            }
            T d2 = (T)0, ru = (T)0;
#pragma unroll
            for (int x = 0; x < D; x++) {
                d2 = d2 + r[x] * r[x];
                ru = ru + r[x] * uv[x];
            }
            const T inv = recip(d2);
            const T mm = mv * tm[t];
            kb = kb + mm * ru * inv;
            if (zero) continue;
            T vb[D], jv[D];
#pragma unroll
            for (int x = 0; x < D; x++) vb[x] = uv[x] - tu[x][t];
            jmul<T, D>(r, vb, inv, jv);
#pragma unroll
            for (int x = 0; x < D; x++) acc[x] = acc[x] + kr * mm * jv[x];
        }
    }
    if (live) {
#pragma unroll
        for (int x = 0; x < D; x++) a.pb[x][v] += (double)acc[x];
    }
    __shared__ double lds[4];
    double r1[1] = {live ? (double)kb : 0.0};
    block_reduce<1, 0>(r1, lds);
    if (threadIdx.x == 0) a.part[(size_t)blockIdx.x * ADJ_NPART + 6] = r1[0];
}

// barnes_hut.rs:112-184 transposed. The scan is the forward walk's (same wave-uniform
// preorder visit, same accept test on the same tree), so each body meets exactly the cells
// and leaf partners its forward force summed.
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_walk(const FaDev d, const AdjDev a) {
    const uint32_t n = d.n;
    const uint32_t p = blockIdx.x * GR_BLOCK + threadIdx.x;
    const bool live = p < n && p < d.st->n_finite;
    const uint32_t pc = p < n ? p : n - 1;
    const WalkCell<T, D> *__restrict__ wc = (const WalkCell<T, D> *)d.wc;
    const T *__restrict__ sm = (const T *)d.sm;
    const uint32_t *__restrict__ order = d.order;
    const T *sp[3] = {(const T *)d.sp[0], (const T *)d.sp[1],
                      (const T *)(D == 3 ? d.sp[2] : d.sp[0])};
    T xi[D], ui[D], pb[D];
#pragma unroll
    for (int x = 0; x < D; x++) xi[x] = sp[x][pc];
    const T mi = sm[pc];
    const uint32_t i = order[pc];
#pragma unroll
    for (int x = 0; x < D; x++) {
        ui[x] = (T)a.ub[x][i];
        pb[x] = (T)0;
    }
    T kb = (T)0;
    const T theta2 = (T)d.theta2, kr = (T)d.scaling_ratio;
    const GrGraph &gg = d.gr[0];
    const uint32_t seed = d.seed, tick = gg.iter_no;
    const uint32_t lo = gg.cell_lo, hi = gg.cell_hi;
    const uint32_t ncells = d.st->n_cells;
    uint32_t next = (live && lo < hi) ? lo : 0xFFFFFFFFu;
    uint32_t k = next;
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
    while (k < end) {
        k = gr_uniform(k);
        const WalkCell<T, D> cell = wc[k];
        const bool act = next == k;
        bool open = false, took = false;
        T cj[D];
#pragma unroll
        for (int x = 0; x < D; x++) cj[x] = (T)0;
        if (act) {
            const bool holds = cell.start <= p && p < cell.end;
            T dx[D];
#pragma unroll
            for (int x = 0; x < D; x++) {
                dx[x] = xi[x] - cell.c[x];
                if constexpr (!std::is_same<T, double>::value) dx[x] = dx[x] - cell.clo[x];
            }
            T d2 = dx[0] * dx[0] + dx[1] * dx[1];
            if (D == 3) d2 = d2 + dx[D - 1] * dx[D - 1];
            if (!holds && d2 > (T)0 && cell.w2 < open_rhs(d2, theta2)) {
                const T inv = recip(d2);
                T jv[D], ru = (T)0;
                jmul<T, D>(dx, ui, inv, jv);
#pragma unroll
                for (int x = 0; x < D; x++) {
                    ru = ru + dx[x] * ui[x];
                    pb[x] = pb[x] + kr * mi * cell.mass * jv[x];
                    cj[x] = -kr * mi * jv[x];
                }
                kb = kb + mi * cell.mass * ru * inv;
                took = true;
                next = cell.skip;
            } else {
                next = k + 1;
                open = true;
            }
            if (next >= hi) next = 0xFFFFFFFFu;
        }
        const bool leaf = cell.skip == k + 1;
        const bool pairs = open && leaf;
        if (__ballot(pairs) != 0) {
            // A body inside the leaf meets its leaf mates from both sides (every walk opens
            // the leaf holding its own body), so it gathers the pair's two terms. A body
            // outside it opened the leaf alone: it hands the partner's term to the leaf,
            // through the leaf's accumulator when the leaf holds one body, else per body.
            const bool inside = cell.start <= p && p < cell.end;
            const bool single = cell.end - cell.start == 1;
            const uint32_t qend = cell.end < n ? cell.end : n;
            for (uint32_t q = cell.start; q < qend; q++) {
                const uint32_t j = order[q];
                if (!pairs || j == i) continue;
                T dd[D];
                bool zero = true;
#pragma unroll
                for (int x = 0; x < D; x++) {
                    const T xj = sp[x][q];
                    dd[x] = i < j ? xi[x] - xj : -(xj - xi[x]);
                    if (dd[x] != (T)0) zero = false;
                }
                if (zero) {
#pragma unroll
                    for (int x = 0; x < D; x++) {
                        const T jg = (T)d_jiggle(seed, tick, (uint32_t)x, i, j);
                        dd[x] = i < j ? jg : -jg;
                    }
                }
                const T mj = sm[q];
                T d2 = (T)0, ru = (T)0;
#pragma unroll
                for (int x = 0; x < D; x++) {
                    d2 = d2 + dd[x] * dd[x];
                    ru = ru + dd[x] * ui[x];
                }
                const T inv = recip(d2);
                kb = kb + mi * mj * ru * inv;
                if (zero) continue;
                T vb[D], jv[D];
                if (inside) {
#pragma unroll
                    for (int x = 0; x < D; x++) vb[x] = ui[x] - (T)a.ub[x][j];
                    jmul<T, D>(dd, vb, inv, jv);
#pragma unroll
                    for (int x = 0; x < D; x++) pb[x] = pb[x] + kr * mi * mj * jv[x];
                } else {
                    jmul<T, D>(dd, ui, inv, jv);
#pragma unroll
                    for (int x = 0; x < D; x++) {
                        pb[x] = pb[x] + kr * mi * mj * jv[x];
                        if (single) {
                            cj[x] = cj[x] - kr * mi * jv[x];
                        } else {
                            exact_add(&a.pscat[((size_t)q * D + x) * ADJ_LIMBS],
                                      -(double)(kr * mi * mj * jv[x]));
                        }
                    }
                    took = took || single;
                }
            }
        }
        if (__ballot(took) != 0) {
#pragma unroll
            for (int x = 0; x < D; x++) {
                const double s = gr_wave_sum((double)cj[x]);
                if (gr_lane() == 0) exact_add(&a.cacc[((size_t)k * D + x) * ADJ_LIMBS], s);
            }
        }
        k = (__ballot(open) != 0 || cell.skip <= k) ? k + 1 : cell.skip;
    }
    if (live) {
#pragma unroll
        for (int x = 0; x < D; x++) a.pb[x][i] += (double)pb[x];
    }
    __shared__ double lds[4];
    double r1[1] = {live ? (double)kb : 0.0};
    block_reduce<1, 0>(r1, lds);
    if (threadIdx.x == 0) a.part[(size_t)blockIdx.x * ADJ_NPART + 6] = r1[0];
}

// A cell's centre is sum m_j p_j / M over its bodies, so body j takes m_j times the
// accumulators of every cell on its root path (found by descending through the children).
template <typename T, int D>
__global__ __launch_bounds__(GR_BLOCK) void k_adj_push(const FaDev d, const AdjDev a) {
    const uint32_t q = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (q >= d.n || q >= d.st->n_finite) return;
    const WalkCell<T, D> *__restrict__ wc = (const WalkCell<T, D> *)d.wc;
    const GrGraph &gg = d.gr[0];
    if (gg.cell_lo >= gg.cell_hi) return;
    uint32_t k = gg.cell_lo;
    if (!(wc[k].start <= q && q < wc[k].end)) return;
    double s[D];
#pragma unroll
    for (int x = 0; x < D; x++) s[x] = 0.0;
    for (;;) {
#pragma unroll
        for (int x = 0; x < D; x++) s[x] += exact_value(&a.cacc[((size_t)k * D + x) * ADJ_LIMBS]);
        const uint32_t end = wc[k].skip;
        if (end == k + 1) break;  // leaf
        uint32_t c = k + 1;
        while (c < end && !(wc[c].start <= q && q < wc[c].end)) c = wc[c].skip;
        if (c >= end) break;
        k = c;
    }
    const uint32_t j = d.order[q];
    const double mj = (double)((const T *)d.sm)[q];
#pragma unroll
    for (int x = 0; x < D; x++)
        a.pb[x][j] += mj * s[x] + exact_value(&a.pscat[((size_t)q * D + x) * ADJ_LIMBS]);
}

__global__ void k_adj_scalar3(const AdjDev a, uint32_t nblk) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (uint32_t b = 0; b < nblk; b++) s += a.part[(size_t)b * ADJ_NPART + 6];
    a.sc->kbar += s;
}

// ------------------------------------------------------------------ launchers

template <typename T, int D>
static void load_impl(FaDev &d, const void *planar) {
    hipLaunchKernelGGL((k_adj_load<T, D>), dim3(gr_blocks(d.n, GR_BLOCK)), dim3(GR_BLOCK), 0,
                       d.stream, d, (const T *)planar);
}

template <typename T, int D>
static void step_impl(FaDev &d, AdjDev &a) {
    hipStream_t s = d.stream;
    const uint32_t nbl = gr_blocks(d.n, GR_BLOCK);
    a.nblk = nbl;
    hipLaunchKernelGGL((k_adj_apply<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
    hipLaunchKernelGGL(k_adj_scalar1, dim3(1), dim3(64), 0, s, d, a);
    hipLaunchKernelGGL((k_adj_node<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
    hipLaunchKernelGGL((k_adj_scalar2<D>), dim3(1), dim3(64), 0, s, d, a);
    // the repulsion transposes read ub of every body and add into pb; the attraction
    // gather too. They only add to pb, so their order only affects rounding.
    if (d.repulsion == REP_TREE) {
        const uint64_t cl = (uint64_t)(d.cap > d.n ? d.cap : d.n) * D * ADJ_LIMBS;
        const uint32_t nclr = gr_blocks(cl, GR_BLOCK);
        hipLaunchKernelGGL((k_adj_clear<D>), dim3(nclr), dim3(GR_BLOCK), 0, s, d, a);
        hipLaunchKernelGGL((k_adj_walk<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
        hipLaunchKernelGGL((k_adj_push<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
    } else {
        hipLaunchKernelGGL((k_adj_dense<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
    }
    hipLaunchKernelGGL(k_adj_scalar3, dim3(1), dim3(64), 0, s, a, nbl);
    hipLaunchKernelGGL((k_adj_attr<T, D>), dim3(nbl), dim3(GR_BLOCK), 0, s, d, a);
    if (a.m > 0)
        hipLaunchKernelGGL((k_adj_wbar<T, D>), dim3(gr_blocks(a.m, GR_BLOCK)), dim3(GR_BLOCK),
                           0, s, d, a);
}

void launch_adjoint_load(FaDev &d, const void *planar) {
    if (d.precision == PREC_FP64) {
        if (d.dim == 2)
            load_impl<double, 2>(d, planar);
        else
            load_impl<double, 3>(d, planar);
    } else {
        if (d.dim == 2)
            load_impl<float, 2>(d, planar);
        else
            load_impl<float, 3>(d, planar);
    }
}

void launch_adjoint_step(FaDev &d, AdjDev &a) {
    if (d.precision == PREC_FP64) {
        if (d.dim == 2)
            step_impl<double, 2>(d, a);
        else
            step_impl<double, 3>(d, a);
    } else {
        if (d.dim == 2)
            step_impl<float, 2>(d, a);
        else
            step_impl<float, 3>(d, a);
    }
}

}  // namespace gr
