// Host side of the C-ABI (include/graph_render_fa2.h): parameter checks, the simple graph
// and its CSR, start positions, device residency of one Fa2State and the iteration loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/graph_render_fa2.h"
#include "d3.h"
#include "libm_trig.h"
#include "fa2_kernels.h"
#include "sort_scan.h"

namespace {

struct LastError {
    gr_status code = GR_OK;
    std::string name, rule;
};
thread_local LastError g_err;

gr_status fail(gr_status code, const char *name, const char *rule) {
    g_err.code = code;
    g_err.name = name ? name : "";
    g_err.rule = rule ? rule : "";
    return code;
}

gr_status ok() {
    g_err = LastError();
    return GR_OK;
}

#define HIP_TRY(x)                                                                         \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(GR_ERR_DEVICE, #x, hipGetErrorString(e_));       \
    } while (0)

// synthetic.rs:44-50 (Mulberry32), the generator of the reference's start positions.
double mulberry32_next(uint32_t &st) {
    st += 0x6D2B79F5u;
    uint32_t a = st;
    uint32_t t = (a ^ (a >> 15)) * (1u | a);
    t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
    return (double)(t ^ (t >> 14)) / 4294967296.0;
}

uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

struct SimpleGraph {
    std::vector<uint32_t> lo, hi;
    std::vector<size_t> raw;  // the raw edge each simple edge keeps (its first occurrence)
    std::vector<double> w;
    std::vector<uint32_t> row_off, nbr;
    std::vector<double> roww;
};

// layout/force/mod.rs:101-120 and :177-180: self-loops and repeated unordered pairs
// dropped (first raw edge kept), then each node's row lists its simple edges in ascending
// id (the stable counting sort of csr.rs:57-95), stored as neighbour ids.
void build_simple(uint32_t n, const uint32_t *src, const uint32_t *dst, size_t m,
                  const double *weights, SimpleGraph &g) {
    size_t cap = 16;
    while (cap < 2 * m + 2) cap <<= 1;
    std::vector<uint64_t> table(cap, ~0ull);
    g.lo.reserve(m);
    g.hi.reserve(m);
    if (weights) g.w.reserve(m);
    for (size_t e = 0; e < m; e++) {
        uint32_t a = src[e], b = dst[e];
        if (a == b) continue;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        uint64_t key = ((uint64_t)lo << 32) | hi;
        size_t h = (size_t)(mix64(key) & (cap - 1));
        bool dup = false;
        while (table[h] != ~0ull) {
            if (table[h] == key) {
                dup = true;
                break;
            }
            h = (h + 1) & (cap - 1);
        }
        if (dup) continue;
        table[h] = key;
        g.lo.push_back(lo);
        g.hi.push_back(hi);
        g.raw.push_back(e);
        if (weights) g.w.push_back(weights[e]);
    }
    size_t ms = g.lo.size();
    g.row_off.assign((size_t)n + 1, 0);
    for (size_t e = 0; e < ms; e++) {
        g.row_off[g.lo[e] + 1]++;
        g.row_off[g.hi[e] + 1]++;
    }
    for (size_t v = 0; v < n; v++) g.row_off[v + 1] += g.row_off[v];
    std::vector<uint32_t> cur(g.row_off.begin(), g.row_off.end() - 1);
    g.nbr.resize(2 * ms);
    if (weights) g.roww.resize(2 * ms);
    for (size_t e = 0; e < ms; e++) {
        uint32_t a = g.lo[e], b = g.hi[e];
        uint32_t ia = cur[a]++, ib = cur[b]++;
        g.nbr[ia] = b;
        g.nbr[ib] = a;
        if (weights) {
            g.roww[ia] = g.w[e];
            g.roww[ib] = g.w[e];
        }
    }
}

bool schedule_valid(const gr_fa2_schedule &sc) {
    const bool wb = sc.walk_block == 64 || sc.walk_block == 128 || sc.walk_block == 256;
    const bool si = sc.sort_items == 4 || sc.sort_items == 8 || sc.sort_items == 16;
    const bool fs = sc.finish_split == 1 || sc.finish_split == 2 || sc.finish_split == 4 ||
                    sc.finish_split == 8;
    return wb && si && fs && sc.walk_xcd <= 1;
}

// A session's size class: log2 of its node count, its edges per node (rounded down to a
// power of two), dimension, precision, and whether it is a batch.
using SizeClass = std::tuple<int, int, uint32_t, int, int>;
SizeClass size_class(uint32_t n, size_t m, uint32_t dim, int precision, uint32_t ngraphs) {
    const int ln = n > 1 ? (int)std::floor(std::log2((double)n)) : 0;
    const double per = n ? (double)m / n : 0.0;
    const int lm = per >= 1.0 ? (int)std::floor(std::log2(per)) : 0;
    return SizeClass{ln, lm, dim, precision, ngraphs > 1 ? 1 : 0};
}

std::mutex g_sched_lock;
std::map<SizeClass, gr_fa2_schedule> g_sched_cache;  // autotuned schedules of this process

}  // namespace

struct gr_fa2_session {
    gr::FaDev d;
    gr_fa2_params params{};
    uint32_t m_simple = 0;
    std::vector<uint32_t> lo, hi;
    std::vector<size_t> simple_raw;
    size_t m_raw = 0;
    std::vector<double> mass;
    std::vector<uint32_t> row_off;  // the CSR row offsets (the finish split is cut from them)
    uint32_t part_cap = 0;          // k_finish workgroups d.part_rows has room for
    gr_fa2_schedule sched{256, 1, 8, 1};
    uint32_t size_key_n = 0;  // the sizes the session's schedule class is keyed on
    size_t size_key_m = 0;
    std::vector<void *> allocs;
    int device = 0;
    bool tree_valid = false;
    // One iteration captured as a HIP graph and replayed; rebuilt after anything that
    // changes the buffers its launches were captured with (a capacity recovery).
    bool use_graph = true;
    hipGraph_t it_graph = nullptr;
    hipGraphExec_t it_exec = nullptr;

    void drop_graph() {
        if (it_exec) (void)hipGraphExecDestroy(it_exec);
        if (it_graph) (void)hipGraphDestroy(it_graph);
        it_exec = nullptr;
        it_graph = nullptr;
    }

    // Enqueues one iteration: the captured graph when there is one, else direct launches.
    // Returns the launch's status (a direct launch's errors surface in hipGetLastError).
    hipError_t iteration() {
        if (use_graph && !it_exec) {
            hipGraph_t g = nullptr;
            if (hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                gr::launch_iteration(d);
                if (hipStreamEndCapture(d.stream, &g) == hipSuccess && g) {
                    if (hipGraphInstantiate(&it_exec, g, nullptr, nullptr, 0) == hipSuccess)
                        it_graph = g;
                    else
                        (void)hipGraphDestroy(g);
                }
            }
            (void)hipGetLastError();
            if (!it_exec) use_graph = false;  // capture unavailable: launch directly
        }
        if (it_exec) return hipGraphLaunch(it_exec, d.stream);
        gr::launch_iteration(d);
        return hipSuccess;
    }

    // The last gr_fa2_session_record(): state checkpoints at the start of each segment of
    // `every` iterations, and the state it ended in.
    struct Recording {
        bool valid = false;
        uint32_t every = 0, nseg = 0;
        std::vector<uint32_t> seg_len;  // iterations applied in each segment
        void *ck_pos = nullptr;         // nseg snapshots of the planar positions
        GrGraph *ck_g = nullptr;
        GrState *ck_st = nullptr;
        void *fin_pos = nullptr;
        GrGraph *fin_g = nullptr;
        GrState *fin_st = nullptr;
        std::vector<void *> allocs;
    } rec;

    ~gr_fa2_session() {
        drop_recording();
        drop_graph();
        for (void *p : allocs) (void)hipFree(p);
        if (d.stream) (void)hipStreamDestroy(d.stream);
    }

    void drop_recording() {
        if (d.stream) (void)hipStreamSynchronize(d.stream);
        for (void *p : rec.allocs) (void)hipFree(p);
        rec = Recording();
    }

    template <typename P>
    gr_status rec_alloc(P **out, size_t bytes) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
        if (e != hipSuccess) return fail(GR_ERR_DEVICE, "hipMalloc", hipGetErrorString(e));
        rec.allocs.push_back(p);
        *out = (P *)p;
        return GR_OK;
    }

    size_t pos_bytes() const { return (size_t)d.n * d.dim * esize(); }

    // k_finish's row split: workgroup b takes the rows from the first one whose
    // (first entry + row index) reaches b/nparts of (entries + rows), so that rows and
    // entries are balanced together; a workgroup gets at most one LDS staging round of rows +
    // entries, or 1/split of one (more, smaller workgroups).
    static uint32_t finish_parts(uint32_t n, uint32_t m2, uint32_t es, uint32_t dim,
                                 bool weighted, uint32_t split) {
        const uint64_t units = (uint64_t)m2 + n;
        const uint32_t per = gr::fin_stage(es, dim, weighted) / (split ? split : 1);
        return gr_xcd_grid(gr_blocks(units, per > 0 ? per : 1));
    }
    // Installs a schedule: the walk's launch shape, the sort's tile and the finish split.
    // The captured iteration holds the old launch parameters, so it is rebuilt.
    gr_status apply_schedule(const gr_fa2_schedule &sc) {
        gr_status r = set_finish_split(sc.finish_split);
        if (r) return r;
        d.walk_block = sc.walk_block;
        d.walk_xcd = sc.walk_xcd ? 1 : 0;
        d.sort_items = sc.sort_items;
        sched = sc;
        drop_graph();
        return GR_OK;
    }
    gr_status set_finish_split(uint32_t split) {
        const uint32_t n = d.n;
        const uint64_t units = (uint64_t)d.m2 + n;
        const uint32_t parts = finish_parts(n, d.m2, (uint32_t)esize(), d.dim, d.weighted != 0, split);
        if (parts > part_cap) return fail(GR_ERR_PARAM, "finish_split", "1, 2, 4 or 8");
        std::vector<uint32_t> part_rows((size_t)parts + 1, n);
        uint32_t r = 0;
        for (uint32_t b = 0; b < parts; b++) {
            const uint64_t cut = units * b / parts;
            while (r < n && (uint64_t)row_off[r] + r < cut) r++;
            part_rows[b] = r;
        }
        part_rows[0] = 0;
        part_rows[parts] = n;
        HIP_TRY(hipStreamSynchronize(d.stream));
        HIP_TRY(gr_copy_wait(d.part_rows, part_rows.data(), part_rows.size() * 4,
                             hipMemcpyHostToDevice, d.stream));
        d.nparts = parts;
        return GR_OK;
    }

    // state -> snapshot (planar positions, graph scalars, session bookkeeping)
    gr_status save_state(void *pos, GrGraph *g, GrState *st) {
        const size_t col = (size_t)d.n * esize();
        for (uint32_t a = 0; a < d.dim; a++)
            HIP_TRY(hipMemcpyAsync((char *)pos + a * col, d.pos[a], col,
                                   hipMemcpyDeviceToDevice, d.stream));
        HIP_TRY(hipMemcpyAsync(g, d.gr, sizeof(GrGraph), hipMemcpyDeviceToDevice, d.stream));
        HIP_TRY(hipMemcpyAsync(st, d.st, sizeof(GrState), hipMemcpyDeviceToDevice, d.stream));
        return GR_OK;
    }

    gr_status load_state(const void *pos, const GrGraph *g, const GrState *st) {
        gr::launch_adjoint_load(d, pos);
        HIP_TRY(hipMemcpyAsync(d.gr, g, sizeof(GrGraph), hipMemcpyDeviceToDevice, d.stream));
        HIP_TRY(hipMemcpyAsync(d.st, st, sizeof(GrState), hipMemcpyDeviceToDevice, d.stream));
        gr::launch_bounds(d);
        tree_valid = false;
        return GR_OK;
    }

    template <typename P>
    gr_status alloc(P **out, size_t bytes) {
        void *p = nullptr;
        if (bytes == 0) bytes = 8;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) return fail(GR_ERR_DEVICE, "hipMalloc", hipGetErrorString(e));
        allocs.push_back(p);
        *out = (P *)p;
        return GR_OK;
    }

    size_t esize() const { return d.precision == gr::PREC_FP64 ? 8 : 4; }

    gr_status alloc_cells(uint32_t cap) {
        d.cap = cap;
        gr_status s;
        // one spare record: the walk may read the record just past the scanned cells
        if ((s = alloc(&d.wc, ((size_t)cap + 1) * gr::walkcell_bytes(d.precision, d.dim))))
            return s;
        if ((s = alloc(&d.clevel, (size_t)cap))) return s;
        const size_t tiles = (cap + 4095) / 4096;
        if ((s = alloc(&d.span, tiles * 32 * 4))) return s;
        if ((s = alloc(&d.span_n, tiles * 4))) return s;
        if ((s = alloc(&d.span_lev, tiles * 32))) return s;
        if ((s = alloc(&d.span_cells, tiles * 32 * 4))) return s;
        d.span_cap = (uint32_t)(tiles * 32);
        if ((s = alloc(&d.span_kids, 2 * tiles * 32 * ((size_t)1 << d.dim) * 4))) return s;
        if ((s = alloc(&d.span_slot, (size_t)cap * 4))) return s;
        if ((s = alloc(&d.span_kval, tiles * 32 * ((size_t)1 << d.dim) * (1 + d.dim) * 8)))
            return s;
        if ((s = alloc(&d.span_lcnt, 128 * 4))) return s;
        if ((s = alloc(&d.raw, (size_t)cap * (1 + d.dim) * 8))) return s;
        if (d.export_bounds && (s = alloc(&d.cbounds, (size_t)cap * 2 * d.dim * 8))) return s;
        return GR_OK;
    }

    gr_status upload_positions(const double *rowmajor) {
        const uint32_t n = d.n, dim = d.dim;
        for (uint32_t a = 0; a < dim; a++) {
            if (d.precision == gr::PREC_FP64) {
                std::vector<double> col(n);
                for (uint32_t i = 0; i < n; i++) col[i] = rowmajor[(size_t)i * dim + a];
                HIP_TRY(gr_copy_wait(d.pos[a], col.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                                     d.stream));
            } else {
                std::vector<float> col(n);
                for (uint32_t i = 0; i < n; i++) col[i] = (float)rowmajor[(size_t)i * dim + a];
                HIP_TRY(gr_copy_wait(d.pos[a], col.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                                     d.stream));
            }
        }
        gr::launch_bounds(d);
        HIP_TRY(hipStreamSynchronize(d.stream));
        tree_valid = false;
        return GR_OK;
    }

    gr_status read_state(GrState &st) {
        HIP_TRY(hipMemcpyAsync(&st, d.st, sizeof(GrState), hipMemcpyDeviceToHost, d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        return GR_OK;
    }

    // After a halted iteration: grow what ran out, clear the flag. The halted iteration
    // applied nothing, so it is simply run again.
    gr_status recover(const GrState &st) {
        bool grew = false;
        if (st.pool_overflow) {
            grew = true;
            uint32_t grow = d.pool_cap * 2;
            void *p1 = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr;
            HIP_TRY(hipMalloc(&p1, (size_t)grow * 4));
            HIP_TRY(hipMalloc(&p2, (size_t)d.lcell_cap * 2 * 16));
            HIP_TRY(hipMalloc(&p3, (size_t)d.lcell_cap * 2 * 2 * d.dim * 8));
            HIP_TRY(hipMalloc(&p4, (size_t)d.lcell_cap * 2 * (1 + d.dim) * 8));
            allocs.push_back(p1);
            allocs.push_back(p2);
            allocs.push_back(p3);
            allocs.push_back(p4);
            d.pool = (int32_t *)p1;
            d.lc_int = (int32_t *)p2;
            d.lc_dbl = (double *)p3;
            d.lc_sum = (double *)p4;
            d.pool_cap = grow;
            d.lcell_cap *= 2;
        }
        if (st.need_cells > d.cap) {
            grew = true;
            uint32_t want = st.need_cells + st.need_cells / 4 + 64;
            gr_status s = alloc_cells(want);
            if (s) return s;
        }
        if (!grew) return fail(GR_ERR_DEVICE, "halt", "iteration halted without a capacity cause");
        drop_graph();  // its launches hold the old buffers
        gr::launch_clear_halt(d);
        return GR_OK;
    }
};

// The median of three timings of `reps` launches (HIP events on the stream they go to).
template <typename F>
static gr_status time_median(hipStream_t stream, uint32_t reps, F launch, double *ms) {
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double t[3];
    for (int k = 0; k < 3; k++) {
        HIP_TRY(hipEventRecord(e0, stream));
        for (uint32_t q = 0; q < reps; q++) launch();
        HIP_TRY(hipEventRecord(e1, stream));
        HIP_TRY(hipEventSynchronize(e1));
        float f = 0.f;
        HIP_TRY(hipEventElapsedTime(&f, e0, e1));
        t[k] = (double)f / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::sort(t, t + 3);
    *ms = t[1];
    HIP_TRY(hipGetLastError());
    return GR_OK;
}

// ... on an FA2 session's stream
template <typename F>
static gr_status time_median(gr_fa2_session *s, uint32_t reps, F launch, double *ms) {
    return time_median(s->d.stream, reps, launch, ms);
}

// The simple graph (layout/force/mod.rs:70-180) for another translation unit of this
// library: self-loops and repeated unordered pairs dropped, (lo, hi) in first-occurrence order.
__attribute__((visibility("hidden"))) void gr_internal_simple_edges(
    uint32_t n, const uint32_t *src, const uint32_t *dst, size_t m, std::vector<uint32_t> &lo,
    std::vector<uint32_t> &hi) {
    SimpleGraph g;
    build_simple(n, src, dst, m, nullptr, g);
    lo.swap(g.lo);
    hi.swap(g.hi);
}

extern "C" {

// Sets the calling thread's last error from another translation unit of this library.
__attribute__((visibility("hidden"))) gr_status gr_internal_fail(gr_status code,
                                                                 const char *name,
                                                                 const char *rule) {
    return code == GR_OK ? ok() : fail(code, name, rule);
}

void gr_fa2_params_default(gr_fa2_params *out) {
    out->max_iter = 100;
    out->jitter_tolerance = 1.0;
    out->scaling_ratio = 2.0;
    out->gravity = 1.0;
    out->seed = 0;
    out->dim = 2;
}

void gr_fa2_ext_default(gr_fa2_ext *out) {
    std::memset(out, 0, sizeof(*out));
    out->struct_size = sizeof(gr_fa2_ext);
    out->theta = -1.0;
    out->precision = GR_PREC_FP64;
    out->repulsion = GR_REP_DEFAULT;
}

// registry/params.rs:144-202 with graph-contract params.rs:147-196's validation.
gr_status gr_fa2_params_decode(const uint8_t *bytes, size_t len, gr_fa2_params *out) {
    gr_fa2_params_default(out);
    if (len == 0) return ok();
    struct Spec {
        const char *name;
        bool integer;
        double min, max;
    };
    static const Spec specs[5] = {{"max_iter", true, 1.0, 10000.0},
                                  {"jitter_tolerance", false, 0.0, 1e6},
                                  {"scaling_ratio", false, 0.0, 1e6},
                                  {"gravity", false, 0.0, 1e6},
                                  {"seed", true, 0.0, 4294967295.0}};
    if (!bytes || len != 5 * sizeof(double))
        return fail(GR_ERR_PARAMS_MALFORMED, "params",
                    "one little-endian f64 per published parameter, and no other bytes");
    double v[5];
    for (int i = 0; i < 5; i++) {
        uint64_t bits = 0;
        for (int b = 0; b < 8; b++) bits |= (uint64_t)bytes[i * 8 + b] << (8 * b);
        std::memcpy(&v[i], &bits, 8);
        const Spec &s = specs[i];
        if (!std::isfinite(v[i]) || v[i] < s.min || v[i] > s.max ||
            (s.integer && v[i] != std::floor(v[i])))
            return fail(GR_ERR_PARAM_OUT_OF_RANGE, s.name,
                        "inside the range the schema publishes, never clamped");
    }
    out->max_iter = (uint32_t)v[0];
    out->jitter_tolerance = v[1];
    out->scaling_ratio = v[2];
    out->gravity = v[3];
    out->seed = (uint32_t)v[4];
    return ok();
}

// A session over `ngraphs` independent graphs laid out one after another: graph b has
// node_counts[b] nodes and edge_counts[b] raw edges whose endpoints are graph-local. One graph
// is the reference's Stage; many are the batched layout, all in the same launches.
static gr_fa2_session *create_impl(uint32_t ngraphs, const uint32_t *node_counts,
                                   const size_t *edge_counts, const uint32_t *source,
                                   const uint32_t *target, const gr_fa2_params *params,
                                   const gr_fa2_ext *ext) {
    if (ngraphs == 0 || !node_counts || !edge_counts) {
        fail(GR_ERR_PARAM, "graphs", "at least one graph with its node and edge counts");
        return nullptr;
    }
    uint64_t n64 = 0, m64 = 0;
    for (uint32_t b = 0; b < ngraphs; b++) {
        n64 += node_counts[b];
        m64 += edge_counts[b];
    }
    if (n64 >= 0xFFFFFFFFull) {  // index.rs: the u32 index space
        fail(GR_ERR_CAPACITY, "nodes", "node count fits the u32 index space");
        return nullptr;
    }
    const uint32_t n = (uint32_t)n64;
    const size_t m = (size_t)m64;
    gr_fa2_params p;
    if (params)
        p = *params;
    else
        gr_fa2_params_default(&p);
    gr_fa2_ext x;
    gr_fa2_ext_default(&x);
    if (ext) {
        if (ext->struct_size != sizeof(gr_fa2_ext)) {
            fail(GR_ERR_PARAM, "ext", "struct_size must equal sizeof(gr_fa2_ext)");
            return nullptr;
        }
        x = *ext;
    }
    if (p.dim != 2 && p.dim != 3) {  // layout/forceatlas2.rs:70-76
        fail(GR_ERR_PARAM, "dim", "2 or 3 coordinates per node");
        return nullptr;
    }
    if (m > 0xFFFFFFFFull / 2) {
        fail(GR_ERR_CAPACITY, "edges", "edge count fits the u32 index space");
        return nullptr;
    }
    if (m > 0 && (!source || !target)) {
        fail(GR_ERR_PARAM, "edges", "source and target columns are required");
        return nullptr;
    }
    // graph-local endpoints, checked, then offset into the session's node space
    std::vector<uint32_t> gsrc, gdst;
    const uint32_t *src_g = source, *dst_g = target;
    {
        size_t e = 0;
        uint32_t off = 0;
        if (ngraphs > 1) {
            gsrc.resize(m);
            gdst.resize(m);
        }
        for (uint32_t b = 0; b < ngraphs; b++) {
            for (size_t q = 0; q < edge_counts[b]; q++, e++) {
                if (source[e] >= node_counts[b] || target[e] >= node_counts[b]) {
                    fail(GR_ERR_PARAM, "edges", "every endpoint is a node index below n");
                    return nullptr;
                }
                if (ngraphs > 1) {
                    gsrc[e] = source[e] + off;
                    gdst[e] = target[e] + off;
                }
            }
            off += node_counts[b];
        }
        if (ngraphs > 1) {
            src_g = gsrc.data();
            dst_g = gdst.data();
        }
    }
    if (x.precision < GR_PREC_FP64 || x.precision > GR_PREC_FP32_KAHAN) {
        fail(GR_ERR_PARAM, "precision", "GR_PREC_FP64, GR_PREC_FP32 or GR_PREC_FP32_KAHAN");
        return nullptr;
    }
    int rep = x.repulsion == GR_REP_DEFAULT ? GR_REP_TREE : x.repulsion;
    if (rep < GR_REP_TREE || rep > GR_REP_PM) {
        fail(GR_ERR_PARAM, "repulsion", "GR_REP_DEFAULT, TREE, DENSE or PM");
        return nullptr;
    }
    double theta = x.theta < 0.0 ? 0.8 : x.theta;
    if (!std::isfinite(theta)) {
        fail(GR_ERR_PARAM, "theta", "a finite opening angle");
        return nullptr;
    }
    if (x.edge_weights)
        for (size_t e = 0; e < m; e++)
            if (!std::isfinite(x.edge_weights[e])) {
                fail(GR_ERR_NON_FINITE, "edge.weight", "finite edge weights");
                return nullptr;
            }
    if (rep == GR_REP_PM) {
        uint32_t side = x.pm_grid ? x.pm_grid : 2048;
        if (side < 64 || side > 4096 || (side & (side - 1))) {
            fail(GR_ERR_PARAM, "pm_grid", "a power of two from 64 to 4096");
            return nullptr;
        }
        if (p.dim != 2 || ngraphs != 1) {
            fail(GR_ERR_UNSUPPORTED, "pm_grid", "the particle mesh runs on one 2D graph");
            return nullptr;
        }
        x.pm_grid = side;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        fail(GR_ERR_DEVICE, "device", "no HIP device is visible");
        return nullptr;
    }

    auto *s = new gr_fa2_session();
    s->params = p;
    gr::FaDev &d = s->d;
    d.n = n;
    d.dim = p.dim;
    d.precision = x.precision;
    d.repulsion = rep == GR_REP_TREE ? gr::REP_TREE : rep == GR_REP_DENSE ? gr::REP_DENSE
                                                                           : gr::REP_PM;
    d.theta2 = theta * theta;
    d.scaling_ratio = p.scaling_ratio;
    d.gravity = p.gravity;
    d.jitter_tolerance = p.jitter_tolerance;
    d.seed = p.seed;
    d.export_bounds = (x.flags & GR_EXT_EXPORT_TREE) ? 1 : 0;
    d.seq_sums = (x.precision == GR_PREC_FP64 && !(x.flags & GR_EXT_BLOCK_SUMS)) ? 1 : 0;
    s->use_graph = (x.flags & GR_EXT_DIRECT_LAUNCH) == 0;
    d.pm_side = x.pm_grid;
    (void)hipGetDevice(&s->device);
    if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess) {
        fail(GR_ERR_DEVICE, "hipStreamCreate", "stream creation failed");
        delete s;
        return nullptr;
    }

    SimpleGraph g;
    build_simple(n, src_g, dst_g, m, x.edge_weights, g);
    s->m_simple = (uint32_t)g.lo.size();
    s->lo = g.lo;
    s->hi = g.hi;
    s->simple_raw = g.raw;
    s->m_raw = m;
    s->mass.resize(n);
    for (uint32_t v = 0; v < n; v++) s->mass[v] = (double)(g.row_off[v + 1] - g.row_off[v]) + 1.0;
    d.m2 = (uint32_t)g.nbr.size();
    d.weighted = x.edge_weights ? 1 : 0;
    std::vector<uint32_t> long_rows;
    for (uint32_t v = 0; v < n; v++)
        if (g.row_off[v + 1] - g.row_off[v] > gr::LONG_ROW) long_rows.push_back(v);
    d.nlong = (uint32_t)long_rows.size();

    const size_t es = s->esize();
    const size_t nn = n ? n : 1;
    gr_status st;
#define TRY_ALLOC(ptr, bytes)                      \
    if ((st = s->alloc(&(ptr), (bytes))) != GR_OK) { \
        delete s;                                  \
        return nullptr;                            \
    }
    for (uint32_t a = 0; a < d.dim; a++) {
        TRY_ALLOC(d.pos[a], nn * es);
        TRY_ALLOC(d.u[a], nn * es);
        TRY_ALLOC(d.rep[a], nn * es);
        TRY_ALLOC(d.sp[a], nn * es);
    }
    TRY_ALLOC(d.pi, nn * d.dim * es);
    TRY_ALLOC(d.pk, nn * 4 * es);
    TRY_ALLOC(d.mass, nn * es);
    TRY_ALLOC(d.sm, nn * es);
    TRY_ALLOC(d.row_off, (nn + 1) * 4);
    TRY_ALLOC(d.nbr, (g.nbr.size() + 1) * 4);
    if (d.weighted) TRY_ALLOC(d.w, (g.nbr.size() + 1) * es);
    TRY_ALLOC(d.long_rows, ((size_t)d.nlong + 1) * 4);
    TRY_ALLOC(d.long_acc, ((size_t)d.nlong + 1) * d.dim * es);
    TRY_ALLOC(d.st, sizeof(GrState));
    TRY_ALLOC(d.wstat, 4 * sizeof(uint64_t));
    // batch tables: node offsets, graph-aligned node chunks, node -> graph
    d.ngraphs = ngraphs;
    {
        // The sort takes whole digits of radix_bits() bits. The key holds the fewest digits
        // (at least 30 bits) that still give WANT levels (the flag and graph id included);
        // points sharing all of them are separated by the deep-group builder, which replays
        // the reference insertion below the last key level, so the tree is the same for any
        // level count.
        const uint32_t gbits = ngraphs > 1 ? (uint32_t)std::ceil(std::log2((double)ngraphs)) : 0;
        const uint32_t maxl = d.dim == 2 ? 31 : 21;
        uint32_t biggest = 1;
        for (uint32_t b = 0; b < ngraphs; b++) biggest = std::max(biggest, node_counts[b]);
        // levels to spare beyond the depth an even spread of the biggest graph fills
        const uint32_t fill = (uint32_t)std::ceil(std::log2((double)biggest) / d.dim);
        const uint32_t want = fill + (d.dim == 2 ? 8 : 5);
        const uint32_t rb = (uint32_t)gr::radix_bits(), max_digits = 64 / rb;
        uint32_t digits = (30 + rb - 1) / rb, levels = 0;
        for (;; digits++) {
            levels = (rb * digits - 1 - gbits) / d.dim;
            if (levels >= want || levels >= maxl || digits == max_digits) break;
        }
        d.levels = levels < maxl ? levels : maxl;
        d.key_bits = d.dim * d.levels + gbits + 1;
    }
    std::vector<uint32_t> gnode(ngraphs + 1, 0), gchunk(ngraphs + 1, 0), chunk(1, 0), chunk_g;
    const uint32_t CHUNK = 2048;
    for (uint32_t b = 0; b < ngraphs; b++) {
        gnode[b + 1] = gnode[b] + node_counts[b];
        for (uint32_t v = gnode[b]; v < gnode[b + 1]; v += CHUNK) {
            chunk.push_back(std::min(v + CHUNK, gnode[b + 1]));
            chunk_g.push_back(b);
        }
        gchunk[b + 1] = (uint32_t)chunk_g.size();
    }
    if (chunk_g.empty()) {  // no nodes at all: one empty chunk keeps every grid non-empty
        chunk.push_back(0);
        chunk_g.push_back(0);
        gchunk[ngraphs] = 1;
    }
    d.nchunks = (uint32_t)chunk_g.size();
    TRY_ALLOC(d.gr, (size_t)ngraphs * sizeof(GrGraph));
    TRY_ALLOC(d.gnode, (ngraphs + 1) * 4);
    TRY_ALLOC(d.gchunk, (ngraphs + 1) * 4);
    TRY_ALLOC(d.chunk, (d.nchunks + 1) * 4);
    TRY_ALLOC(d.chunk_g, d.nchunks * 4);
    TRY_ALLOC(d.nodeval, nn * 2 * 8);
    if (ngraphs > 1) TRY_ALLOC(d.gidn, nn * 4);
    TRY_ALLOC(d.keys, nn * 8);
    TRY_ALLOC(d.keys_alt, nn * 8);
    TRY_ALLOC(d.vals, nn * 4);
    TRY_ALLOC(d.vals_alt, nn * 4);
    TRY_ALLOC(d.stall, nn);
    TRY_ALLOC(d.order, nn * 4);
    TRY_ALLOC(d.leaf, nn * 4);
    TRY_ALLOC(d.flag, nn * 4);
    TRY_ALLOC(d.gid, nn * 4);
    TRY_ALLOC(d.dflag, nn * 4);
    TRY_ALLOC(d.gstart, nn * 4);
    TRY_ALLOC(d.gcells, nn * 4);
    TRY_ALLOC(d.gbase, nn * 4);
    TRY_ALLOC(d.ginfo, nn * 4);
    TRY_ALLOC(d.glocal, nn * 4);
    TRY_ALLOC(d.gkey, nn * 8);
    TRY_ALLOC(d.chain_next, nn * 4);
    d.pool_cap = 1u << 22;
    d.lcell_cap = 1u << 16;
    TRY_ALLOC(d.pool, (size_t)d.pool_cap * 4);
    TRY_ALLOC(d.lc_int, (size_t)d.lcell_cap * 16);
    TRY_ALLOC(d.lc_dbl, (size_t)d.lcell_cap * 2 * d.dim * 8);
    TRY_ALLOC(d.lc_sum, (size_t)d.lcell_cap * (1 + d.dim) * 8);
    s->row_off = g.row_off;
    s->part_cap = gr_fa2_session::finish_parts(n, d.m2, (uint32_t)es, d.dim, d.weighted != 0, 8);
    TRY_ALLOC(d.part_rows, ((size_t)s->part_cap + 1) * 4);
    TRY_ALLOC(d.bpart, (size_t)d.nchunks * 10 * 8);
    TRY_ALLOC(d.spart, (size_t)d.nchunks * 2 * 8);
    TRY_ALLOC(d.mpart, (size_t)d.nchunks * 8);
    TRY_ALLOC(d.scan_tmp, gr::scan_tmp_bytes((uint32_t)nn));
    TRY_ALLOC(d.sort_tmp, gr::radix_tmp_bytes((uint32_t)nn));
#undef TRY_ALLOC
#define TRY_ALLOC2(ptr, bytes)                     \
    if ((st = s->alloc(&(ptr), (bytes))) != GR_OK) { \
        delete s;                                  \
        return nullptr;                            \
    }
    if (s->alloc_cells(2 * (uint32_t)nn + 64) != GR_OK) {
        delete s;
        return nullptr;
    }
    if (d.repulsion == gr::REP_PM) {
        const size_t P2 = (size_t)d.pm_side * d.pm_side;
        TRY_ALLOC2(d.pm_a, P2 * sizeof(float2));
        TRY_ALLOC2(d.pm_b, P2 * sizeof(float2));
        TRY_ALLOC2(d.pm_k, P2 * sizeof(float2));
        TRY_ALLOC2(d.pm_acc, P2 * sizeof(unsigned long long));
        TRY_ALLOC2(d.pm_tw, (d.pm_side / 2) * sizeof(float2));
        TRY_ALLOC2(d.pm_tiles, ((size_t)(d.pm_side / 16) * (d.pm_side / 16) + 2) * 4);
        TRY_ALLOC2(d.pm_halo, (size_t)(d.pm_side / 16) * (d.pm_side / 16) * 33 * 8);
        TRY_ALLOC2(d.pm_fr, gr::pm_frame_bytes());
        std::vector<float2> tw(d.pm_side / 2);
        gr::pm_twiddles(d.pm_side, tw.data());
        std::vector<uint8_t> zero(gr::pm_frame_bytes(), 0);
        if (gr_copy_wait(d.pm_tw, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice,
                         d.stream) != hipSuccess ||
            gr_copy_wait(d.pm_fr, zero.data(), zero.size(), hipMemcpyHostToDevice,
                         d.stream) != hipSuccess ||
            gr_fill_wait(d.pm_a, 0, P2 * sizeof(float2), d.stream) != hipSuccess) {
            fail(GR_ERR_DEVICE, "hipMemcpy", "particle-mesh setup failed");
            delete s;
            return nullptr;
        }
        // fixed-point deposit: every cell sum stays below 2^60
        double total = 0.0;
        for (double m : s->mass) total += m;
        d.pm_fix = std::ldexp(1.0, 60 - (int)std::ceil(std::log2(total + 1.0)));
    }

    auto upload = [&](void *dst, const void *src, size_t bytes) -> bool {
        if (bytes == 0) return true;
        return gr_copy_wait(dst, src, bytes, hipMemcpyHostToDevice, d.stream) == hipSuccess;
    };
    bool good = upload(d.row_off, g.row_off.data(), ((size_t)n + 1) * 4) &&
                upload(d.nbr, g.nbr.data(), g.nbr.size() * 4) &&
                upload(d.long_rows, long_rows.data(), long_rows.size() * 4);
    if (d.precision == gr::PREC_FP64) {
        good = good && upload(d.mass, s->mass.data(), (size_t)n * 8);
        if (d.weighted) good = good && upload(d.w, g.roww.data(), g.roww.size() * 8);
    } else {
        std::vector<float> mf(s->mass.begin(), s->mass.end());
        good = good && upload(d.mass, mf.data(), (size_t)n * 4);
        if (d.weighted) {
            std::vector<float> wf(g.roww.begin(), g.roww.end());
            good = good && upload(d.w, wf.data(), wf.size() * 4);
        }
    }
    GrState init;
    std::memset(&init, 0, sizeof(init));
    std::vector<GrGraph> ginit(ngraphs);
    uint32_t live = 0;
    for (uint32_t b = 0; b < ngraphs; b++) {
        GrGraph &gg = ginit[b];
        std::memset(&gg, 0, sizeof(gg));
        gg.swing = gg.traction = gg.speed = gg.speed_efficiency = 1.0;  // state.rs:90-93
        gg.done = node_counts[b] == 0 ? 1u : 0u;  // an empty graph has nothing to run
        live += gg.done ? 0u : 1u;
    }
    init.live = live;
    init.done = live == 0 ? 1u : 0u;
    good = good && upload(d.st, &init, sizeof(init));
    good = good && upload(d.gr, ginit.data(), ginit.size() * sizeof(GrGraph)) &&
           upload(d.gnode, gnode.data(), gnode.size() * 4) &&
           upload(d.gchunk, gchunk.data(), gchunk.size() * 4) &&
           upload(d.chunk, chunk.data(), chunk.size() * 4) &&
           upload(d.chunk_g, chunk_g.data(), chunk_g.size() * 4);
    if (ngraphs > 1) {
        std::vector<uint32_t> gidn(n);
        for (uint32_t b = 0; b < ngraphs; b++)
            for (uint32_t v = gnode[b]; v < gnode[b + 1]; v++) gidn[v] = b;
        good = good && upload(d.gidn, gidn.data(), (size_t)n * 4);
    }
    if (!good) {
        fail(GR_ERR_DEVICE, "hipMemcpy", "upload failed");
        delete s;
        return nullptr;
    }
    std::vector<double> start((size_t)n * d.dim);
    if (x.start) {
        std::memcpy(start.data(), x.start, start.size() * 8);
    } else {  // state.rs:272-283, each graph its own stream from the same seed
        for (uint32_t b = 0; b < ngraphs; b++) {
            uint32_t rs = p.seed;
            for (size_t q = (size_t)gnode[b] * d.dim; q < (size_t)gnode[b + 1] * d.dim; q++)
                start[q] = mulberry32_next(rs);
        }
    }
    {
        gr_fa2_schedule sc;
        bool cached = false;
        {
            std::lock_guard<std::mutex> hold(g_sched_lock);
            auto it = g_sched_cache.find(size_class(n, m, d.dim, d.precision, ngraphs));
            if (it != g_sched_cache.end()) {
                sc = it->second;
                cached = true;
            }
        }
        if (!cached) gr_fa2_schedule_select(n, m, d.dim, d.precision, ngraphs, &sc);
        if (s->apply_schedule(sc) != GR_OK) {
            delete s;
            return nullptr;
        }
    }
    s->size_key_n = n;
    s->size_key_m = m;
    if (n > 0 && s->upload_positions(start.data()) != GR_OK) {
        delete s;
        return nullptr;
    }
    ok();
    return s;
}

gr_fa2_session *gr_fa2_session_create(uint32_t n, const uint32_t *source,
                                      const uint32_t *target, size_t m,
                                      const gr_fa2_params *params, const gr_fa2_ext *ext) {
    return create_impl(1, &n, &m, source, target, params, ext);
}

gr_fa2_session *gr_fa2_batch_create(uint32_t ngraphs, const uint32_t *node_counts,
                                    const size_t *edge_counts, const uint32_t *source,
                                    const uint32_t *target, const gr_fa2_params *params,
                                    const gr_fa2_ext *ext) {
    return create_impl(ngraphs, node_counts, edge_counts, source, target, params, ext);
}

void gr_fa2_session_release(gr_fa2_session *s) { delete s; }

// Runs iterations until `iters` more are applied or the run exits early, growing capacity
// and re-running after a halted iteration. *applied: iterations applied.
static gr_status run_iters(gr_fa2_session *s, uint32_t iters, uint32_t *applied) {
    *applied = 0;
    GrState st;
    gr_status r = s->read_state(st);
    if (r) return r;
    const uint32_t start = st.iters_done;
    uint32_t remaining = iters;
    const uint32_t batch = 32;
    int recoveries = 0;
    while (remaining > 0 && !st.done) {
        uint32_t k = remaining < batch ? remaining : batch;
        const uint32_t before = st.iters_done;
        for (uint32_t q = 0; q < k; q++) HIP_TRY(s->iteration());
        HIP_TRY(hipGetLastError());
        if ((r = s->read_state(st))) return r;
        // launches that applied nothing, with neither a halt nor an early exit to say why,
        // would otherwise be repeated forever
        if (!st.halt && !st.done && st.iters_done == before)
            return fail(GR_ERR_DEVICE, "iteration", "launched iterations applied nothing");
        if (st.halt) {
            if (++recoveries > 8) return fail(GR_ERR_CAPACITY, "tree", "cell capacity");
            if ((r = s->recover(st))) return r;
            if ((r = s->read_state(st))) return r;
        }
        remaining = iters - (st.iters_done - start);
    }
    *applied = st.iters_done - start;
    return GR_OK;
}

gr_status gr_fa2_session_step(gr_fa2_session *s, uint32_t iters, uint32_t *done_out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (done_out) *done_out = 0;
    if (s->d.n == 0 || iters == 0) return ok();
    s->drop_recording();
    uint32_t applied = 0;
    gr_status r = run_iters(s, iters, &applied);
    if (r) return r;
    s->tree_valid = s->d.repulsion == gr::REP_TREE;
    if (done_out) *done_out = applied;
    return ok();
}

gr_status gr_fa2_session_set_params(gr_fa2_session *s, const gr_fa2_params *params,
                                    double theta) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (!params) return fail(GR_ERR_PARAM, "params", "a parameter block");
    gr_fa2_params p = *params;
    if (p.dim != s->d.dim)
        return fail(GR_ERR_UNSUPPORTED, "dim", "the dimension the session was created with");
    // the published ranges (registry/tunable.rs:229-240), refused rather than clamped
    struct Check {
        const char *name;
        double v, lo, hi;
    } checks[3] = {{"jitter_tolerance", p.jitter_tolerance, 0.0, 1e6},
                   {"scaling_ratio", p.scaling_ratio, 0.0, 1e6},
                   {"gravity", p.gravity, 0.0, 1e6}};
    for (const Check &c : checks)
        if (!std::isfinite(c.v) || c.v < c.lo || c.v > c.hi)
            return fail(GR_ERR_PARAM_OUT_OF_RANGE, c.name,
                        "inside the range the schema publishes, never clamped");
    if (!std::isfinite(theta)) return fail(GR_ERR_PARAM, "theta", "a finite opening angle");
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    s->params = p;
    s->d.scaling_ratio = p.scaling_ratio;
    s->d.gravity = p.gravity;
    s->d.jitter_tolerance = p.jitter_tolerance;
    s->d.seed = p.seed;
    if (theta >= 0.0) s->d.theta2 = theta * theta;
    s->drop_graph();      // its launches carry the old values
    s->drop_recording();  // a reverse sweep would replay them with the new ones
    return ok();
}

gr_status gr_fa2_session_record(gr_fa2_session *s, uint32_t iters, uint32_t checkpoint_every,
                                uint32_t *done_out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (done_out) *done_out = 0;
    if (checkpoint_every == 0)
        return fail(GR_ERR_PARAM, "checkpoint_every", "at least one iteration per segment");
    if (s->d.ngraphs != 1)
        return fail(GR_ERR_UNSUPPORTED, "adjoint", "a session of one graph");
    if (s->d.repulsion != gr::REP_TREE && s->d.repulsion != gr::REP_DENSE)
        return fail(GR_ERR_UNSUPPORTED, "adjoint", "tree or dense repulsion");
    s->drop_recording();
    auto &rec = s->rec;
    rec.every = checkpoint_every;
    if (s->d.n == 0 || iters == 0) {
        rec.valid = true;
        return ok();
    }
    const uint32_t maxseg = (uint32_t)(((uint64_t)iters + checkpoint_every - 1) / checkpoint_every);
    const size_t pb = s->pos_bytes();
    gr_status r;
    if ((r = s->rec_alloc(&rec.ck_pos, (size_t)maxseg * pb)) ||
        (r = s->rec_alloc(&rec.ck_g, (size_t)maxseg * sizeof(GrGraph))) ||
        (r = s->rec_alloc(&rec.ck_st, (size_t)maxseg * sizeof(GrState))) ||
        (r = s->rec_alloc(&rec.fin_pos, pb)) || (r = s->rec_alloc(&rec.fin_g, sizeof(GrGraph))) ||
        (r = s->rec_alloc(&rec.fin_st, sizeof(GrState)))) {
        s->drop_recording();
        return r;
    }
    uint32_t total = 0;
    for (uint32_t seg = 0; seg < maxseg; seg++) {
        if ((r = s->save_state((char *)rec.ck_pos + (size_t)seg * pb, rec.ck_g + seg,
                               rec.ck_st + seg))) {
            s->drop_recording();
            return r;
        }
        const uint32_t want = iters - total < checkpoint_every ? iters - total : checkpoint_every;
        uint32_t got = 0;
        if ((r = run_iters(s, want, &got))) {
            s->drop_recording();
            return r;
        }
        rec.seg_len.push_back(got);
        rec.nseg++;
        total += got;
        if (got < want) break;  // the run exited early
    }
    if ((r = s->save_state(rec.fin_pos, rec.fin_g, rec.fin_st))) {
        s->drop_recording();
        return r;
    }
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    rec.valid = true;
    s->tree_valid = s->d.repulsion == gr::REP_TREE;
    if (done_out) *done_out = total;
    return ok();
}

gr_status gr_fa2_session_adjoint(gr_fa2_session *s, const double *pos_bar, double *weight_bar,
                                 double *param_bar, double *start_bar) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    auto &rec = s->rec;
    if (!rec.valid)
        return fail(GR_ERR_UNSUPPORTED, "adjoint", "a recording (gr_fa2_session_record)");
    gr::FaDev &d = s->d;
    const uint32_t n = d.n, dim = d.dim;
    if (n > 0 && !pos_bar) return fail(GR_ERR_PARAM, "pos_bar", "n*dim loss gradients");
    if (weight_bar) std::memset(weight_bar, 0, s->m_raw * sizeof(double));
    if (param_bar) param_bar[0] = param_bar[1] = 0.0;
    if (n == 0 || rec.nseg == 0) {
        if (start_bar && n) std::memcpy(start_bar, pos_bar, (size_t)n * dim * 8);
        return ok();
    }
    // temporaries of this call
    std::vector<void *> tmp;
    struct Free {
        std::vector<void *> &v;
        hipStream_t st;
        ~Free() {
            (void)hipStreamSynchronize(st);
            for (void *p : v) (void)hipFree(p);
        }
    } release{tmp, d.stream};
    auto get = [&](void **out, size_t bytes) -> gr_status {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 8);
        if (e != hipSuccess) return fail(GR_ERR_DEVICE, "hipMalloc", hipGetErrorString(e));
        tmp.push_back(p);
        *out = p;
        return GR_OK;
    };
    gr::AdjDev a;
    const uint32_t nbl = gr_blocks(n, GR_BLOCK);
    // the tape holds one segment: size it by the longest one recorded (checkpoint_every
    // may exceed the iterations run)
    size_t every = 1;
    for (uint32_t L : rec.seg_len) every = std::max<size_t>(every, L);
    const size_t pb = s->pos_bytes();
    void *tape_pos = nullptr, *tape_g = nullptr, *tape_st = nullptr, *cacc = nullptr;
    uint32_t cacc_cap = 0;
    gr_status r;
    for (uint32_t x = 0; x < dim; x++) {
        if ((r = get((void **)&a.pb[x], (size_t)n * 8)) ||
            (r = get((void **)&a.ub[x], (size_t)n * 8)))
            return r;
    }
    a.m = s->m_simple;
    if ((r = get((void **)&a.pscat, (size_t)n * dim * 8 * gr::ADJ_LIMBS)) ||
        (r = get((void **)&a.wb, (size_t)a.m * 8)) || (r = get((void **)&a.elo, (size_t)a.m * 4)) ||
        (r = get((void **)&a.ehi, (size_t)a.m * 4)) ||
        (r = get((void **)&a.sc, sizeof(gr::AdjScal))) ||
        (r = get((void **)&a.part, (size_t)nbl * gr::ADJ_NPART * 8)) ||
        (r = get(&tape_pos, every * pb)) || (r = get(&tape_g, (every + 1) * sizeof(GrGraph))) ||
        (r = get(&tape_st, (every + 1) * sizeof(GrState))))
        return r;
    {
        std::vector<double> col(n);
        for (uint32_t x = 0; x < dim; x++) {
            for (uint32_t i = 0; i < n; i++) col[i] = pos_bar[(size_t)i * dim + x];
            HIP_TRY(gr_copy_wait(a.pb[x], col.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                                 d.stream));
        }
        if (a.m) {
            HIP_TRY(gr_copy_wait(a.elo, s->lo.data(), (size_t)a.m * 4, hipMemcpyHostToDevice,
                                 d.stream));
            HIP_TRY(gr_copy_wait(a.ehi, s->hi.data(), (size_t)a.m * 4, hipMemcpyHostToDevice,
                                 d.stream));
        }
        HIP_TRY(gr_fill_wait(a.wb, 0, (size_t)a.m * 8, d.stream));
        HIP_TRY(gr_fill_wait(a.sc, 0, sizeof(gr::AdjScal), d.stream));
    }
    GrGraph *tg = (GrGraph *)tape_g;
    GrState *ts = (GrState *)tape_st;
    GrState st;
    for (uint32_t seg = rec.nseg; seg-- > 0;) {
        const uint32_t L = rec.seg_len[seg];
        if (L == 0) continue;
        // recompute the segment from its checkpoint, keeping every iteration's start state
        int tries = 0;
        for (;;) {
            if ((r = s->load_state((char *)rec.ck_pos + (size_t)seg * pb, rec.ck_g + seg,
                                   rec.ck_st + seg)))
                return r;
            for (uint32_t j = 0; j < L; j++) {
                if ((r = s->save_state((char *)tape_pos + (size_t)j * pb, tg + j, ts + j)))
                    return r;
                HIP_TRY(s->iteration());
            }
            HIP_TRY(hipMemcpyAsync(tg + L, d.gr, sizeof(GrGraph), hipMemcpyDeviceToDevice,
                                   d.stream));
            HIP_TRY(hipGetLastError());
            if ((r = s->read_state(st))) return r;
            if (!st.halt) break;
            if (++tries > 8) return fail(GR_ERR_CAPACITY, "tree", "cell capacity");
            if ((r = s->recover(st))) return r;
        }
        if (cacc_cap < d.cap) {
            if ((r = get(&cacc, (size_t)d.cap * dim * 8 * gr::ADJ_LIMBS))) return r;
            cacc_cap = d.cap;
            a.cacc = (unsigned long long *)cacc;
        }
        for (uint32_t j = L; j-- > 0;) {
            if ((r = s->load_state((char *)tape_pos + (size_t)j * pb, tg + j, ts + j))) return r;
            gr::launch_forces(d, gr::PART_ATTRACTION | gr::PART_REPULSION | gr::PART_GRAVITY);
            a.g0 = tg + j;
            a.g1 = tg + j + 1;
            gr::launch_adjoint_step(d, a);
        }
        HIP_TRY(hipGetLastError());
        if ((r = s->read_state(st))) return r;
        if (st.halt) return fail(GR_ERR_CAPACITY, "tree", "halted while reversing");
    }
    // outputs
    if (start_bar) {
        std::vector<double> col(n);
        for (uint32_t x = 0; x < dim; x++) {
            HIP_TRY(gr_copy_wait(col.data(), a.pb[x], (size_t)n * 8, hipMemcpyDeviceToHost,
                                 d.stream));
            for (uint32_t i = 0; i < n; i++) start_bar[(size_t)i * dim + x] = col[i];
        }
    }
    if (weight_bar && a.m) {
        std::vector<double> wb(a.m);
        HIP_TRY(gr_copy_wait(wb.data(), a.wb, (size_t)a.m * 8, hipMemcpyDeviceToHost, d.stream));
        for (uint32_t e = 0; e < a.m; e++) weight_bar[s->simple_raw[e]] = wb[e];
    }
    if (param_bar) {
        gr::AdjScal sc;
        HIP_TRY(gr_copy_wait(&sc, a.sc, sizeof(sc), hipMemcpyDeviceToHost, d.stream));
        param_bar[0] = sc.kbar;
        param_bar[1] = sc.gbar;
    }
    // leave the session as the recording left it
    if ((r = s->load_state(rec.fin_pos, rec.fin_g, rec.fin_st))) return r;
    HIP_TRY(hipStreamSynchronize(d.stream));
    return ok();
}

static gr_status read_columns(gr_fa2_session *s, void *const *cols, double *out) {
    const uint32_t n = s->d.n, dim = s->d.dim;
    for (uint32_t a = 0; a < dim; a++) {
        if (s->d.precision == gr::PREC_FP64) {
            std::vector<double> c(n);
            HIP_TRY(gr_copy_wait(c.data(), cols[a], (size_t)n * 8, hipMemcpyDeviceToHost,
                                 s->d.stream));
            for (uint32_t i = 0; i < n; i++) out[(size_t)i * dim + a] = c[i];
        } else {
            std::vector<float> c(n);
            HIP_TRY(gr_copy_wait(c.data(), cols[a], (size_t)n * 4, hipMemcpyDeviceToHost,
                                 s->d.stream));
            for (uint32_t i = 0; i < n; i++) out[(size_t)i * dim + a] = (double)c[i];
        }
    }
    return GR_OK;
}

gr_status gr_fa2_session_positions(gr_fa2_session *s, double *out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    gr_status r = read_columns(s, s->d.pos, out);
    return r ? r : ok();
}

gr_status gr_fa2_session_set_positions(gr_fa2_session *s, const double *in) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    if (s->d.n == 0) return ok();
    s->drop_recording();
    gr_status r = s->upload_positions(in);
    return r ? r : ok();
}

gr_status gr_fa2_session_forces(gr_fa2_session *s, uint32_t parts, double *out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (s->d.n == 0) return ok();
    GrState st;
    gr_status r;
    for (int tries = 0; tries < 8; tries++) {
        // the bounds partials of the current positions are always current: they are
        // written at upload and by every applied step
        gr::launch_forces(s->d, parts);
        HIP_TRY(hipGetLastError());
        if ((r = s->read_state(st))) return r;
        if (!st.halt) break;
        if ((r = s->recover(st))) return r;
    }
    if (st.halt) return fail(GR_ERR_CAPACITY, "tree", "cell capacity");
    s->tree_valid = s->d.repulsion == gr::REP_TREE;
    r = read_columns(s, s->d.u, out);
    return r ? r : ok();
}

gr_status gr_fa2_session_graph_scalars(gr_fa2_session *s, uint32_t graph, double *out7) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (graph >= s->d.ngraphs) return fail(GR_ERR_PARAM, "graph", "a graph of the session");
    GrGraph g;
    HIP_TRY(hipMemcpyAsync(&g, s->d.gr + graph, sizeof(g), hipMemcpyDeviceToHost, s->d.stream));
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    out7[0] = g.swing;
    out7[1] = g.traction;
    out7[2] = g.speed;
    out7[3] = g.speed_efficiency;
    out7[4] = g.iter_no;
    out7[5] = g.iters_done;
    out7[6] = g.done;
    return ok();
}

gr_status gr_fa2_session_scalars(gr_fa2_session *s, double *out7) {
    return gr_fa2_session_graph_scalars(s, 0, out7);
}

gr_status gr_fa2_session_graph(gr_fa2_session *s, uint32_t *m_out, uint32_t *lo, uint32_t *hi,
                               double *mass) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (m_out) *m_out = s->m_simple;
    if (lo) std::memcpy(lo, s->lo.data(), s->lo.size() * 4);
    if (hi) std::memcpy(hi, s->hi.data(), s->hi.size() * 4);
    if (mass) std::memcpy(mass, s->mass.data(), s->mass.size() * 8);
    return ok();
}

gr_status gr_fa2_session_tree(gr_fa2_session *s, uint32_t *ncells_out, uint32_t cap,
                              double *bounds, uint32_t *skip, uint32_t *start, uint32_t *end,
                              double *bodies, uint32_t *order) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (s->d.repulsion != gr::REP_TREE || !s->tree_valid)
        return fail(GR_ERR_UNSUPPORTED, "tree", "a tree session after forces or a step");
    GrState st;
    gr_status r = s->read_state(st);
    if (r) return r;
    const uint32_t C = st.n_cells, dim = s->d.dim;
    if (ncells_out) *ncells_out = C;
    if (order && s->d.n)
        HIP_TRY(gr_copy_wait(order, s->d.order, (size_t)s->d.n * 4, hipMemcpyDeviceToHost,
                             s->d.stream));
    if (C == 0) return ok();
    if (cap < C) {  // a size query (no cell output) is the first call of two
        if (bounds || skip || start || end || bodies)
            return fail(GR_ERR_CAPACITY, "cap", "cap >= *ncells_out cells in every output");
        return ok();
    }
    if (bounds) {
        if (!s->d.export_bounds)
            return fail(GR_ERR_UNSUPPORTED, "bounds", "create the session with GR_EXT_EXPORT_TREE");
        HIP_TRY(gr_copy_wait(bounds, s->d.cbounds, (size_t)C * 2 * dim * 8, hipMemcpyDeviceToHost,
                             s->d.stream));
    }
    if (skip || start || end) {
        size_t wb = gr::walkcell_bytes(s->d.precision, dim);
        std::vector<uint8_t> buf((size_t)C * wb);
        HIP_TRY(gr_copy_wait(buf.data(), s->d.wc, buf.size(), hipMemcpyDeviceToHost, s->d.stream));
        size_t es = s->esize();
        // [clo[dim] in the FP32 precisions], c[dim], mass, w2, then skip, start, end
        size_t off = es * (dim + 2) + (s->d.precision == gr::PREC_FP64 ? 0 : es * dim);
        for (uint32_t k = 0; k < C; k++) {
            const uint8_t *c = buf.data() + (size_t)k * wb + off;
            uint32_t v[3];
            std::memcpy(v, c, 12);
            if (skip) skip[k] = v[0];
            if (start) start[k] = v[1];
            if (end) end[k] = v[2];
        }
    }
    if (bodies) {
        std::vector<double> raw((size_t)C * (1 + dim));
        HIP_TRY(gr_copy_wait(raw.data(), s->d.raw, raw.size() * 8, hipMemcpyDeviceToHost,
                             s->d.stream));
        for (uint32_t k = 0; k < C; k++) {
            double m = raw[(size_t)k * (1 + dim)];
            bodies[(size_t)k * (1 + dim)] = m;
            for (uint32_t a = 0; a < dim; a++)
                bodies[(size_t)k * (1 + dim) + 1 + a] = raw[(size_t)k * (1 + dim) + 1 + a] / m;
        }
    }
    return ok();
}

// The table of measured schedules: what gr_fa2_session_autotune chose on MI355X for the
// benchmark workloads (`bench.py --workload autotune`): 10K-node SBM 1.057x (FP32) and
// 1.020x (FP64) over the default {256, 1, 8, 1}, the 4096-graph batch 1.038x, the 250K 3D
// and 1M graphs within noise. Keyed on node count, precision and batch; every entry gives
// the same bits.
gr_status gr_fa2_schedule_select(uint32_t n, size_t m, uint32_t dim, int32_t precision,
                                 uint32_t ngraphs, gr_fa2_schedule *out) {
    if (!out) return fail(GR_ERR_PARAM, "out", "a schedule to fill");
    (void)m;
    (void)dim;
    if (ngraphs > 1)
        *out = gr_fa2_schedule{64, 1, 16, 1};
    else if (n <= 131072)
        *out = gr_fa2_schedule{64, 1, 4, precision == GR_PREC_FP64 ? 2u : 8u};
    else
        *out = gr_fa2_schedule{64, 1, 8, 1};
    return ok();
}

gr_status gr_fa2_session_get_schedule(gr_fa2_session *s, gr_fa2_schedule *out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (!out) return fail(GR_ERR_PARAM, "out", "a schedule to fill");
    *out = s->sched;
    return ok();
}

gr_status gr_fa2_session_set_schedule(gr_fa2_session *s, const gr_fa2_schedule *sched) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (!sched || !schedule_valid(*sched))
        return fail(GR_ERR_PARAM, "schedule",
                    "walk_block 64/128/256, walk_xcd 0/1, sort_items 4/8/16, finish_split 1/2/4/8");
    gr_status r = s->apply_schedule(*sched);
    return r ? r : ok();
}

gr_status gr_fa2_session_autotune(gr_fa2_session *s, uint32_t reps, gr_fa2_schedule *out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (s->d.repulsion != gr::REP_TREE)
        return fail(GR_ERR_UNSUPPORTED, "autotune", "a tree session");
    if (reps == 0) reps = 5;
    gr_fa2_schedule best = s->sched;
    if (s->d.n == 0) {
        if (out) *out = best;
        return ok();
    }
    s->drop_recording();
    gr_status r;
    std::vector<double> tmp((size_t)s->d.n * s->d.dim);
    if ((r = gr_fa2_session_forces(s, 7, tmp.data()))) return r;  // a current tree
    // walk: workgroup size x XCD mapping
    double bt = 1e300;
    for (uint32_t wb : {64u, 128u, 256u})
        for (uint32_t xcd : {0u, 1u}) {
            gr_fa2_schedule c = best;
            c.walk_block = wb;
            c.walk_xcd = xcd;
            if ((r = s->apply_schedule(c))) return r;
            double ms;
            if ((r = time_median(s, reps, [&] { gr::launch_walk(s->d); }, &ms))) return r;
            if (ms < bt) {
                bt = ms;
                best = c;
            }
        }
    // the tree build's radix-sort tile
    bt = 1e300;
    gr_fa2_schedule base = best;
    for (uint32_t items : {4u, 8u, 16u}) {
        gr_fa2_schedule c = base;
        c.sort_items = items;
        if ((r = s->apply_schedule(c))) return r;
        double ms;
        if ((r = time_median(s, reps, [&] { gr::launch_tree(s->d); }, &ms))) return r;
        if (ms < bt) {
            bt = ms;
            best = c;
        }
    }
    // the force finish's split
    bt = 1e300;
    base = best;
    if ((r = s->apply_schedule(base))) return r;
    gr::launch_walk(s->d);  // the finish reads the walk's repulsion of the current tree
    for (uint32_t split : {1u, 2u, 4u, 8u}) {
        gr_fa2_schedule c = base;
        c.finish_split = split;
        if ((r = s->apply_schedule(c))) return r;
        double ms;
        if ((r = time_median(s, reps, [&] { gr::launch_finish(s->d); }, &ms))) return r;
        if (ms < bt) {
            bt = ms;
            best = c;
        }
    }
    if ((r = s->apply_schedule(best))) return r;
    GrState st;
    if ((r = s->read_state(st))) return r;
    if (st.halt) return fail(GR_ERR_CAPACITY, "tree", "halted while tuning");
    // the timed stages rewrote the force buffers; the state (positions, speed) is untouched
    if ((r = gr_fa2_session_forces(s, 7, tmp.data()))) return r;
    {
        std::lock_guard<std::mutex> hold(g_sched_lock);
        g_sched_cache[size_class(s->size_key_n, s->size_key_m, s->d.dim, s->d.precision,
                                 s->d.ngraphs)] = best;
    }
    if (out) *out = best;
    return ok();
}

gr_status gr_fa2_session_time(gr_fa2_session *s, const char *stage, uint32_t reps,
                              double *ms_out) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (!stage || reps == 0) return fail(GR_ERR_PARAM, "stage", "a stage name and reps > 0");
    std::string name(stage);
    bool walk = name == "walk", tree = name == "tree", it = name == "iteration",
         pm = name == "pm";
    if (!walk && !tree && !it && !pm)
        return fail(GR_ERR_PARAM, "stage", "walk, tree, iteration or pm");
    if ((walk || tree) && s->d.repulsion != gr::REP_TREE)
        return fail(GR_ERR_UNSUPPORTED, "stage", "walk/tree timing needs a tree session");
    if (pm && s->d.repulsion != gr::REP_PM)
        return fail(GR_ERR_UNSUPPORTED, "stage", "pm timing needs a PM session");
    if (it) s->drop_recording();  // timed iterations move the state
    if (walk && !s->tree_valid) {
        double *tmp = (double *)std::malloc((size_t)s->d.n * s->d.dim * 8 + 8);
        gr_status r = gr_fa2_session_forces(s, 7, tmp);
        std::free(tmp);
        if (r) return r;
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, s->d.stream));
    for (uint32_t q = 0; q < reps; q++) {
        if (walk) gr::launch_walk(s->d);
        if (tree) gr::launch_tree(s->d);
        if (it) HIP_TRY(s->iteration());
        if (pm) gr::launch_forces(s->d, 2);
    }
    HIP_TRY(hipEventRecord(e1, s->d.stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = (double)ms / reps;
    GrState st;
    gr_status r = s->read_state(st);
    if (r) return r;
    if (st.halt) return fail(GR_ERR_CAPACITY, "tree", "halted while timing");
    return ok();
}

gr_status gr_fa2_session_walk_stats(gr_fa2_session *s, uint64_t *out4) {
    if (!s) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session handle");
    if (!out4) return fail(GR_ERR_PARAM, "out", "four counters");
    if (s->d.repulsion != gr::REP_TREE)
        return fail(GR_ERR_UNSUPPORTED, "walk", "a tree session");
    for (int q = 0; q < 4; q++) out4[q] = 0;
    if (s->d.n == 0) return ok();
    if ((uint64_t)(s->d.cap + 1) * gr::walkcell_bytes(s->d.precision, s->d.dim) >= (1ull << 31))
        return fail(GR_ERR_UNSUPPORTED, "walk", "a cell arena under 2 GiB");
    if (!s->tree_valid) {
        std::vector<double> tmp((size_t)s->d.n * s->d.dim);
        gr_status r = gr_fa2_session_forces(s, 7, tmp.data());
        if (r) return r;
    }
    HIP_TRY(hipMemsetAsync(s->d.wstat, 0, 4 * sizeof(uint64_t), s->d.stream));
    gr::launch_walk_count(s->d);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(out4, s->d.wstat, 4 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           s->d.stream));
    HIP_TRY(hipStreamSynchronize(s->d.stream));
    return ok();
}

// mesh.rs:31-34: ceil(sqrt n) rounded up to a power of two, held to 128..1024
uint32_t gr_pm_side_for(uint32_t n) {
    const double root = std::ceil(std::sqrt((double)n));
    uint32_t side = 1;
    while ((double)side < root) side <<= 1;
    return side < 128 ? 128 : (side > 1024 ? 1024 : side);
}

gr_status gr_pm_charge(uint32_t n, const double *xy, uint32_t side, double charge, double alpha,
                       double distance_min, double distance_max, double *out,
                       double *frame_out) {
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (frame_out)
        for (int k = 0; k < 7; k++) frame_out[k] = 0.0;
    if (!std::isfinite(charge) || !std::isfinite(alpha) || !(distance_min >= 0.0) ||
        !(distance_max >= 0.0) || std::isnan(distance_max))
        return fail(GR_ERR_PARAM, "charge", "finite charge and alpha, distances >= 0");
    if (side == 0) side = gr_pm_side_for(n);
    if (n == 0) return ok();
    gr_fa2_params p;
    gr_fa2_params_default(&p);
    gr_fa2_ext x;
    gr_fa2_ext_default(&x);
    x.repulsion = GR_REP_PM;
    x.precision = GR_PREC_FP64;
    x.pm_grid = side;
    x.start = xy;
    x.flags = GR_EXT_DIRECT_LAUNCH | GR_EXT_BLOCK_SUMS;  // the mesh reads no global sum
    gr_fa2_session *s = gr_fa2_session_create(n, nullptr, nullptr, 0, &p, &x);
    if (!s) return g_err.code;
    gr::FaDev &d = s->d;
    d.pm_law = 1;
    d.pm_dmin2 = distance_min * distance_min;
    d.pm_dmax2 = distance_max * distance_max;
    d.pm_strength = charge * alpha;
    d.pm_fix = std::ldexp(1.0, 60 - (int)std::ceil(std::log2((double)n + 1.0)));  // unit charges
    gr::launch_forces(d, gr::PART_REPULSION);
    gr_status r = GR_OK;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(d.stream) != hipSuccess)
        r = fail(GR_ERR_DEVICE, "pm", "the particle-mesh launches failed");
    if (!r) r = read_columns(s, d.rep, out);
    if (!r && frame_out) {
        std::vector<uint8_t> fr(gr::pm_frame_bytes());
        if (gr_copy_wait(fr.data(), d.pm_fr, fr.size(), hipMemcpyDeviceToHost,
                         d.stream) != hipSuccess) {
            r = fail(GR_ERR_DEVICE, "pm", "frame readback");
        } else {
            gr::pm_frame_read(fr.data(), frame_out);
        }
    }
    gr_fa2_session_release(s);
    return r ? r : ok();
}

// d3's many-body pass over the quadtree (barnes_hut/charge.rs) on an FP64 2D session that
// holds only the positions: the FA2 tree build, then the d3 bodies, node ids and walk
// (d3.hip).
struct D3Pass {
    gr_fa2_session *s;
    gr::D3Dev g;
    std::vector<uint32_t> lvl;  // the level table read back
    uint32_t cap = 0;           // cells the buffers hold (a tick's tree may be larger)
    bool have_ids = false;

    // The tree of the session's positions; grows the arena when it runs out.
    gr_status tree() {
        gr::FaDev &d = s->d;
        GrState st;
        gr_status r;
        for (int tries = 0; tries < 8; tries++) {
            gr::launch_tree_arena(d);  // d3 forms its own bodies
            HIP_TRY(hipGetLastError());
            if ((r = s->read_state(st))) return r;
            if (!st.halt) break;
            if ((r = s->recover(st))) return r;
        }
        if (st.halt) return fail(GR_ERR_CAPACITY, "tree", "cell capacity");
        g.n = d.n;
        g.nf = st.n_finite;
        g.C = st.n_cells;
        g.wc = d.wc;
        g.order = d.order;
        g.sx = (const double *)d.sp[0];
        g.sy = (const double *)d.sp[1];
        g.cb = d.cbounds;
        g.stream = d.stream;
        return GR_OK;
    }

    // The pass's buffers for the current tree's cell count, kept while later trees fit.
    gr_status alloc(bool ids) {
        if (g.C <= cap && (have_ids || !ids) && cap > 0) return GR_OK;
        const size_t c = std::max<size_t>(g.C, (size_t)cap + cap / 4) + 64;
        gr_status r;
        if ((r = s->alloc(&g.body, c * sizeof(gr::D3Body))) ||
            (r = s->alloc(&g.diff, (c + 1) * 4)) || (r = s->alloc(&g.depth, c * 4)) ||
            (r = s->alloc(&g.amin, c * 4)) || (r = s->alloc(&g.bmin, c * 4)) ||
            (r = s->alloc(&g.keys, c * 8)) || (r = s->alloc(&g.keys_alt, c * 8)) ||
            (r = s->alloc(&g.vals, c * 4)) || (r = s->alloc(&g.vals_alt, c * 4)) ||
            (r = s->alloc(&g.lvl, (c + 3) * 4)) ||
            (r = s->alloc(&g.tmp, gr::d3_tmp_bytes((uint32_t)c))) ||
            (r = s->alloc(&g.out, (size_t)(g.n ? g.n : 1) * 16)) ||
            (ids && (r = s->alloc(&g.nid_out, c * 4))))
            return r;
        cap = (uint32_t)c;
        have_ids = have_ids || ids;
        return GR_OK;
    }

    // Bodies and node ids: leaves, depths, then the internal cells deepest level first (one
    // launch per depth; the level table is read back once).
    gr_status bodies() {
        const uint32_t C = g.C;
        if (C == 0) return GR_OK;
        gr::d3_launch_cells(g);
        const uint32_t *sorted = gr::d3_launch_depth(g);
        HIP_TRY(hipGetLastError());
        // the table's head and the first 64 depths in one read; deeper trees read the rest
        const size_t first = std::min<size_t>((size_t)C + 3, 66);
        lvl.assign(first, 0);
        HIP_TRY(hipMemcpyAsync(lvl.data(), g.lvl, first * 4, hipMemcpyDeviceToHost, g.stream));
        HIP_TRY(hipStreamSynchronize(g.stream));
        const uint32_t maxd = lvl[0], nint = lvl[1];
        if (nint > C || (nint > 0 && maxd >= C))
            return fail(GR_ERR_DEVICE, "tree", "a level table outside the arena");
        if (nint > 0 && (size_t)maxd + 3 > first) {
            lvl.resize((size_t)maxd + 3);
            HIP_TRY(gr_copy_wait(lvl.data() + first, g.lvl + first, (lvl.size() - first) * 4,
                                 hipMemcpyDeviceToHost, g.stream));
        }
        if (nint > 0)
            for (uint32_t l = maxd + 1; l-- > 0;) {
                const uint32_t a = lvl[2 + l], b = l == maxd ? nint : lvl[3 + l];
                if (a > b || b > nint)
                    return fail(GR_ERR_DEVICE, "tree", "a level table out of order");
                if (b > a) gr::d3_launch_level(g, sorted, a, b - a);
            }
        gr::d3_launch_ids(g);
        HIP_TRY(hipGetLastError());
        return GR_OK;
    }

    // The whole pass again on the resident positions: tree, bodies, walk.
    gr_status again() {
        const uint32_t C = g.C;
        gr_status r;
        if ((r = tree())) return r;
        if (g.C != C) return fail(GR_ERR_DEVICE, "tree", "the same positions gave another tree");
        if ((r = bodies())) return r;
        gr::d3_launch_walk(g);
        HIP_TRY(hipGetLastError());
        return GR_OK;
    }
};

// Writes the deltas when `out` is given and the node ids when `node_id` is given and
// cap >= the cell count (*ncells_out); with neither it is a size query. reps > 0 then times
// the whole pass and the walk alone on the resident positions (ms_out[0], ms_out[1]: medians
// of three timings of `reps` runs) and gives the cell count (ms_out[2]).
static gr_status d3_pass(gr_fa2_session *s, double theta, double charge, double alpha,
                         double dmin, double dmax, uint32_t seed, uint32_t tick, double *out,
                         uint32_t cap, uint32_t *ncells_out, uint32_t *node_id, uint32_t reps,
                         double *ms_out) {
    D3Pass p{s, {}, {}, 0, false};
    gr_status r;
    if ((r = p.tree())) return r;
    gr::D3Dev &g = p.g;
    const uint32_t C = g.C, n = g.n;
    if (ncells_out) *ncells_out = C;
    const bool ids = node_id && cap >= C;
    if (node_id && !ids) return fail(GR_ERR_CAPACITY, "cap", "cap >= *ncells_out node ids");
    if (!out && !ids) return GR_OK;
    g.theta = theta;
    g.theta2 = theta * theta;
    g.dmin2 = dmin * dmin;
    g.dmax2 = dmax * dmax;
    g.charge = charge;
    g.alpha = alpha;
    g.seed = seed;
    g.tick = tick;
    if ((r = p.alloc(ids)) || (r = p.bodies())) return r;
    if (out) gr::d3_launch_walk(g);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g.stream));
    if (out) HIP_TRY(gr_copy_wait(out, g.out, (size_t)n * 16, hipMemcpyDeviceToHost, g.stream));
    if (ids && C) HIP_TRY(gr_copy_wait(node_id, g.nid_out, (size_t)C * 4, hipMemcpyDeviceToHost,
                                       g.stream));
    if (reps && ms_out) {
        gr_status e = GR_OK;
        if ((r = time_median(s, reps, [&] { if (!e) e = p.again(); }, &ms_out[0]))) return r;
        if (e) return e;
        if ((r = time_median(s, reps, [&] { gr::d3_launch_walk(g); }, &ms_out[1]))) return r;
        ms_out[2] = (double)C;
        // the timed runs recomputed the same deltas: they must carry the same bits
        std::vector<double> again((size_t)n * 2);
        HIP_TRY(hipStreamSynchronize(g.stream));
        HIP_TRY(gr_copy_wait(again.data(), g.out, (size_t)n * 16, hipMemcpyDeviceToHost, g.stream));
        if (out && std::memcmp(again.data(), out, (size_t)n * 16) != 0)
            return fail(GR_ERR_DEVICE, "d3", "a timed run gave other deltas");
    }
    return GR_OK;
}

static gr_fa2_session *d3_session(uint32_t n, const double *xy) {
    gr_fa2_params p;
    gr_fa2_params_default(&p);
    gr_fa2_ext x;
    gr_fa2_ext_default(&x);
    x.repulsion = GR_REP_TREE;
    x.precision = GR_PREC_FP64;
    x.start = xy;
    // the bounds of every cell are kept: a body's opening threshold is its cell's width
    x.flags = GR_EXT_DIRECT_LAUNCH | GR_EXT_BLOCK_SUMS | GR_EXT_EXPORT_TREE;
    return gr_fa2_session_create(n, nullptr, nullptr, 0, &p, &x);
}

gr_status gr_d3_charge_bh(uint32_t n, const double *xy, double theta, double charge,
                          double alpha, double distance_min, double distance_max, uint32_t seed,
                          uint32_t tick, double *out) {
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (!std::isfinite(charge) || !std::isfinite(alpha) || !(distance_min >= 0.0) ||
        !(distance_max >= 0.0) || std::isnan(distance_max))
        return fail(GR_ERR_PARAM, "charge", "finite charge and alpha, distances >= 0");
    if (!(theta >= 0.0)) return fail(GR_ERR_PARAM, "theta", "an opening angle >= 0");
    if (n == 0) return ok();
    gr_fa2_session *s = d3_session(n, xy);
    if (!s) return g_err.code;
    gr_status r = d3_pass(s, theta, charge, alpha, distance_min, distance_max, seed, tick, out, 0,
                          nullptr, nullptr, 0, nullptr);
    gr_fa2_session_release(s);
    return r ? r : ok();
}

gr_status gr_d3_charge_bh_time(uint32_t n, const double *xy, double theta, double charge,
                               double alpha, double distance_min, double distance_max,
                               uint32_t seed, uint32_t tick, uint32_t reps, double *out,
                               double *ms_out) {
    if (!ms_out || reps == 0) return fail(GR_ERR_PARAM, "reps", "reps >= 1 and ms_out[3]");
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (!std::isfinite(charge) || !std::isfinite(alpha) || !(distance_min >= 0.0) ||
        !(distance_max >= 0.0) || std::isnan(distance_max))
        return fail(GR_ERR_PARAM, "charge", "finite charge and alpha, distances >= 0");
    if (!(theta >= 0.0)) return fail(GR_ERR_PARAM, "theta", "an opening angle >= 0");
    for (int k = 0; k < 3; k++) ms_out[k] = 0.0;
    if (n == 0) return ok();
    gr_fa2_session *s = d3_session(n, xy);
    if (!s) return g_err.code;
    gr_status r = d3_pass(s, theta, charge, alpha, distance_min, distance_max, seed, tick, out, 0,
                          nullptr, nullptr, reps, ms_out);
    gr_fa2_session_release(s);
    return r ? r : ok();
}

gr_status gr_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed,
                        uint32_t tick, double *out) {
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (!std::isfinite(radius) || radius < 0.0)
        return fail(GR_ERR_PARAM, "radius", "a finite radius >= 0");
    if (n == 0) return ok();
    gr_fa2_session *s = d3_session(n, xy);
    if (!s) return g_err.code;
    gr_status r = GR_OK;
    {
        D3Pass p{s, {}, {}, 0, false};
        gr::D3SimDev sd;
        double *dl = nullptr;
        if (!(r = p.tree()) && !(r = s->alloc(&dl, (size_t)n * 16))) {
            sd.n = n;
            sd.dl = dl;
            sd.seed = seed;
            sd.reach2 = (2.0 * radius) * (2.0 * radius);
            sd.stream = s->d.stream;
            gr::D3CollideTree t;
            t.C = p.g.C;
            t.n = n;
            t.wc = p.g.wc;
            t.cb = p.g.cb;
            t.order = p.g.order;
            t.sx = p.g.sx;
            t.sy = p.g.sy;
            gr::d3sim_collide(sd, t, tick);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(sd.stream) != hipSuccess ||
                gr_copy_wait(out, dl, (size_t)n * 16, hipMemcpyDeviceToHost,
                             sd.stream) != hipSuccess)
                r = fail(GR_ERR_DEVICE, "hip", "the collide pass failed on the device");
        }
    }
    gr_fa2_session_release(s);
    return r ? r : ok();
}

gr_status gr_quadtree_node_ids(uint32_t n, const double *xy, uint32_t cap, uint32_t *ncells_out,
                               uint32_t *node_id) {
    if (ncells_out) *ncells_out = 0;
    if (n > 0 && !xy) return fail(GR_ERR_PARAM, "xy", "n*2 positions");
    if (n == 0) return ok();
    gr_fa2_session *s = d3_session(n, xy);
    if (!s) return g_err.code;
    gr_status r = d3_pass(s, 0.9, -90.0, 1.0, 1.0, 520.0, 0, 0, nullptr, cap, ncells_out, node_id,
                          0, nullptr);
    gr_fa2_session_release(s);
    return r ? r : ok();
}

// ---- the frozen layout.force.barnes_hut (barnes_hut.rs run_under, sim.rs tick) ----

// Two FP64 tree sessions over the nodes, sharing the charge session's stream: the charge
// tree over the positions (its position columns are the simulation's x, y) and the collide
// tree over the projected positions (its columns are px, py); the rest of the tick is d3.hip.
struct D3Sim {
    gr_fa2_session *cs = nullptr, *ks = nullptr;
    hipStream_t ks_stream = nullptr;  // the collide session's own stream, restored for release
    D3Pass charge{nullptr, {}, {}, 0, false};
    D3Pass collide{nullptr, {}, {}, 0, false};  // its tree only
    gr::D3SimDev sd;
    // the centring sum runs on its own stream beside the link and charge passes: ev_x marks
    // the positions the last tick left, ev_c the shift it computed
    hipStream_t side = nullptr;
    hipEvent_t ev_x = nullptr, ev_c = nullptr;
    // layout.force.particle_mesh: the many-body and collide passes on the mesh instead
    bool mesh = false;
    gr::PmD3Dev pm;
    bool k_valid = false;  // the spectrum in pm.ks was built for (k_step, k_reach)
    int k_step = 0, k_reach = 0;
    ~D3Sim() {
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        if (ev_x) (void)hipEventDestroy(ev_x);
        if (ev_c) (void)hipEventDestroy(ev_c);
        if (cs) (void)hipStreamSynchronize(cs->d.stream);
        if (ks) {
            if (ks_stream) ks->d.stream = ks_stream;
            gr_fa2_session_release(ks);
        }
        if (cs) gr_fa2_session_release(cs);
    }
};

static gr_status d3sim_check(const gr_d3_params &p) {
    const double v[12] = {p.charge_strength, p.theta, p.distance_min, p.distance_max,
                          p.link_distance, p.link_strength_scale, p.collide_radius,
                          p.center_strength, p.velocity_decay, p.alpha_decay, p.alpha_min,
                          p.initial_alpha};
    static const char *names[12] = {"charge_strength", "theta", "distance_min", "distance_max",
                                    "link_distance", "link_strength_scale", "collide_radius",
                                    "center_strength", "velocity_decay", "alpha_decay",
                                    "alpha_min", "initial_alpha"};
    for (int k = 0; k < 12; k++)
        if (!std::isfinite(v[k])) return fail(GR_ERR_PARAM, names[k], "a finite value");
    if (p.theta < 0.0 || p.distance_min < 0.0 || p.distance_max < 0.0)
        return fail(GR_ERR_PARAM, "theta", "theta and distances >= 0");
    return GR_OK;
}

static gr_status d3sim_setup(D3Sim &S, uint32_t n, const uint32_t *src, const uint32_t *dst,
                             const double *strength, size_t m_raw, const gr_d3_params &p) {
    // the simple graph (layout/force/mod.rs:101-120): a pair keeps its first raw edge's
    // strength; each node's row lists its edges in ascending index
    SimpleGraph g;
    build_simple(n, src, dst, m_raw, nullptr, g);
    const uint32_t m = (uint32_t)g.lo.size();
    std::vector<uint32_t> deg(n, 0), roff((size_t)n + 1, 0), slot((size_t)2 * m);
    for (uint32_t e = 0; e < m; e++) {
        deg[g.lo[e]]++;
        deg[g.hi[e]]++;
    }
    for (uint32_t v = 0; v < n; v++) roff[v + 1] = roff[v] + deg[v];
    std::vector<uint32_t> fillp(roff.begin(), roff.end() - 1);
    for (uint32_t e = 0; e < m; e++) {
        slot[2 * (size_t)e] = fillp[g.lo[e]]++;
        slot[2 * (size_t)e + 1] = fillp[g.hi[e]]++;
    }
    // link.rs geometry
    std::vector<double> ld(m), ls(m), lb(m);
    for (uint32_t e = 0; e < m; e++) {
        const double sv = strength ? strength[g.raw[e]] : 1.0;
        const double dlo = (double)deg[g.lo[e]], dhi = (double)deg[g.hi[e]];
        ld[e] = p.link_distance / std::fmax(0.4, sv);
        ls[e] = std::fmin(0.7, p.link_strength_scale * sv);
        lb[e] = dlo / (dlo + dhi);
    }
    // seed.rs: node i at 12 sqrt(i + 1) along the golden-angle spiral, libm's cos and sin
    std::vector<double> xy((size_t)n * 2);
    for (uint32_t i = 0; i < n; i++) {
        const double r = 12.0 * std::sqrt((double)i + 1.0);
        const double a = (double)i * 2.399963229728653;
        xy[2 * (size_t)i] = gr_libm::cos(a) * r;
        xy[2 * (size_t)i + 1] = gr_libm::sin(a) * r;
    }
    if (!(S.cs = d3_session(n, xy.data()))) return g_err.code;
    if (!(S.ks = d3_session(n, xy.data()))) return g_err.code;
    S.ks_stream = S.ks->d.stream;
    S.ks->d.stream = S.cs->d.stream;
    S.charge.s = S.cs;
    S.collide.s = S.ks;
    gr::D3SimDev &sd = S.sd;
    gr_fa2_session *c = S.cs;
    sd.n = n;
    sd.m = m;
    sd.x = (double *)c->d.pos[0];
    sd.y = (double *)c->d.pos[1];
    sd.px = (double *)S.ks->d.pos[0];
    sd.py = (double *)S.ks->d.pos[1];
    sd.stream = c->d.stream;
    sd.seed = 0;  // session.rs SEED
    sd.decay = p.velocity_decay;
    sd.center_strength = p.center_strength;
    sd.reach2 = (2.0 * p.collide_radius) * (2.0 * p.collide_radius);
    const size_t nn = n ? n : 1, mm = m ? m : 1;
    uint32_t *lo = nullptr, *hi = nullptr, *ro = nullptr, *sl = nullptr;
    double *dld = nullptr, *dls = nullptr, *dlb = nullptr;
    gr_status r;
    if ((r = c->alloc(&sd.vx, nn * 8)) || (r = c->alloc(&sd.vy, nn * 8)) ||
        (r = c->alloc(&sd.dl, nn * 16)) || (r = c->alloc(&sd.sh, 2 * mm * 16)) ||
        (r = c->alloc(&sd.shift, 16)) || (r = c->alloc(&lo, mm * 4)) ||
        (r = c->alloc(&hi, mm * 4)) || (r = c->alloc(&ro, (nn + 1) * 4)) ||
        (r = c->alloc(&sl, 2 * mm * 4)) || (r = c->alloc(&dld, mm * 8)) ||
        (r = c->alloc(&dls, mm * 8)) || (r = c->alloc(&dlb, mm * 8)))
        return r;
    if (hipStreamCreateWithFlags(&S.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&S.ev_x, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&S.ev_c, hipEventDisableTiming) != hipSuccess)
        return fail(GR_ERR_DEVICE, "hipStreamCreate", "stream or event creation failed");
    HIP_TRY(hipMemsetAsync(sd.vx, 0, nn * 8, sd.stream));
    HIP_TRY(hipMemsetAsync(sd.vy, 0, nn * 8, sd.stream));
    HIP_TRY(hipStreamSynchronize(sd.stream));
    if (m) {
        HIP_TRY(gr_copy_wait(lo, g.lo.data(), (size_t)m * 4, hipMemcpyHostToDevice, sd.stream));
        HIP_TRY(gr_copy_wait(hi, g.hi.data(), (size_t)m * 4, hipMemcpyHostToDevice, sd.stream));
        HIP_TRY(gr_copy_wait(sl, slot.data(), (size_t)2 * m * 4, hipMemcpyHostToDevice, sd.stream));
        HIP_TRY(gr_copy_wait(dld, ld.data(), (size_t)m * 8, hipMemcpyHostToDevice, sd.stream));
        HIP_TRY(gr_copy_wait(dls, ls.data(), (size_t)m * 8, hipMemcpyHostToDevice, sd.stream));
        HIP_TRY(gr_copy_wait(dlb, lb.data(), (size_t)m * 8, hipMemcpyHostToDevice, sd.stream));
    }
    HIP_TRY(gr_copy_wait(ro, roff.data(), ((size_t)n + 1) * 4, hipMemcpyHostToDevice, sd.stream));
    sd.lo = lo;
    sd.hi = hi;
    sd.roff = ro;
    sd.slot = sl;
    sd.ld = dld;
    sd.ls = dls;
    sd.lb = dlb;
    gr::D3Dev &cg = S.charge.g;
    cg.theta = p.theta;
    cg.theta2 = p.theta * p.theta;
    cg.dmin2 = p.distance_min * p.distance_min;
    cg.dmax2 = p.distance_max * p.distance_max;
    cg.charge = p.charge_strength;
    cg.seed = 0;
    return GR_OK;
}

// mesh.rs Mesh::new on the device: fft.rs Plan's twiddle tables from libm's cos and sin,
// the transform and spectrum grids, and a fresh grid (the identity order, collide.rs
// Grid::new).
static gr_status pmd3_setup(D3Sim &S) {
    gr_fa2_session *c = S.cs;
    gr::PmD3Dev &m = S.pm;
    const uint32_t n = S.sd.n, side = gr_pm_side_for(n);
    m.side = side;
    m.bits = 0;
    while ((1u << m.bits) < side) m.bits++;
    m.nparts = std::min<uint32_t>(gr_blocks(n, GR_BLOCK), 256);
    uint32_t buckets = 4;
    while (buckets < 2 * n) buckets <<= 1;
    m.bucket_bits = 0;
    while ((1u << m.bucket_bits) < buckets) m.bucket_bits++;
    m.dmin2 = S.charge.g.dmin2;
    m.dmax2 = S.charge.g.dmax2;
    const size_t p2 = (size_t)side * side, nn = n ? n : 1;
    double2 *tf = nullptr, *ti = nullptr;
    gr_status r;
    if ((r = c->alloc(&tf, side * 16)) || (r = c->alloc(&ti, side * 16)) ||
        (r = c->alloc(&m.a, p2 * 16)) || (r = c->alloc(&m.b, p2 * 16)) ||
        (r = c->alloc(&m.ks, p2 * 16)) || (r = c->alloc(&m.fr, sizeof(gr::PmD3Frame))) ||
        (r = c->alloc(&m.part, (size_t)m.nparts * 32)) || (r = c->alloc(&m.origin, 16)) ||
        (r = c->alloc(&m.order, nn * 4)) || (r = c->alloc(&m.slot, nn * 4)) ||
        (r = c->alloc(&m.cell, nn * 4)) || (r = c->alloc(&m.fxy, nn * 16)) ||
        (r = c->alloc(&m.keys, nn * 8)) || (r = c->alloc(&m.keys_alt, nn * 8)) ||
        (r = c->alloc(&m.vals, nn * 4)) || (r = c->alloc(&m.vals_alt, nn * 4)) ||
        (r = c->alloc(&m.sort_tmp, gr::radix_tmp_bytes((uint32_t)nn))) ||
        (r = c->alloc(&m.cstart, (p2 + 1) * 4)) ||
        (r = c->alloc(&m.bstart, ((size_t)buckets + 1) * 4)) || (r = c->alloc(&m.at, nn * 16)))
        return r;
    // fft.rs Plan::new: root k = (cos, sin)(-tau k / side); forward[k] = root[(k mod half) *
    // side / (2 half)], half the largest power of two <= max(k, 1); inverse conjugated
    std::vector<double2> root(side / 2), fwd(side), inv(side);
    const double tau = 6.283185307179586;
    for (uint32_t k = 0; k < side / 2; k++) {
        const double angle = -tau * (double)k / (double)side;
        root[k] = make_double2(gr_libm::cos(angle), gr_libm::sin(angle));
    }
    for (uint32_t k = 0; k < side; k++) {
        uint32_t half = 1;
        while (half * 2 <= std::max<uint32_t>(k, 1)) half *= 2;
        fwd[k] = root[(k % half) * (side / (2 * half))];
        inv[k] = make_double2(fwd[k].x, -fwd[k].y);
    }
    HIP_TRY(gr_copy_wait(tf, fwd.data(), side * 16, hipMemcpyHostToDevice, c->d.stream));
    HIP_TRY(gr_copy_wait(ti, inv.data(), side * 16, hipMemcpyHostToDevice, c->d.stream));
    m.tw_fwd = tf;
    m.tw_inv = ti;
    std::vector<uint32_t> iota(nn);
    for (uint32_t i = 0; i < nn; i++) iota[i] = i;
    HIP_TRY(gr_copy_wait(m.order, iota.data(), nn * 4, hipMemcpyHostToDevice, c->d.stream));
    HIP_TRY(gr_copy_wait(m.slot, iota.data(), nn * 4, hipMemcpyHostToDevice, c->d.stream));
    S.mesh = true;
    S.k_valid = false;
    return GR_OK;
}

// One tick (sim.rs tick): alpha decays, then link, many-body, center, collide, integrate.
static gr_status d3sim_tick(D3Sim &S, double alpha, uint32_t tick) {
    gr::D3SimDev &sd = S.sd;
    // the centring sum reads only the positions the last tick left: beside link and charge
    HIP_TRY(hipEventRecord(S.ev_x, sd.stream));
    HIP_TRY(hipStreamWaitEvent(S.side, S.ev_x, 0));
    gr::d3sim_center_sum(sd, S.side);
    HIP_TRY(hipEventRecord(S.ev_c, S.side));
    gr::d3sim_link(sd, alpha, tick);
    gr_status r;
    if (S.mesh) {
        // particle_mesh/charge.rs: the mesh's frame over the positions; no frame, no pass
        gr::pmd3_frame(sd, S.pm);
        gr::PmD3Frame f;
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(&f, S.pm.fr, sizeof(f), hipMemcpyDeviceToHost, sd.stream));
        HIP_TRY(hipStreamSynchronize(sd.stream));
        if (f.valid) {
            if (f.cells < 2 || (uint32_t)f.cells > S.pm.side || f.reach < 0)
                return fail(GR_ERR_DEVICE, "mesh", "a frame outside the mesh");
            gr::pmd3_deposit(sd, S.pm, (uint32_t)f.cells);
            if (!S.k_valid || S.k_step != f.step || S.k_reach != f.reach) {
                gr::pmd3_spectrum(S.pm, f.h, f.reach, sd.stream);
                S.k_valid = true;
                S.k_step = f.step;
                S.k_reach = f.reach;
            }
            gr::pmd3_solve_read(sd, S.pm, (uint32_t)f.cells, S.charge.g.charge * alpha);
            gr::d3sim_merge(sd, sd.dl);
        }
    } else {
        // many-body over the quadtree of the positions (bounds of the moved positions first)
        gr::launch_bounds(S.cs->d);
        if ((r = S.charge.tree()) || (r = S.charge.alloc(false))) return r;
        S.charge.g.alpha = alpha;
        S.charge.g.tick = tick;
        if ((r = S.charge.bodies())) return r;
        gr::d3_launch_walk(S.charge.g);
        gr::d3sim_merge(sd, S.charge.g.out);
    }
    // center, then collide over the quadtree of the projected positions
    HIP_TRY(hipStreamWaitEvent(sd.stream, S.ev_c, 0));
    gr::d3sim_center_project(sd);
    if (S.mesh) {
        // particle_mesh/collide.rs: the hashed grid; no pass at zero reach
        if (sd.reach2 != 0.0) {
            gr::pmd3_collide(sd, S.pm, tick);
            gr::d3sim_merge(sd, sd.dl);
        }
        gr::d3sim_integrate(sd);
        HIP_TRY(hipGetLastError());
        return GR_OK;
    }
    gr::launch_bounds(S.ks->d);
    if ((r = S.collide.tree())) return r;
    gr::D3CollideTree t;
    t.C = S.collide.g.C;
    t.n = sd.n;
    t.wc = S.collide.g.wc;
    t.cb = S.collide.g.cb;
    t.order = S.collide.g.order;
    t.sx = S.collide.g.sx;
    t.sy = S.collide.g.sy;
    gr::d3sim_collide(sd, t, tick);
    gr::d3sim_merge(sd, sd.dl);
    gr::d3sim_integrate(sd);
    HIP_TRY(hipGetLastError());
    return GR_OK;
}

void gr_d3_params_default(gr_d3_params *out) {
    if (!out) return;
    out->charge_strength = -90.0;
    out->theta = 0.9;
    out->distance_min = 1.0;
    out->distance_max = 520.0;
    out->link_distance = 60.0;
    out->link_strength_scale = 0.15;
    out->collide_radius = 16.0;
    out->center_strength = 1.0;
    out->velocity_decay = 1.0 - 0.42;
    out->alpha_decay = 0.06;
    out->alpha_min = 0.001;
    out->initial_alpha = 1.0;
}

// The resident simulation: ForceSession (layout/force/session.rs:118-270) over D3Sim.
struct gr_d3_session {
    D3Sim S;
    gr_d3_params p;
    double alpha = 1.0, alpha_target = 0.0;
    uint32_t tick = 0;
};

static gr_status d3_edges_check(uint32_t n, const uint32_t *source, const uint32_t *target,
                                size_t m) {
    if (n >= 0xFFFFFFFFu) return fail(GR_ERR_CAPACITY, "nodes", "node count fits the u32 index space");
    if (m >= 0xFFFFFFFFull) return fail(GR_ERR_CAPACITY, "edges", "edge count fits the u32 index space");
    if (m > 0 && (!source || !target)) return fail(GR_ERR_PARAM, "edges", "source and target columns");
    for (size_t e = 0; e < m; e++)
        if (source[e] >= n || target[e] >= n)
            return fail(GR_ERR_PARAM, "edges", "edge endpoints below the node count");
    return GR_OK;
}

gr_d3_session *gr_d3_session_create(uint32_t n, const uint32_t *source, const uint32_t *target,
                                    const double *strength, size_t m,
                                    const gr_d3_params *params) {
    if (d3_edges_check(n, source, target, m)) return nullptr;
    gr_d3_params p;
    gr_d3_params_default(&p);
    if (params) p = *params;
    if (d3sim_check(p)) return nullptr;
    gr_d3_session *h = new (std::nothrow) gr_d3_session;
    if (!h) {
        fail(GR_ERR_DEVICE, "host", "out of host memory");
        return nullptr;
    }
    h->p = p;
    h->alpha = p.initial_alpha;
    if (n > 0 && d3sim_setup(h->S, n, source, target, strength, m, p)) {
        delete h;
        return nullptr;
    }
    ok();
    return h;
}

void gr_d3_session_release(gr_d3_session *h) { delete h; }

gr_status gr_d3_session_step(gr_d3_session *h, uint32_t ticks, double *alpha_out,
                             int32_t *settled_out) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    gr_status r;
    for (uint32_t t = 0; t < ticks; t++) {
        // sim.rs tick: alpha moves toward its target (simulation.js:45)
        h->alpha += (h->alpha_target - h->alpha) * h->p.alpha_decay;
        if (h->S.sd.n && (r = d3sim_tick(h->S, h->alpha, h->tick))) return r;
        h->tick++;
    }
    if (h->S.sd.n) HIP_TRY(hipStreamSynchronize(h->S.sd.stream));
    if (alpha_out) *alpha_out = h->alpha;
    // session.rs report: settled when alpha is below alpha_min and no target holds it up
    if (settled_out) *settled_out = h->alpha < h->p.alpha_min && h->alpha_target == 0.0;
    return ok();
}

gr_status gr_d3_session_state(gr_d3_session *h, double *x, double *y, double *vx, double *vy) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    const size_t bytes = (size_t)h->S.sd.n * 8;
    if (!bytes) return ok();
    HIP_TRY(hipStreamSynchronize(h->S.sd.stream));
    if (x) HIP_TRY(gr_copy_wait(x, h->S.sd.x, bytes, hipMemcpyDeviceToHost, h->S.sd.stream));
    if (y) HIP_TRY(gr_copy_wait(y, h->S.sd.y, bytes, hipMemcpyDeviceToHost, h->S.sd.stream));
    if (vx) HIP_TRY(gr_copy_wait(vx, h->S.sd.vx, bytes, hipMemcpyDeviceToHost, h->S.sd.stream));
    if (vy) HIP_TRY(gr_copy_wait(vy, h->S.sd.vy, bytes, hipMemcpyDeviceToHost, h->S.sd.stream));
    return ok();
}

gr_status gr_d3_session_time(gr_d3_session *h, uint32_t reps, double *ms_out,
                             uint32_t *cells_out) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    if (!ms_out || !cells_out) return fail(GR_ERR_PARAM, "out", "two times and two counts");
    D3Sim &S = h->S;
    if (!h->tick || !S.sd.n) return fail(GR_ERR_PARAM, "session", "a session stepped at least once");
    if (reps == 0) reps = 1;
    gr_status r;
    if (S.mesh) {
        // the centring chain (it rewrites only the shift the next tick recomputes) and the
        // hashed grid's pushes over the last tick's grid order (scratch deltas only)
        if (S.sd.reach2 == 0.0)
            return fail(GR_ERR_UNSUPPORTED, "session", "a mesh session with a collide reach");
        if ((r = time_median(S.sd.stream, reps, [&] { gr::d3sim_center_sum(S.sd, S.sd.stream); },
                             &ms_out[0])) ||
            (r = time_median(S.sd.stream, reps,
                             [&] { gr::pmd3_collide_gather(S.sd, S.pm, h->tick - 1); },
                             &ms_out[1])))
            return r;
        cells_out[0] = S.pm.side * S.pm.side;
        cells_out[1] = 1u << S.pm.bucket_bits;
        return ok();
    }
    gr::D3CollideTree t;
    t.C = S.collide.g.C;
    t.n = S.sd.n;
    t.wc = S.collide.g.wc;
    t.cb = S.collide.g.cb;
    t.order = S.collide.g.order;
    t.sx = S.collide.g.sx;
    t.sy = S.collide.g.sy;
    // both walks write only their scratch deltas: the state is left as the last tick left it
    if ((r = time_median(S.cs, reps, [&] { gr::d3_launch_walk(S.charge.g); }, &ms_out[0])) ||
        (r = time_median(S.cs, reps, [&] { gr::d3sim_collide(S.sd, t, h->tick - 1); },
                         &ms_out[1])))
        return r;
    cells_out[0] = S.charge.g.C;
    cells_out[1] = S.collide.g.C;
    return ok();
}

double gr_d3_session_alpha(const gr_d3_session *h) { return h ? h->alpha : NAN; }

gr_status gr_d3_session_reheat(gr_d3_session *h, double alpha) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    // session.rs reheat: ALPHA's range, finite and 0..=1, taken exactly
    if (!(alpha >= 0.0 && alpha <= 1.0)) return fail(GR_ERR_PARAM, "alpha", "alpha: finite, 0..=1");
    h->alpha = alpha;
    return ok();
}

gr_status gr_d3_session_set_alpha_target(gr_d3_session *h, double target) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    // session.rs set_alpha_target: finite, 0 <= target < 1
    if (!(target >= 0.0 && target < 1.0))
        return fail(GR_ERR_PARAM, "alpha_target", "alpha_target: finite, 0..<1");
    h->alpha_target = target;
    return ok();
}

// One mesh pass over given positions: a session of n edge-free nodes on a fresh mesh whose
// position (or projected position) columns are set to xy.
static gr_status pm_pass_session(uint32_t n, const double *xy, double radius, double dmin,
                                 double dmax, bool projected, gr_d3_session **out) {
    gr_d3_params p;
    gr_d3_params_default(&p);
    p.collide_radius = radius;
    p.distance_min = dmin;
    p.distance_max = dmax;
    gr_d3_session *h = gr_d3_session_create(n, nullptr, nullptr, nullptr, 0, &p);
    if (!h) return g_err.code;
    *out = h;
    gr_status r = gr_d3_session_with_particle_mesh(h);
    if (r) return r;
    std::vector<double> x(n), y(n);
    for (uint32_t i = 0; i < n; i++) {
        x[i] = xy[2 * (size_t)i];
        y[i] = xy[2 * (size_t)i + 1];
    }
    gr::D3SimDev &sd = h->S.sd;
    HIP_TRY(gr_copy_wait(projected ? sd.px : sd.x, x.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                         sd.stream));
    HIP_TRY(gr_copy_wait(projected ? sd.py : sd.y, y.data(), (size_t)n * 8, hipMemcpyHostToDevice,
                         sd.stream));
    return GR_OK;
}

gr_status gr_pm_d3_charge(uint32_t n, const double *xy, double charge, double alpha,
                          double distance_min, double distance_max, double *out,
                          double *frame_out) {
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (frame_out)
        for (int k = 0; k < 7; k++) frame_out[k] = 0.0;
    if (!std::isfinite(charge) || !std::isfinite(alpha) || !(distance_min >= 0.0) ||
        !(distance_max >= 0.0) || !std::isfinite(distance_max))
        return fail(GR_ERR_PARAM, "charge", "finite charge, alpha and distances >= 0");
    if (n == 0) return ok();
    for (size_t q = 0; q < (size_t)n * 2; q++) out[q] = 0.0;
    gr_d3_session *h = nullptr;
    gr_status r = pm_pass_session(n, xy, 16.0, distance_min, distance_max, false, &h);
    if (!r) {
        D3Sim &S = h->S;
        gr::pmd3_frame(S.sd, S.pm);
        gr::PmD3Frame f;
        // the session's stream does not wait for the null stream's copies: read on it
        if (hipMemcpyAsync(&f, S.pm.fr, sizeof(f), hipMemcpyDeviceToHost, S.sd.stream) !=
                hipSuccess ||
            hipStreamSynchronize(S.sd.stream) != hipSuccess) {
            r = fail(GR_ERR_DEVICE, "mesh", "frame readback");
        } else if (f.valid) {
            gr::pmd3_deposit(S.sd, S.pm, (uint32_t)f.cells);
            gr::pmd3_spectrum(S.pm, f.h, f.reach, S.sd.stream);
            gr::pmd3_solve_read(S.sd, S.pm, (uint32_t)f.cells, charge * alpha);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(S.sd.stream) != hipSuccess ||
                gr_copy_wait(out, S.sd.dl, (size_t)n * 16, hipMemcpyDeviceToHost,
                             S.sd.stream) != hipSuccess)
                r = fail(GR_ERR_DEVICE, "mesh", "the mesh pass failed on the device");
            if (frame_out) {
                frame_out[0] = f.step;
                frame_out[1] = f.h;
                frame_out[2] = f.ox;
                frame_out[3] = f.oy;
                frame_out[4] = f.cells;
                frame_out[5] = f.reach;
                frame_out[6] = S.pm.side;
            }
        }
    }
    gr_d3_session_release(h);
    return r ? r : ok();
}

gr_status gr_pm_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed,
                           uint32_t tick, double *out, uint32_t *order_out) {
    if (n > 0 && (!xy || !out)) return fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (!std::isfinite(radius) || radius < 0.0)
        return fail(GR_ERR_PARAM, "radius", "a finite radius >= 0");
    if (n == 0) return ok();
    for (size_t q = 0; q < (size_t)n * 2; q++) out[q] = 0.0;
    gr_d3_session *h = nullptr;
    gr_status r = pm_pass_session(n, xy, radius, 1.0, 520.0, true, &h);
    if (!r && h->S.sd.reach2 != 0.0) {
        D3Sim &S = h->S;
        S.sd.seed = seed;
        gr::pmd3_collide(S.sd, S.pm, tick);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(S.sd.stream) != hipSuccess ||
            gr_copy_wait(out, S.sd.dl, (size_t)n * 16, hipMemcpyDeviceToHost,
                         S.sd.stream) != hipSuccess ||
            (order_out &&
             gr_copy_wait(order_out, S.pm.order, (size_t)n * 4, hipMemcpyDeviceToHost,
                          S.sd.stream) != hipSuccess))
            r = fail(GR_ERR_DEVICE, "mesh", "the collide pass failed on the device");
    }
    gr_d3_session_release(h);
    return r ? r : ok();
}

gr_status gr_d3_session_with_particle_mesh(gr_d3_session *h) {
    if (!h) return fail(GR_ERR_INVALID_HANDLE, "session", "a live session");
    if (h->S.mesh) return ok();  // a mesh is made once; the reference's builder replaces it
    if (!h->S.sd.n) {
        h->S.mesh = true;
        return ok();
    }
    gr_status r = pmd3_setup(h->S);
    return r ? r : ok();
}

static gr_status d3_layout(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, bool mesh, float *out_x, float *out_y,
                           double *ms_ticks_out);

gr_status gr_d3_layout_run(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, float *out_x, float *out_y, double *ms_ticks_out) {
    return d3_layout(n, source, target, strength, m, params, ticks, false, out_x, out_y,
                     ms_ticks_out);
}

gr_status gr_pm_layout_run(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, float *out_x, float *out_y, double *ms_ticks_out) {
    return d3_layout(n, source, target, strength, m, params, ticks, true, out_x, out_y,
                     ms_ticks_out);
}

static gr_status d3_layout(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, bool mesh, float *out_x, float *out_y,
                           double *ms_ticks_out) {
    if (n >= 0xFFFFFFFFu) return fail(GR_ERR_CAPACITY, "nodes", "node count fits the u32 index space");
    auto zero = [&]() {
        if (out_x) std::memset(out_x, 0, (size_t)n * 4);
        if (out_y) std::memset(out_y, 0, (size_t)n * 4);
    };
    if (ms_ticks_out) *ms_ticks_out = 0.0;
    if (n > 0 && (!out_x || !out_y)) return fail(GR_ERR_PARAM, "out", "two output columns");
    gr_status r = d3_edges_check(n, source, target, m);
    gr_d3_params p;
    gr_d3_params_default(&p);
    if (params) p = *params;
    if (!r) r = d3sim_check(p);
    if (r) {
        zero();
        return r;
    }
    if (n == 0) return ok();
    gr_d3_session *h = gr_d3_session_create(n, source, target, strength, m, &p);
    if (!h) {
        zero();
        return g_err.code;
    }
    if (mesh && (r = gr_d3_session_with_particle_mesh(h))) {
        gr_d3_session_release(h);
        zero();
        return r;
    }
    std::vector<double> x(n), y(n);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ms_ticks_out && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
                         hipEventRecord(e0, h->S.sd.stream) != hipSuccess))
        r = fail(GR_ERR_DEVICE, "hipEvent", "event creation failed");
    if (!r) r = gr_d3_session_step(h, ticks, nullptr, nullptr);
    if (!r && ms_ticks_out) {
        float ms = 0.f;
        if (hipEventRecord(e1, h->S.sd.stream) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            r = fail(GR_ERR_DEVICE, "hipEvent", "event timing failed");
        *ms_ticks_out = ms;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!r) r = gr_d3_session_state(h, x.data(), y.data(), nullptr, nullptr);
    gr_d3_session_release(h);
    if (r) {
        zero();
        return r;
    }
    // planar_points (layout/force/mod.rs:47-59): refused if any coordinate is not finite
    for (uint32_t i = 0; i < n; i++)
        if (!std::isfinite(x[i]) || !std::isfinite(y[i])) {
            zero();
            return fail(GR_ERR_NON_FINITE, "node.x", "finite coordinates");
        }
    for (uint32_t i = 0; i < n; i++) {
        out_x[i] = (float)x[i];
        out_y[i] = (float)y[i];
    }
    return ok();
}

gr_status gr_layout_run(const char *layout_id, uint32_t n, const uint32_t *source,
                        const uint32_t *target, size_t m, const uint8_t *params,
                        size_t params_len, const gr_fa2_ext *ext, float *out_x, float *out_y,
                        float *out_z) {
    // index.rs: the u32 index space holds at most u32::MAX - 1 nodes; no output buffer of
    // this length exists to clear
    if (n >= 0xFFFFFFFFu) return fail(GR_ERR_CAPACITY, "nodes", "node count fits the u32 index space");
    auto zero = [&]() {
        if (out_x) std::memset(out_x, 0, (size_t)n * 4);
        if (out_y) std::memset(out_y, 0, (size_t)n * 4);
        if (out_z) std::memset(out_z, 0, (size_t)n * 4);
    };
    if (!layout_id) {
        zero();
        return fail(GR_ERR_UNKNOWN_LAYOUT, "layout_id", "a registered layout id");
    }
    std::string id(layout_id);
    if (id == "layout.force.barnes_hut" || id == "layout.force.particle_mesh") {
        // registered with no published parameters (registry/layouts.rs:92-97 and
        // registry/force.rs:291-298); the frozen ForceParams, 112 ticks; the edges' strength
        // column arrives as ext->edge_weights
        if (params_len != 0) {
            zero();
            return fail(GR_ERR_PARAMS_MALFORMED, "params",
                        "one little-endian f64 per published parameter, and no other bytes");
        }
        if (ext && ext->struct_size != sizeof(gr_fa2_ext)) {
            zero();
            return fail(GR_ERR_PARAM, "ext", "struct_size must equal sizeof(gr_fa2_ext)");
        }
        // of the extensions, the strength column and the opening angle apply here
        if (ext && (ext->start || ext->precision != GR_PREC_FP64 ||
                    (ext->repulsion != GR_REP_DEFAULT && ext->repulsion != GR_REP_TREE))) {
            zero();
            return fail(GR_ERR_UNSUPPORTED, "ext",
                        "the force layouts take edge_weights and theta only");
        }
        gr_d3_params p;
        gr_d3_params_default(&p);
        if (ext && ext->theta >= 0.0) p.theta = ext->theta;
        return d3_layout(n, source, target, ext ? ext->edge_weights : nullptr, m, &p, 112,
                         id == "layout.force.particle_mesh", out_x, out_y, nullptr);
    }
    if (id == "layout.forceatlas2.forcesim") {
        // registered with no published parameters (registry/forceatlas2_forcesim.rs:67-72):
        // the empty buffer is the defaults, any byte is malformed; no gr_fa2_ext applies
        if (params_len != 0) {
            zero();
            return fail(GR_ERR_PARAMS_MALFORMED, "params",
                        "one little-endian f64 per published parameter, and no other bytes");
        }
        if (ext) {
            zero();
            return fail(GR_ERR_UNSUPPORTED, "ext", "no extension applies to this layout");
        }
        return gr_forcesim_run(n, source, target, m, nullptr, out_x, out_y, out_z);
    }
    uint32_t dim;
    int rep;
    if (id == "layout.forceatlas2") {
        dim = 2;
        rep = GR_REP_DENSE;
    } else if (id == "layout.forceatlas2.barnes_hut") {
        dim = 2;
        rep = GR_REP_TREE;
    } else if (id == "layout.forceatlas2.3d") {
        dim = 3;
        rep = GR_REP_DENSE;
    } else {
        zero();
        return fail(GR_ERR_UNKNOWN_LAYOUT, "layout_id", "a registered layout id");
    }
    gr_fa2_params p;
    gr_status r = gr_fa2_params_decode(params, params_len, &p);
    if (r) {
        zero();
        return r;
    }
    p.dim = dim;
    gr_fa2_ext x;
    gr_fa2_ext_default(&x);
    if (ext) {
        if (ext->struct_size != sizeof(gr_fa2_ext)) {
            zero();
            return fail(GR_ERR_PARAM, "ext", "struct_size must equal sizeof(gr_fa2_ext)");
        }
        x = *ext;
    }
    if (x.repulsion == GR_REP_DEFAULT) x.repulsion = rep;
    if (!out_x || !out_y || (dim == 3 && !out_z)) {
        zero();
        return fail(GR_ERR_PARAM, "output", "one f32 column per coordinate");
    }
    std::vector<double> pos((size_t)n * dim);
    if (n > 0) {
        gr_fa2_session *s = gr_fa2_session_create(n, source, target, m, &p, &x);
        if (!s) {
            zero();
            return g_err.code;
        }
        uint32_t done = 0;
        r = gr_fa2_session_step(s, p.max_iter, &done);
        if (!r) r = gr_fa2_session_positions(s, pos.data());
        gr_fa2_session_release(s);
        if (r) {
            zero();
            return r;
        }
    }
    // layout/forceatlas2.rs:86-88: the check reads the first `dim` values of the flattened
    // rows, i.e. node 0's coordinates.
    for (uint32_t a = 0; a < dim && n > 0; a++)
        if (!std::isfinite(pos[a])) {
            zero();
            return fail(GR_ERR_NON_FINITE, "node.x", "finite positions");
        }
    for (uint32_t i = 0; i < n; i++) {
        out_x[i] = (float)pos[(size_t)i * dim];
        out_y[i] = (float)pos[(size_t)i * dim + 1];
        if (dim == 3) out_z[i] = (float)pos[(size_t)i * dim + 2];
    }
    return ok();
}

gr_status gr_last_error(const char **name, const char **rule) {
    if (name) *name = g_err.name.c_str();
    if (rule) *rule = g_err.rule.c_str();
    return g_err.code;
}

const char *gr_version(void) { return "graph_render_fa2 0.1.0 (gfx950)"; }

}  // extern "C"
