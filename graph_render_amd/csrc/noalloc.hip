// gr_pm_charge_noalloc: the particle-mesh charge pass of gr_pm_charge with every device
// buffer carved out of one file-scope array, for callers that may not allocate device
// memory. No hipMalloc / hipFree: the array is the library's own, one copy per GPU, held
// until the process ends.
//
// Limits: n <= GR_NOALLOC_MAX_NODES, mesh side <= GR_NOALLOC_MAX_SIDE; the array is sized
// for both at once. Calls on one GPU share its array (and keep the last call's contents),
// so they are serialised by a process-wide lock; calls that overlap on one GPU from
// several processes race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/graph_render_fa2.h"
#include "fa2_kernels.h"
#include "sort_scan.h"

extern "C" gr_status gr_internal_fail(gr_status code, const char *name, const char *rule);

namespace {

constexpr size_t SCRATCH_BYTES = (size_t)64 << 20;
constexpr uint32_t CHUNK = 2048;  // node chunk of the bounds partials, as in a session
constexpr int MAX_DEVICES = 64;

// Device memory of this entry point: planar and interleaved positions, masses, the sort's
// keys and values, its histograms, the bounds partials, the graph head, the mesh grids and
// the frame record. 256-byte aligned.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_lock;
void *g_base[MAX_DEVICES] = {};  // this array's address on each device, once looked up

struct Carve {
    char *base;
    size_t off = 0;
    template <typename P>
    void take(P *&out, size_t bytes) {
        out = (P *)(base + off);
        off += (bytes + 255) & ~(size_t)255;
    }
    void take(void *&out, size_t bytes) {
        out = (void *)(base + off);
        off += (bytes + 255) & ~(size_t)255;
    }
};

// Lays a charge pass of n nodes on a side x side mesh out from `base` (null: sizing only);
// returns the bytes used.
size_t layout(char *base, uint32_t n, uint32_t side, gr::FaDev &d) {
    Carve c{base};
    const size_t nn = n ? n : 1;
    const uint32_t nchunks = n ? (n + CHUNK - 1) / CHUNK : 1;
    const size_t P2 = (size_t)side * side;
    const size_t tiles = (size_t)(side / 16) * (side / 16);
    c.take(d.pos[0], nn * 8);
    c.take(d.pos[1], nn * 8);
    c.take(d.pi, nn * 2 * 8);
    c.take(d.pk, nn * 4 * 8);
    c.take(d.mass, nn * 8);
    c.take(d.rep[0], nn * 8);
    c.take(d.rep[1], nn * 8);
    c.take(d.keys, nn * 8);
    c.take(d.keys_alt, nn * 8);
    c.take(d.vals, nn * 4);
    c.take(d.vals_alt, nn * 4);
    c.take(d.sort_tmp, gr::radix_tmp_bytes((uint32_t)nn));
    c.take(d.bpart, (size_t)nchunks * 10 * 8);
    c.take(d.gr, sizeof(GrGraph));
    c.take(d.st, sizeof(GrState));
    c.take(d.gnode, 2 * 4);
    c.take(d.gchunk, 2 * 4);
    c.take(d.chunk, ((size_t)nchunks + 1) * 4);
    c.take(d.chunk_g, (size_t)nchunks * 4);
    c.take(d.pm_a, P2 * sizeof(float2));
    c.take(d.pm_b, P2 * sizeof(float2));
    c.take(d.pm_k, P2 * sizeof(float2));
    c.take(d.pm_acc, P2 * sizeof(unsigned long long));
    c.take(d.pm_tw, (size_t)(side / 2) * sizeof(float2));
    c.take(d.pm_tiles, (tiles + 2) * 4);
    c.take(d.pm_halo, tiles * 33 * 8);
    c.take(d.pm_fr, gr::pm_frame_bytes());
    d.nchunks = nchunks;
    return c.off;
}

// Positions from the interleaved upload to the planar columns, unit masses, the one-graph
// tables, a fresh graph head and state, and a cleared frame record (so the kernel spectrum
// is rebuilt for this call's mesh and law).
__global__ __launch_bounds__(GR_BLOCK) void k_na_setup(const gr::FaDev d, uint32_t frame_words) {
    const uint32_t n = d.n;
    const double *xy = (const double *)d.pi;
    const uint32_t stride = gridDim.x * GR_BLOCK;
    const uint32_t t0 = blockIdx.x * GR_BLOCK + threadIdx.x;
    for (uint32_t i = t0; i < n; i += stride) {
        ((double *)d.pos[0])[i] = xy[2 * (size_t)i];
        ((double *)d.pos[1])[i] = xy[2 * (size_t)i + 1];
        ((double *)d.mass)[i] = 1.0;
    }
    for (uint32_t c = t0; c <= d.nchunks; c += stride) {
        const uint64_t v = (uint64_t)c * CHUNK;
        d.chunk[c] = v < n ? (uint32_t)v : n;
        if (c < d.nchunks) d.chunk_g[c] = 0;
    }
    for (uint32_t w = t0; w < frame_words; w += stride) ((uint32_t *)d.pm_fr)[w] = 0u;
    if (t0 == 0) {
        d.gnode[0] = 0;
        d.gnode[1] = n;
        d.gchunk[0] = 0;
        d.gchunk[1] = d.nchunks;
        GrGraph g = {};
        g.swing = g.traction = g.speed = g.speed_efficiency = 1.0;
        *d.gr = g;
        GrState s = {};
        s.live = 1;
        *d.st = s;
    }
}

// The field (planar) back to row-major pairs, into the upload buffer.
__global__ __launch_bounds__(GR_BLOCK) void k_na_out(const gr::FaDev d) {
    const uint32_t i = blockIdx.x * GR_BLOCK + threadIdx.x;
    if (i >= d.n) return;
    double *xy = (double *)d.pi;
    xy[2 * (size_t)i] = ((const double *)d.rep[0])[i];
    xy[2 * (size_t)i + 1] = ((const double *)d.rep[1])[i];
}

// This device's copy of `scratch`, looked up on the first call made with it current.
gr_status device_base(char **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, "device", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        return gr_internal_fail(GR_ERR_DEVICE, "device", "a device index below 64");
    if (!g_base[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return gr_internal_fail(GR_ERR_DEVICE, "hipGetSymbolAddress",
                                    e != hipSuccess ? hipGetErrorString(e) : "null address");
        g_base[dev] = p;
    }
    *out = (char *)g_base[dev];
    return GR_OK;
}

}  // namespace

extern "C" {

gr_status gr_pm_charge_noalloc(uint32_t n, const double *xy, uint32_t side, double charge,
                               double alpha, double distance_min, double distance_max,
                               double *out, double *frame_out) {
    if (n > 0 && (!xy || !out)) return gr_internal_fail(GR_ERR_PARAM, "xy", "n*2 positions and an output");
    if (frame_out)
        for (int k = 0; k < 7; k++) frame_out[k] = 0.0;
    if (!std::isfinite(charge) || !std::isfinite(alpha) || !(distance_min >= 0.0) ||
        !(distance_max >= 0.0) || std::isnan(distance_max))
        return gr_internal_fail(GR_ERR_PARAM, "charge", "finite charge and alpha, distances >= 0");
    if (side == 0) side = gr_pm_side_for(n);
    if (side < 64 || side > 4096 || (side & (side - 1)))
        return gr_internal_fail(GR_ERR_PARAM, "side", "a power of two from 64 to 4096");
    if (n > GR_NOALLOC_MAX_NODES)
        return gr_internal_fail(GR_ERR_CAPACITY, "n",
                                "at most GR_NOALLOC_MAX_NODES (262144) nodes without allocation");
    if (side > GR_NOALLOC_MAX_SIDE)
        return gr_internal_fail(GR_ERR_CAPACITY, "side",
                                "a mesh side of at most GR_NOALLOC_MAX_SIDE (1024) without allocation");
    if (n == 0) return gr_internal_fail(GR_OK, nullptr, nullptr);

    std::lock_guard<std::mutex> hold(g_lock);
    char *base = nullptr;
    gr_status r = device_base(&base);
    if (r) return r;
    gr::FaDev d;
    d.n = n;
    d.dim = 2;
    d.precision = gr::PREC_FP64;
    d.repulsion = gr::REP_PM;
    d.ngraphs = 1;
    d.pm_side = side;
    d.pm_law = 1;
    d.pm_dmin2 = distance_min * distance_min;
    d.pm_dmax2 = distance_max * distance_max;
    d.pm_strength = charge * alpha;
    d.pm_fix = std::ldexp(1.0, 60 - (int)std::ceil(std::log2((double)n + 1.0)));  // unit charges
    d.stream = hipStreamPerThread;  // the calling thread's own stream: none is created
    if (layout(base, n, side, d) > SCRATCH_BYTES)  // the limits above keep every call inside
        return gr_internal_fail(GR_ERR_CAPACITY, "scratch", "the call's buffers fit the array");

    // the twiddles as a session makes them (host cos/sin), so the transforms are the same
    float2 tw[GR_NOALLOC_MAX_SIDE / 2];
    gr::pm_twiddles(side, tw);
    hipError_t e = gr_copy_wait(d.pm_tw, tw, (size_t)(side / 2) * sizeof(float2),
                                hipMemcpyHostToDevice, d.stream);
    if (e == hipSuccess)
        e = gr_copy_wait(d.pi, xy, (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice,
                         d.stream);
    if (e != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, "hipMemcpy", hipGetErrorString(e));
    const uint32_t frame_words = (uint32_t)(gr::pm_frame_bytes() / 4);
    hipLaunchKernelGGL(k_na_setup, dim3(gr_blocks(n, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream, d,
                       frame_words);
    gr::launch_pm_charge(d);
    hipLaunchKernelGGL(k_na_out, dim3(gr_blocks(n, GR_BLOCK)), dim3(GR_BLOCK), 0, d.stream, d);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(d.stream);
    if (e != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, "pm", hipGetErrorString(e));
    e = gr_copy_wait(out, d.pi, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, d.stream);
    if (e != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, "hipMemcpy", hipGetErrorString(e));
    if (frame_out) {
        unsigned char fr[256];
        const size_t fb = gr::pm_frame_bytes();
        if (fb > sizeof(fr)) return gr_internal_fail(GR_ERR_DEVICE, "pm", "frame record size");
        e = gr_copy_wait(fr, d.pm_fr, fb, hipMemcpyDeviceToHost, d.stream);
        if (e != hipSuccess) return gr_internal_fail(GR_ERR_DEVICE, "hipMemcpy", hipGetErrorString(e));
        gr::pm_frame_read(fr, frame_out);
    }
    return gr_internal_fail(GR_OK, nullptr, nullptr);
}

}  // extern "C"
