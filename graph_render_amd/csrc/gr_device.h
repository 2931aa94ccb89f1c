// Shared device-side definitions for the FA2 / Barnes-Hut HIP path (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GR_WAVE 64
#define GR_BLOCK 256

// One graph's scalars, resident on the device: the adaptive-speed state of
// layout/forceatlas2/state.rs:151-171 (swing, traction, speed, speed_efficiency, iter_no),
// its centroid, extent and root square, and where its tree sits in the session's arena.
// A single-graph session is a batch of one.
struct GrGraph {
    double swing, traction, speed, speed_efficiency;
    double moved;                   // last applied step's sum |dp| (schedule.rs:62-77)
    double mean[3];                 // centroid of the current positions (state.rs:208-220)
    double root_lo[3], root_hi[3];  // root square of the current tree (quadtree.rs:157-195)
    double box_lo[3], box_hi[3];    // extent of the finite positions
    uint32_t iter_no;               // jiggle tick, incremented once per iteration
    uint32_t iters_done;            // iterations applied to this graph
    uint32_t done;                  // set once an iteration moved < 1e-10 (state.rs:122-128)
    uint32_t n_finite;              // its points inside the tree
    uint32_t cell_lo, cell_hi;      // its cells in the arena: [cell_lo, cell_hi)
    uint32_t has_root, pad;
};

// Session-wide bookkeeping of the launch-ahead iteration loop.
struct GrState {
    uint32_t iters_done;  // iteration launches that moved at least one graph
    uint32_t done;        // every graph has exited early
    uint32_t halt;        // a capacity check failed: the iteration is not applied
    uint32_t n_finite;    // points inside the trees (all graphs)
    uint32_t n_groups;    // distinct-key groups of the current trees
    uint32_t n_cells;     // cells of the current trees
    uint32_t need_cells;  // cells the current iteration asked for (capacity check)
    uint32_t pool_used;   // deep-group scratch used
    uint32_t pool_overflow;
    uint32_t any_deep;
    uint32_t lcell_used;  // deep-group local cells used
    uint32_t span_count;  // unused
    uint32_t live;        // graphs not yet done
    uint32_t pad[3];
};

__device__ __forceinline__ uint64_t gr_ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t gr_lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t gr_lanemask_lt() {
    uint32_t l = __lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

template <typename T>
__device__ __forceinline__ T gr_shfl_xor(T v, int m) {
    return __shfl_xor(v, m, GR_WAVE);
}

// Wave-wide sum in a fixed butterfly order: every lane ends with the same value, and the
// association depends only on lane positions, so repeated runs give the same bits.
__device__ __forceinline__ double gr_wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, GR_WAVE);
    return v;
}

__device__ __forceinline__ uint32_t gr_wave_sum_u32(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, GR_WAVE);
    return v;
}

__device__ __forceinline__ double gr_wave_min(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m, GR_WAVE));
    return v;
}

__device__ __forceinline__ double gr_wave_max(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, GR_WAVE));
    return v;
}

__device__ __forceinline__ uint32_t gr_uniform(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// A comparison as a wave mask (bit l: lane l's result; inactive lanes read 0), so tests can
// be combined with scalar mask operations. The codes are the compiler's integer predicates.
enum { GR_CMP_EQ = 32, GR_CMP_ULT = 36, GR_CMP_ULE = 37 };
template <int PRED>
__device__ __forceinline__ uint64_t wave_cmp_u32(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_uicmp(a, b, PRED);
}
// a < b (ordered: false when either is NaN) as a wave mask
__device__ __forceinline__ uint64_t wave_lt(float a, float b) {
    return __builtin_amdgcn_fcmpf(a, b, 4);
}
__device__ __forceinline__ uint64_t wave_lt(double a, double b) {
    return __builtin_amdgcn_fcmp(a, b, 4);
}

// A raw buffer resource over [p, p + bytes): a load through it at a byte offset at or past
// `bytes` returns 0 and never faults, so a tile's loads can all be issued unconditionally
// (straight-line code keeps them in flight together; a per-load bounds branch makes each
// one wait for the last).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t gr_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t gr_load_u32(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0);
}

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8). This
// maps them so that XCD x owns logical blocks [x*q, (x+1)*q): consecutive bodies or rows
// share one XCD's L2. Launch gr_xcd_grid(nb) workgroups; a logical index >= nb is idle.
#define GR_XCDS 8u
__device__ __forceinline__ uint32_t gr_xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = (nb + GR_XCDS - 1) / GR_XCDS;
    return (b % GR_XCDS) * q + b / GR_XCDS;
}
static inline uint32_t gr_xcd_grid(uint32_t nb) {
    return ((nb + GR_XCDS - 1) / GR_XCDS) * GR_XCDS;
}

static inline uint32_t gr_blocks(uint64_t n, uint32_t per) {
    uint64_t b = (n + per - 1) / per;
    return b == 0 ? 1u : (uint32_t)b;
}

// Host copies and fills on the caller's own stream, waited for before they return. Nothing in
// the library goes on the legacy (null) stream: a legacy-stream copy made while another thread
// captures a graph on its stream is refused (hipErrorStreamCaptureImplicit), so sessions on
// different threads would interfere.
static inline hipError_t gr_copy_wait(void *dst, const void *src, size_t bytes,
                                      hipMemcpyKind kind, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
static inline hipError_t gr_fill_wait(void *dst, int value, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const hipError_t e = hipMemsetAsync(dst, value, bytes, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
