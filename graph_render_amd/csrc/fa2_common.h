// Device definitions shared by the FA2 kernels (fa2_kernels.hip) and the reverse-mode
// kernels (adjoint.hip): the walk record of a tree cell, the coincidence jiggle, the pair
// force and the early-exit/halt test.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <type_traits>

#include "fa2_kernels.h"

namespace gr {

// The FP32 precisions keep each centre as an unevaluated pair hi + lo of floats (about 48
// bits), so the offset of a nearby body from it is exact to float rounding even where
// coordinates are large and the cell is small. FP64 keeps the reference's one double.
template <typename T, int D>
struct CentreLo {
    T clo[D];
};
template <int D>
struct CentreLo<double, D> {};

template <typename T, int D>
struct WalkCell : CentreLo<T, D> {
    T c[D];
    T mass;
    T w2;  // opening threshold: w² (FP64) or w²/θ² (FP32), see open_threshold
    uint32_t skip, start, end;
    // dword index of `mass` in the record (the lo centre, if any, then the centre)
    static constexpr uint32_t MASS_DWORD =
        (std::is_same<T, double>::value ? 0u : (uint32_t)D) + (uint32_t)(D * sizeof(T) / 4);
    // dword index of `start` (`end` follows it): after mass, w2 and skip
    static constexpr uint32_t START_DWORD = MASS_DWORD + (uint32_t)(2 * sizeof(T) / 4) + 1;
};
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winvalid-offsetof"
template <typename C>
constexpr bool mass_dword_ok() {
    return offsetof(C, mass) == 4 * C::MASS_DWORD && offsetof(C, start) == 4 * C::START_DWORD &&
           offsetof(C, end) == 4 * (C::START_DWORD + 1);
}
static_assert(mass_dword_ok<WalkCell<float, 2>>() && mass_dword_ok<WalkCell<float, 3>>() &&
                  mass_dword_ok<WalkCell<double, 2>>() && mass_dword_ok<WalkCell<double, 3>>(),
              "MASS_DWORD and START_DWORD name the record's fields");
#pragma clang diagnostic pop

__device__ __forceinline__ uint32_t graph_of(const FaDev &d, uint32_t node) {
    return d.gidn ? d.gidn[node] : 0u;
}

// rng.rs:27-84: MurmurHash3 fmix64, the three-round fold and the 53-bit nudge.
__device__ __forceinline__ uint64_t d_fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__device__ __forceinline__ double d_jiggle_inline(uint32_t seed, uint32_t tick, uint32_t pass,
                                                 uint32_t i, uint32_t j) {
    uint32_t lo = i <= j ? i : j, hi = i <= j ? j : i;
    uint64_t a = d_fmix64((uint64_t)seed ^ ((uint64_t)tick << 32));
    uint64_t b = d_fmix64(a ^ ((uint64_t)pass << 32) ^ (uint64_t)lo);
    uint64_t w = d_fmix64(b ^ ((uint64_t)hi << 32) ^ a) >> 11;
    if (w == (1ull << 52)) w = (1ull << 52) + 1;
    double u = (double)w * (1.0 / (double)(1ull << 53));
    return (u - 0.5) * 1e-6;
}

// The same, kept out of line: the FA2 walks take it on rare branches only.
static __device__ __noinline__ double d_jiggle(uint32_t seed, uint32_t tick, uint32_t pass,
                                               uint32_t i, uint32_t j) {
    return d_jiggle_inline(seed, tick, pass, i, j);
}

// Iteration kernels stop once the run has exited early or a capacity check failed; a
// force evaluation (ignore_done) still runs on an exited state.
__device__ __forceinline__ bool halted(const FaDev &d) {
    const GrState *st = d.st;
    return st->halt != 0 || (!d.ignore_done && st->done != 0);
}

// The first eight words of GrState, read in one go at a kernel's start: the halt test and
// the counts the kernel sizes itself by then cost one memory round trip, not two.
struct StateHead {
    uint32_t iters_done, done, halt, n_finite, n_groups, n_cells, need_cells, pool_used;
};
static_assert(offsetof(GrState, done) == 4 && offsetof(GrState, halt) == 8 &&
                  offsetof(GrState, n_finite) == 12 && offsetof(GrState, n_groups) == 16 &&
                  offsetof(GrState, n_cells) == 20 && offsetof(GrState, need_cells) == 24,
              "StateHead mirrors the head of GrState");
__device__ __forceinline__ StateHead state_head(const FaDev &d) {
    const uint4 *q = (const uint4 *)d.st;
    const uint4 a = q[0], b = q[1];
    // an empty statement that takes all eight words: both loads are issued and waited for
    // here, ahead of the halt branch (a word used only past the branch would otherwise be
    // loaded there, a second round trip)
    asm volatile("" ::"s"(a.x), "s"(a.y), "s"(a.z), "s"(a.w), "s"(b.x), "s"(b.y), "s"(b.z),
                 "s"(b.w));
    return StateHead{a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
}
__device__ __forceinline__ bool halted(const FaDev &d, const StateHead &h) {
    return h.halt != 0 || (!d.ignore_done && h.done != 0);
}

// m_a*m_b/d2*k. FP64 divides exactly as the reference does; the FP32 precisions use the
// hardware reciprocal (about 1 ulp), well inside their 1e-5 force tolerance.
// a * b + c: FP32 fuses it into one rounding (FP32 results are held to a tolerance, not
// to bits); FP64 rounds the product and the sum apart, as the reference's f64 code does.
template <typename T>
__device__ __forceinline__ T mul_add(T a, T b, T c) {
    if constexpr (std::is_same<T, float>::value) {
        return __builtin_fmaf(a, b, c);
    } else {
        return a * b + c;
    }
}

// The hardware reciprocal (FP32 only).
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double fast_rcp(double x) { return 1.0 / x; }

// The walk record's opening threshold (its `w2` field). FP64 keeps the reference's w² and
// tests w² < θ²·d² (barnes_hut.rs); FP32 keeps w²/θ², formed once per cell, and tests it
// against d² (θ = 0 gives an infinite threshold: nothing is accepted, as in FP64).
template <typename T>
__device__ __forceinline__ T open_threshold(double w, double theta2) {
    if constexpr (std::is_same<T, double>::value) {
        return w * w;
    } else {
        return (T)(w * w / theta2);
    }
}
template <typename T>
__device__ __forceinline__ T open_rhs(T d2, T theta2) {
    if constexpr (std::is_same<T, double>::value) {
        return theta2 * d2;
    } else {
        return d2;
    }
}

template <typename T>
__device__ __forceinline__ T pair_force(T mm, T d2, T kr) {
    if constexpr (std::is_same<T, double>::value) {
        return mm / d2 * kr;
    } else {
        return mm * __builtin_amdgcn_rcpf(d2) * kr;
    }
}

// Block reduction of NV values (min, max or sum) in a fixed order; thread 0 gets the result.
template <int NV, int OP>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *lds) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; q++) {
        if (OP == 0) v[q] = gr_wave_sum(v[q]);
        if (OP == 1) v[q] = gr_wave_min(v[q]);
        if (OP == 2) v[q] = gr_wave_max(v[q]);
        if (lane == 0) lds[wave * NV + q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double r = lds[q];
            for (int w = 1; w < GR_BLOCK / GR_WAVE; w++) {
                double o = lds[w * NV + q];
                if (OP == 0) r = r + o;
                if (OP == 1) r = fmin(r, o);
                if (OP == 2) r = fmax(r, o);
            }
            v[q] = r;
        }
    }
    __syncthreads();
}

}  // namespace gr
