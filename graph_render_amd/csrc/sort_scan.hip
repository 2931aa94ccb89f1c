// Device-wide exclusive scan (u32) and a stable LSD radix sort of (u64 key, u32 value)
// pairs, both hand-written for gfx950: 256-thread blocks of four wave64s, LDS histograms,
// ballot-based stable ranking inside each wave, no library calls.
#include "gr_device.h"
#include "sort_scan.h"

namespace gr {

constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = GR_BLOCK * SCAN_ITEMS;
constexpr uint32_t SCAN_SUMMED_TILES = 4 * GR_BLOCK;  // up to here the down-sweep adds the partials

// Block-wide inclusive scan of one value per thread (256 threads), LDS-based.
__device__ __forceinline__ uint32_t block_inclusive_scan(uint32_t v, uint32_t *lds) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // wave inclusive scan with shuffles
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    if (lane == 63) lds[wave] = v;
    __syncthreads();
    uint32_t add = 0;
    for (uint32_t w = 0; w < wave; w++) add += lds[w];
    __syncthreads();
    return v + add;
}

__device__ __forceinline__ uint32_t load_count(const uint32_t *d_count, uint32_t max_n) {
    if (!d_count) return max_n;
    uint32_t c = *d_count;
    return c < max_n ? c : max_n;
}

__global__ __launch_bounds__(GR_BLOCK) void k_scan_reduce(const uint32_t *in, uint32_t max_n,
                                                          const uint32_t *d_count,
                                                          uint32_t *partials) {
    __shared__ uint32_t lds[8];
    const uint32_t n = load_count(d_count, max_n);
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    const uint32_t tn = base < n ? (n - base < SCAN_TILE ? (uint32_t)(n - base) : SCAN_TILE) : 0;
    const __amdgpu_buffer_rsrc_t rin = gr_rsrc(in + base, tn * 4);
    uint32_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; r++)
        s += gr_load_u32(rin, ((uint32_t)r * GR_BLOCK + threadIdx.x) * 4);
    s = gr_wave_sum_u32(s);
    if ((threadIdx.x & 63u) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partials[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

// Exclusive scan of row[0, cnt) in place by one workgroup, in rounds of 4 * GR_BLOCK values,
// four consecutive per thread: their local scan, then one block scan of the thread totals.
// Returns the grand total (in every thread).
__device__ uint32_t block_scan_row(uint32_t *row, uint32_t cnt, uint32_t *lds) {
    constexpr uint32_t PER = 4;
    const __amdgpu_buffer_rsrc_t r = gr_rsrc(row, cnt * 4);
    uint32_t carry = 0;
    for (uint32_t b = 0; b < cnt; b += PER * GR_BLOCK) {
        const uint32_t i0 = b + PER * threadIdx.x;
        uint32_t v[PER], loc = 0;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) v[k] = gr_load_u32(r, (i0 + k) * 4);
#pragma unroll
        for (uint32_t k = 0; k < PER; k++) {
            const uint32_t t = v[k];
            v[k] = loc;
            loc += t;
        }
        const uint32_t inc = block_inclusive_scan(loc, lds);
        const uint32_t ex = carry + inc - loc;
#pragma unroll
        for (uint32_t k = 0; k < PER; k++)
            if (i0 + k < cnt) row[i0 + k] = ex + v[k];
        // the round's total from the wave totals block_inclusive_scan left in lds, read by
        // every thread before the next round writes them again
        carry += lds[0] + lds[1] + lds[2] + lds[3];
        __syncthreads();
    }
    return carry;
}

// One block scans the per-tile partials (exclusive) and writes the grand total.
__global__ __launch_bounds__(GR_BLOCK) void k_scan_partials(uint32_t *partials, uint32_t nparts,
                                                            uint32_t *d_total) {
    __shared__ uint32_t lds[8];
    const uint32_t total = block_scan_row(partials, nparts, lds);
    if (threadIdx.x == 0 && d_total) *d_total = total;
}

// Wave-wide inclusive scan of one value per lane (shuffles, no barriers).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = threadIdx.x & 63u;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(v, o, 64);
        if (lane >= (uint32_t)o) v += t;
    }
    return v;
}

// Each wave scans its own SCAN_ITEMS * 64 consecutive values of the tile in rounds of 64,
// carrying its running total; one exchange of the wave totals then gives every wave its
// base, so the tile takes one barrier instead of two per round. SUMMED: the partials are the
// tiles' plain sums (no k_scan_partials launch): each workgroup adds those of the tiles
// before it, and the last one writes the grand total.
template <bool SUMMED>
__global__ __launch_bounds__(GR_BLOCK) void k_scan_down(const uint32_t *in, uint32_t *out,
                                                        uint32_t max_n, const uint32_t *d_count,
                                                        const uint32_t *partials,
                                                        uint32_t *d_total) {
    constexpr uint32_t WAVES = GR_BLOCK / GR_WAVE;
    constexpr uint32_t PER_WAVE = SCAN_ITEMS * GR_WAVE;
    __shared__ uint32_t wtot[WAVES];
    __shared__ uint32_t before[WAVES];
    const uint32_t n = load_count(d_count, max_n);
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    const uint32_t tn = base < n ? (n - base < SCAN_TILE ? (uint32_t)(n - base) : SCAN_TILE) : 0;
    const __amdgpu_buffer_rsrc_t rin = gr_rsrc(in + base, tn * 4);
    const uint32_t wave = threadIdx.x / GR_WAVE, lane = threadIdx.x % GR_WAVE;
    uint32_t vin[SCAN_ITEMS];  // the wave's values in flight before the first round
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; r++)
        vin[r] = gr_load_u32(rin, (wave * PER_WAVE + (uint32_t)r * GR_WAVE + lane) * 4);
    uint32_t tile_base;
    if (SUMMED) {
        uint32_t t = 0;
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += GR_BLOCK) t += partials[i];
        t = gr_wave_sum_u32(t);
        if (lane == 0) before[wave] = t;
        __syncthreads();
        tile_base = 0;
#pragma unroll
        for (uint32_t w = 0; w < WAVES; w++) tile_base += before[w];
        if (d_total && threadIdx.x == 0 && blockIdx.x + 1 == gridDim.x)
            *d_total = tile_base + partials[blockIdx.x];
    } else {
        tile_base = partials[blockIdx.x];
    }
    uint32_t run = 0, ex[SCAN_ITEMS];
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; r++) {
        const uint32_t inc = wave_inclusive_scan(vin[r]);
        ex[r] = run + inc - vin[r];
        run += __shfl(inc, 63, 64);
    }
    if (lane == 0) wtot[wave] = run;
    __syncthreads();
    uint32_t wbase = tile_base;
    for (uint32_t w = 0; w < wave; w++) wbase += wtot[w];
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; r++) {
        const uint64_t i = base + wave * PER_WAVE + (uint32_t)r * GR_WAVE + lane;
        if (i < n) out[i] = wbase + ex[r];
    }
}

size_t scan_tmp_bytes(uint32_t max_n) {
    return (size_t)gr_blocks(max_n, SCAN_TILE) * sizeof(uint32_t) + 256;
}

void scan_exclusive(const uint32_t *in, uint32_t *out, uint32_t max_n, const uint32_t *d_count,
                    uint32_t *d_total, void *tmp, hipStream_t stream) {
    uint32_t nb = gr_blocks(max_n, SCAN_TILE);
    uint32_t *partials = (uint32_t *)tmp;
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(GR_BLOCK), 0, stream, in, max_n, d_count,
                       partials);
    if (nb <= SCAN_SUMMED_TILES) {  // each workgroup adds at most 4 partials per thread
        hipLaunchKernelGGL(k_scan_down<true>, dim3(nb), dim3(GR_BLOCK), 0, stream, in, out,
                           max_n, d_count, partials, d_total);
        return;
    }
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(GR_BLOCK), 0, stream, partials, nb,
                       d_total);
    hipLaunchKernelGGL(k_scan_down<false>, dim3(nb), dim3(GR_BLOCK), 0, stream, in, out, max_n,
                       d_count, partials, nullptr);
}

// ------------------------------------------------------------------ radix sort

// RS_ITEMS keys per thread, RS_TILE = 256 * RS_ITEMS keys per tile: 4, 8 (default) or 16,
// chosen per session by its schedule; the sort's result does not depend on it.
constexpr int RS_BITS = 8;
constexpr int RS_BUCKETS = 1 << RS_BITS;
constexpr int RS_DPT = RS_BUCKETS / GR_BLOCK;  // digits per thread in the per-digit steps
static_assert(RS_BUCKETS % GR_BLOCK == 0, "whole digits per thread");

template <int RS_ITEMS>
__global__ __launch_bounds__(GR_BLOCK) void k_rs_hist(const uint64_t *keys, uint32_t n, int shift,
                                                      uint32_t *hist, uint32_t nb) {
    constexpr int RS_TILE = GR_BLOCK * RS_ITEMS;
    __shared__ uint32_t h[RS_BUCKETS];
#pragma unroll
    for (int q = 0; q < RS_DPT; q++) h[q * GR_BLOCK + threadIdx.x] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
    uint32_t dg[RS_ITEMS];  // every key's digit first (all loads in flight), then the counts
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint64_t i = base + (uint64_t)r * GR_BLOCK + threadIdx.x;
        const uint64_t key = keys[i < n ? i : n - 1];  // n >= 1: an unconditional load
        dg[r] = i < n ? (uint32_t)(key >> shift) & (RS_BUCKETS - 1) : RS_BUCKETS;
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++)
        if (dg[r] < RS_BUCKETS) atomicAdd(&h[dg[r]], 1u);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RS_DPT; q++) {
        const uint32_t dgt = q * GR_BLOCK + threadIdx.x;
        hist[(size_t)dgt * nb + blockIdx.x] = h[dgt];
    }
}

// Block d scans row d of the digit-major histogram in place (exclusive) and writes the
// digit's total.
__global__ __launch_bounds__(GR_BLOCK) void k_rs_scan(uint32_t *hist, uint32_t nb,
                                                      uint32_t *totals) {
    __shared__ uint32_t lds[8];
    const uint32_t total = block_scan_row(hist + (size_t)blockIdx.x * nb, nb, lds);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// Stable scatter. A tile's keys are ranked round by round (256 consecutive keys per round),
// inside a wave by ballot matching on the 8 digit bits and across the block's four waves by
// LDS counts, into the tile's digit-sorted order in LDS; the tile is then written out as one
// contiguous run per digit, so global stores are coalesced. The tile's per-digit counts come
// from the (already scanned) block histogram.
// Stable scatter of one digit pass. Wave w takes the tile's keys [w * 512, (w + 1) * 512) in
// eight rounds of 64 and ranks each key among the same-digit keys before it in the wave,
// with wave-private digit counters in LDS (a wave's LDS reads and writes happen in program
// order, so no block barrier is needed between rounds); one prefix over the waves then
// gives each wave's first slot per digit, the keys go to their tile-local slots in LDS in
// input order, and from there to their global slots.
template <int RS_ITEMS>
__global__ __launch_bounds__(GR_BLOCK) void k_rs_scatter(const uint64_t *kin, const uint32_t *vin,
                                                         uint64_t *kout, uint32_t *vout,
                                                         uint32_t n, int shift,
                                                         const uint32_t *hist, uint32_t nb,
                                                         const uint32_t *totals) {
    constexpr int RS_TILE = GR_BLOCK * RS_ITEMS;
    constexpr uint32_t WAVES = GR_BLOCK / GR_WAVE;
    constexpr uint32_t PER_WAVE = RS_ITEMS * GR_WAVE;  // consecutive keys per wave
    __shared__ uint64_t skey[RS_TILE];
    __shared__ uint32_t sval[RS_TILE];
    __shared__ uint32_t gbase[RS_BUCKETS];  // global position of the tile's first key per digit
    __shared__ uint32_t lbase[RS_BUCKETS];  // tile-local position of the same
    __shared__ uint32_t wcnt[WAVES][RS_BUCKETS];  // per wave and digit: count, then first slot
    __shared__ uint32_t lds[8];
    const uint32_t tid = threadIdx.x, wave = tid / GR_WAVE, lane = tid % GR_WAVE;
    const uint32_t b = blockIdx.x;
    const uint64_t tile = (uint64_t)b * RS_TILE;
    const uint32_t tn = tile + RS_TILE <= n ? RS_TILE : (uint32_t)(n - tile);
    // the wave's keys and values, all loads in flight at once (64 consecutive per round)
    uint64_t kk[RS_ITEMS];
    uint32_t vv[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = wave * PER_WAVE + (uint32_t)r * GR_WAVE + lane;
        const uint64_t at = tile + (li < tn ? li : tn - 1);  // tn >= 1: unconditional loads
        kk[r] = kin[at];
        vv[r] = vin[at];
    }
#pragma unroll
    for (uint32_t w = 0; w < WAVES; w++)
#pragma unroll
        for (int q = 0; q < RS_DPT; q++) wcnt[w][q * GR_BLOCK + tid] = 0;
    {
        // thread tid takes digits [tid * RS_DPT, (tid + 1) * RS_DPT): the digits' totals and
        // this tile's counts, scanned locally and then across the block
        uint32_t t[RS_DPT], h[RS_DPT], c[RS_DPT], tsum = 0, csum = 0;
#pragma unroll
        for (int q = 0; q < RS_DPT; q++) {
            const uint32_t dgt = tid * RS_DPT + q;
            t[q] = totals[dgt];
            h[q] = hist[(size_t)dgt * nb + b];
            const uint32_t hn = hist[(size_t)dgt * nb + (b + 1 < nb ? b + 1 : b)];
            c[q] = (b + 1 < nb ? hn : t[q]) - h[q];
            tsum += t[q];
            csum += c[q];
        }
        uint32_t tex = block_inclusive_scan(tsum, lds) - tsum;
        uint32_t cex = block_inclusive_scan(csum, lds) - csum;
#pragma unroll
        for (int q = 0; q < RS_DPT; q++) {
            const uint32_t dgt = tid * RS_DPT + q;
            gbase[dgt] = tex + h[q];
            lbase[dgt] = cex;
            tex += t[q];
            cex += c[q];
        }
    }
    __syncthreads();  // counters cleared (block_inclusive_scan's barriers come before it too)
    const uint64_t below = gr_lanemask_lt();
    uint32_t rank_in_wave[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = wave * PER_WAVE + (uint32_t)r * GR_WAVE + lane;
        const bool valid = li < tn;
        const uint32_t d = (uint32_t)(kk[r] >> shift) & (RS_BUCKETS - 1);
        uint64_t peers = gr_ballot(valid);
#pragma unroll
        for (int bit = 0; bit < RS_BITS; bit++) {
            const uint64_t bb = gr_ballot((d >> bit) & 1u);
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        const uint32_t cnt = (uint32_t)__popcll(peers);
        // the digit's count in this wave so far; the group's last lane then advances it
        const uint32_t before = valid ? wcnt[wave][d] : 0u;
        rank_in_wave[r] = before + rank;
        if (valid && rank == cnt - 1) wcnt[wave][d] = before + cnt;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RS_DPT; q++) {  // per digit: each wave's first slot, in wave order
        const uint32_t dgt = q * GR_BLOCK + tid;
        uint32_t acc = lbase[dgt];
#pragma unroll
        for (uint32_t w = 0; w < WAVES; w++) {
            const uint32_t c = wcnt[w][dgt];
            wcnt[w][dgt] = acc;
            acc += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const uint32_t li = wave * PER_WAVE + (uint32_t)r * GR_WAVE + lane;
        if (li < tn) {
            const uint32_t d = (uint32_t)(kk[r] >> shift) & (RS_BUCKETS - 1);
            const uint32_t pos = wcnt[wave][d] + rank_in_wave[r];
            skey[pos] = kk[r];
            sval[pos] = vv[r];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < tn; i += GR_BLOCK) {
        const uint64_t key = skey[i];
        const uint32_t d = (uint32_t)(key >> shift) & (RS_BUCKETS - 1);
        const uint32_t pos = gbase[d] + (i - lbase[d]);
        kout[pos] = key;
        vout[pos] = sval[i];
    }
}

size_t radix_tmp_bytes(uint32_t n) {  // sized for the smallest tile (the most tiles)
    uint32_t nb = gr_blocks(n, GR_BLOCK * 4);
    return ((size_t)RS_BUCKETS * nb + RS_BUCKETS) * sizeof(uint32_t) + 256;
}

// Sorts bits [0, end_bit) of the keys, rounded up to whole RS_BITS digits, ping-ponging
// between (keys, vals) and (keys_alt, vals_alt): after an even number of digits the result
// is in (keys, vals), after an odd number in the alternates. Nothing is copied.
template <int RS_ITEMS>
static void radix_sort_impl(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                            uint32_t *vals_alt, uint32_t n, int end_bit, void *tmp,
                            hipStream_t stream) {
    constexpr int RS_TILE = GR_BLOCK * RS_ITEMS;
    uint32_t nb = gr_blocks(n, RS_TILE);
    uint32_t *hist = (uint32_t *)tmp;
    uint32_t *totals = hist + (size_t)RS_BUCKETS * nb;
    int passes = (end_bit + RS_BITS - 1) / RS_BITS;
    uint64_t *ki = keys, *ko = keys_alt;
    uint32_t *vi = vals, *vo = vals_alt;
    for (int p = 0; p < passes; p++) {
        int shift = p * RS_BITS;
        hipLaunchKernelGGL(k_rs_hist<RS_ITEMS>, dim3(nb), dim3(GR_BLOCK), 0, stream, ki, n,
                           shift, hist, nb);
        hipLaunchKernelGGL(k_rs_scan, dim3(RS_BUCKETS), dim3(GR_BLOCK), 0, stream, hist, nb,
                           totals);
        hipLaunchKernelGGL(k_rs_scatter<RS_ITEMS>, dim3(nb), dim3(GR_BLOCK), 0, stream, ki, vi,
                           ko, vo, n, shift, hist, nb, totals);
        uint64_t *tk = ki;
        ki = ko;
        ko = tk;
        uint32_t *tv = vi;
        vi = vo;
        vo = tv;
    }
}

void radix_sort_pairs(uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                      uint32_t n, int end_bit, void *tmp, hipStream_t stream, int items) {
    if (n == 0) return;
    if (items == 4)
        radix_sort_impl<4>(keys, vals, keys_alt, vals_alt, n, end_bit, tmp, stream);
    else if (items == 16)
        radix_sort_impl<16>(keys, vals, keys_alt, vals_alt, n, end_bit, tmp, stream);
    else
        radix_sort_impl<8>(keys, vals, keys_alt, vals_alt, n, end_bit, tmp, stream);
}

int radix_digits(int end_bit) { return (end_bit + RS_BITS - 1) / RS_BITS; }
int radix_bits() { return RS_BITS; }

}  // namespace gr
