// Device buffers of one FA2 session and the host-side launchers of its kernels
// (fa2_kernels.hip). Everything here is resident in HBM for the session's life; nothing is
// allocated per iteration.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "gr_device.h"

namespace gr {

enum Repulsion : int { REP_TREE = 0, REP_DENSE = 1, REP_PM = 2, REP_NONE = 3 };
enum Precision : int { PREC_FP64 = 0, PREC_FP32 = 1, PREC_FP32_KAHAN = 2 };

// Rows with more CSR entries than this are summed by k_long_rows (one workgroup per row).
constexpr uint32_t LONG_ROW = 128;

// CSR entries whose neighbour positions (and weights) one k_finish round stages in LDS:
// 24 KiB of dynamic LDS per workgroup. A workgroup of k_finish is also given about this
// many rows + entries, so that its rows take one round.
constexpr uint32_t FIN_LDS = 24576;
constexpr uint32_t fin_stage(uint32_t elem_bytes, uint32_t dim, bool weighted) {
    return FIN_LDS / (elem_bytes * (dim + (weighted ? 1u : 0u)));
}

// Force parts of one evaluation (gr_fa2_session_forces): the reference's iterate() order.
enum Parts : uint32_t { PART_ATTRACTION = 1, PART_REPULSION = 2, PART_GRAVITY = 4 };

struct FaDev {
    uint32_t n = 0, dim = 2;
    int precision = PREC_FP64;
    int repulsion = REP_TREE;
    double theta2 = 0.8 * 0.8;
    double scaling_ratio = 2.0, gravity = 1.0, jitter_tolerance = 1.0;
    uint32_t seed = 0;
    int weighted = 0;
    uint32_t m2 = 0;  // CSR entries (2 per simple edge)

    // state, element type T (double or float) chosen by precision
    void *pos[3] = {};
    void *pi = nullptr;  // positions again, interleaved (x, y[, z]) per node, for gathers
    void *pk = nullptr;  // (x, y, [z,] mass) per node in 4 slots: one line for the tree gather
    void *mass = nullptr;
    void *u[3] = {};
    void *rep[3] = {};
    uint32_t *row_off = nullptr, *nbr = nullptr;
    void *w = nullptr;
    // rows longer than LONG_ROW entries (ascending node ids): their attraction is summed by
    // k_long_rows with parallel gathers into LDS, one workgroup per row
    uint32_t nlong = 0;
    uint32_t *long_rows = nullptr;
    void *long_acc = nullptr;  // dim sums per long row
    GrState *st = nullptr;

    // batch layout: graph b owns nodes [gnode[b], gnode[b+1]) and node chunks
    // [gchunk[b], gchunk[b+1]); chunk c holds nodes [chunk[c], chunk[c+1]) of graph chunk_g[c]
    uint32_t ngraphs = 1;
    uint32_t levels = 31;  // key levels below the root (the graph id sits above them)
    uint32_t key_bits = 64;  // key bits the sort orders: levels, graph id, non-finite flag
    GrGraph *gr = nullptr;
    uint32_t *gnode = nullptr, *gidn = nullptr;  // gidn: node -> graph (null for one graph)
    uint32_t nchunks = 0;
    uint32_t *chunk = nullptr, *chunk_g = nullptr, *gchunk = nullptr;
    double *nodeval = nullptr;  // per node: swing and traction terms

    // tree
    uint64_t *keys = nullptr, *keys_alt = nullptr;
    uint32_t *vals = nullptr, *vals_alt = nullptr;
    uint8_t *stall = nullptr;
    uint32_t *order = nullptr;
    uint32_t *leaf = nullptr;  // per tree position: the leaf cell holding it
    void *sp[3] = {};
    void *sm = nullptr;
    uint32_t *flag = nullptr, *gid = nullptr, *dflag = nullptr;
    uint32_t *gstart = nullptr, *gcells = nullptr, *gbase = nullptr, *ginfo = nullptr;
    uint32_t *glocal = nullptr;
    uint64_t *gkey = nullptr;
    uint32_t cap = 0;  // cell capacity
    void *wc = nullptr;
    uint8_t *clevel = nullptr;  // cell depth; 255 for cells summed by the deep builder
    uint32_t *span = nullptr;   // cells whose subtree crosses a centre-of-mass tile
    uint32_t *span_n = nullptr; // per tile: how many of its span slots are used
    uint8_t *span_lev = nullptr; // the level of each span slot's cell
    uint32_t *span_cells = nullptr; // the span cells bucketed by level (32 slots per tile)
    uint32_t span_cap = 0;          // span slots (32 per tile)
    uint32_t *span_kids = nullptr;  // child references per span slot: by cell, then by slot
    double *span_kval = nullptr;    // per span slot: its final children's sums (lane-major)
    uint32_t *span_slot = nullptr;  // per cell: its slot in span_cells (span cells only)
    uint32_t *span_lcnt = nullptr;  // span cells per level [0, 33), then bucket fills [33, 66)
    double *raw = nullptr;      // unnormalised (mass, m*x...) per cell
    double *cbounds = nullptr;  // 2*dim per cell when the session exports the tree
    int export_bounds = 0;

    // deep-group builder pools
    uint32_t pool_cap = 0, lcell_cap = 0;
    int32_t *pool = nullptr;
    int32_t *lc_int = nullptr;
    double *lc_dbl = nullptr;
    double *lc_sum = nullptr;  // unnormalised sums of local cells, (1 + dim) each
    int32_t *chain_next = nullptr;

    // reductions
    uint32_t nparts = 0;          // workgroups of k_finish, a multiple of the XCD count
    uint32_t *part_rows = nullptr;  // k_finish workgroup b: rows [part_rows[b], part_rows[b+1])
    double *bpart = nullptr;  // bounds/mean partials
    double *spart = nullptr;  // swing/traction partials
    double *mpart = nullptr;  // moved partials

    void *scan_tmp = nullptr;
    void *sort_tmp = nullptr;

    // particle mesh (pm.hip): side x side complex FP32 grids
    uint32_t pm_side = 0;
    float2 *pm_a = nullptr;   // density, then field, x along rows
    float2 *pm_b = nullptr;   // transposed grid (y along rows)
    float2 *pm_k = nullptr;   // kernel spectrum, transposed layout, 1/side^2 folded in
    float2 *pm_tw = nullptr;  // twiddles exp(-2 pi i t / side), t < side/2
    unsigned long long *pm_acc = nullptr;  // fixed-point deposit
    uint32_t *pm_tiles = nullptr;  // deposit tiles: first sorted particle of each, (side/16)^2 + 1
    unsigned long long *pm_halo = nullptr;  // deposit tiles: 33 cells beyond each tile
    void *pm_fr = nullptr;    // frame and the rung the spectrum was built for
    double pm_fix = 0.0;      // fixed-point scale of the deposit (a power of two)
    // law: 0 FA2 (mass deposit, G = r/|r|^2, infinite reach, read times k*m_i); 1 d3's
    // many-body (unit charge, G = -r/l with l raised below dmin^2 and cut at dmax^2, reach
    // ceil(dmax/h), read times charge*alpha; particle_mesh/kernel.rs, charge.rs)
    int pm_law = 0;
    double pm_dmin2 = 0.0, pm_dmax2 = 0.0, pm_strength = 0.0;

    int ignore_done = 0;  // set on force evaluations, which run after an early exit too
    // The global sums (centroid, swing, traction, moved) added one term after another in
    // node order, as the reference does, instead of by block reductions: FP64 trajectories
    // are then the reference's bit for bit.
    int seq_sums = 0;
    uint64_t *wstat = nullptr;  // the measuring walk's four counters

    // the schedule (none of it moves a result bit): walk workgroup size, XCD-contiguous
    // walk workgroups, radix-sort keys per thread (a tile is 256 of them)
    uint32_t walk_block = 256;
    int walk_xcd = 1;
    uint32_t walk_group = 64;  // tree positions per walk wave (set per launch from n)
    uint32_t sort_items = 8;

    hipStream_t stream = nullptr;
};

size_t walkcell_bytes(int precision, uint32_t dim);
size_t pm_frame_bytes();
// step, h, origin x, origin y, cells, reach, valid from a frame record (pm.hip)
void pm_frame_read(const void *frame, double *out7);
void pm_twiddles(uint32_t side, float2 *host_out);

// Positions changed from the host: recompute the bounds/mean partials.
void launch_bounds(FaDev &d);
// One full iteration (after launch_bounds or a previous iteration's apply).
void launch_iteration(FaDev &d);
// One force evaluation on the current positions: builds the tree when needed and leaves the
// force in d.u; no state moves (swing/traction are not accumulated, nothing applied).
void launch_forces(FaDev &d, uint32_t parts);
// The tree build alone (after launch_bounds): keys, sort, topology, centre of mass.
void launch_tree(FaDev &d);
// The tree's arena alone (cells, bounds, order), without the centres of mass.
void launch_tree_arena(FaDev &d);
// The walk alone over the current tree, for timing.
void launch_walk(FaDev &d);
// The force finish alone (attraction, gravity and the swing/traction terms over the last
// walk's repulsion), for timing.
void launch_finish(FaDev &d);
// The walk's measuring variant: adds its scan steps, acting lanes, accepted cells and leaf
// pairs to d.wstat (cleared by the caller).
void launch_walk_count(FaDev &d);
// Clears the halt flag left by a failed capacity check.
void launch_clear_halt(FaDev &d);
// One particle-mesh charge pass of a 2D FP64 single-graph state whose planar positions,
// masses and tables are in place: bounds partials, the graph's extent, then the mesh
// (pm.hip), leaving the field times the law's strength in d.rep.
void launch_pm_charge(FaDev &d);

// ---- reverse mode (adjoint.hip) ----

// Scalar adjoints carried backward through a recording, resident on the device.
struct AdjScal {
    double sbar;        // d loss / d speed entering the step being reversed
    double Sbar, Tbar;  // d loss / d cumulative swing and traction totals
    double kbar, gbar;  // d loss / d scaling_ratio, d gravity, summed over steps
    double mubar[3];    // d loss / d centroid of the step being reversed
    double pad[8];
};

constexpr int ADJ_NPART = 8;  // partial sums per block
// One exact accumulator: ADJ_LIMBS - 1 signed 32-bit digit sums (held in 64-bit words) of a
// fixed-point number with its lowest bit at 2^ADJ_BASE, and a count of non-finite terms.
constexpr int ADJ_LIMBS = 9;
constexpr int ADJ_BASE = -180;

struct AdjDev {
    double *pb[3] = {};  // d loss / d positions, one column per axis
    double *ub[3] = {};  // d loss / d force u of the step being reversed
    // exact accumulators (ADJ_LIMBS integer limbs per value, adjoint.hip), so the order in
    // which waves add into them moves no bit
    unsigned long long *cacc = nullptr;  // per tree cell: sum over accepting bodies of
                                         // -k*m_i*J(r)*ub_i
    unsigned long long *pscat = nullptr; // per tree position: pair terms for bodies of
                                         // many-body leaves
    double *wb = nullptr;    // d loss / d weight, per simple edge, summed over steps
    uint32_t *elo = nullptr, *ehi = nullptr;  // simple edges (lo < hi)
    uint32_t m = 0;
    AdjScal *sc = nullptr;
    double *part = nullptr;  // ADJ_NPART doubles per block
    uint32_t nblk = 0;       // blocks of GR_BLOCK nodes (>= node chunks)
    const GrGraph *g0 = nullptr;  // the graph's scalars before the step being reversed
    const GrGraph *g1 = nullptr;  // and after it
};

// Copies n*dim planar positions (one column per axis, the session's element type) into
// the state (planar and interleaved copies).
void launch_adjoint_load(FaDev &d, const void *planar);
// One reversed iteration. Needs the state of the iteration's start with its force u
// evaluated (launch_forces with every part) and a.g0 / a.g1 set; updates a.pb, a.wb and the
// scalar adjoints in a.sc.
void launch_adjoint_step(FaDev &d, AdjDev &a);

}  // namespace gr
