// d3's many-body force over the quadtree (the reference's layout.force Barnes-Hut charge
// pass, crates/graph-core/src/layout/force/barnes_hut/charge.rs) on a tree the FA2 build
// made: the d3 bodies, the cells' insertion node ids and the walk (d3.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gr {

// One cell as the walk reads it (charge.rs:86-94: centre, opening threshold, point count,
// the index past its subtree, its first point), plus the node id its jiggle is keyed on.
struct D3Body {
    double comx, comy, open;
    uint32_t count, skip, start, nid;
};
static_assert(sizeof(D3Body) == 40, "a d3 body is 40 bytes");

struct D3Dev {
    uint32_t n = 0;   // points
    uint32_t nf = 0;  // finite points: tree positions [0, nf)
    uint32_t C = 0;   // cells
    // the tree (an FP64 2D FA2 session's arena): cell records, the point order, the
    // positions in tree order and the exported cell bounds (x0, y0, x1, y1)
    const void *wc = nullptr;
    const uint32_t *order = nullptr;
    const double *sx = nullptr, *sy = nullptr;
    const double *cb = nullptr;
    // the pass's own buffers
    D3Body *body = nullptr;
    uint32_t *diff = nullptr;    // C + 1: +1 past each internal cell, -1 at its skip
    uint32_t *depth = nullptr;   // C: prefix of diff, then each cell's depth
    uint32_t *amin = nullptr;    // C: the smallest point index of the cell's subtree
    uint32_t *bmin = nullptr;    // C: the insertion that created the cell (see d3.hip)
    uint64_t *keys = nullptr, *keys_alt = nullptr;  // C: sort keys (depth, then pusher)
    uint32_t *vals = nullptr, *vals_alt = nullptr;  // C: cells in sort order
    uint32_t *lvl = nullptr;     // C + 3: deepest internal depth, internal count, depth starts
    void *tmp = nullptr;         // scan and sort scratch
    double *out = nullptr;       // n * 2: the velocity each node gains
    uint32_t *nid_out = nullptr; // C, or null
    double theta = 0.9, theta2 = 0.81;
    double dmin2 = 1.0, dmax2 = 520.0 * 520.0, charge = -90.0, alpha = 1.0;
    uint32_t seed = 0, tick = 0;
    hipStream_t stream = nullptr;
};

size_t d3_tmp_bytes(uint32_t C);
// Leaf bodies, the structure of every body, the leaves' smallest and pushing points and the
// depth difference array (d.diff cleared here).
void d3_launch_cells(const D3Dev &d);
// Depths, then the cells sorted by depth with the level table in d.lvl (deepest internal
// depth, internal cell count, the first sorted slot of each depth); returns the sorted cells.
const uint32_t *d3_launch_depth(const D3Dev &d);
// The internal cells of one depth, [off, off + cnt) of the sorted cells, from their children.
void d3_launch_level(const D3Dev &d, const uint32_t *bucket, uint32_t off, uint32_t cnt);
// Node ids: each cell's rank by (the point whose insertion pushed it, preorder index).
void d3_launch_ids(const D3Dev &d);
// The walk: every node's delta into d.out.
void d3_launch_walk(const D3Dev &d);

// ---- the rest of the Barnes-Hut force simulation's tick (barnes_hut/sim.rs) ----

struct D3SimDev {
    uint32_t n = 0, m = 0;  // nodes, simple edges
    double *x = nullptr, *y = nullptr;    // positions (the charge tree session's columns)
    double *vx = nullptr, *vy = nullptr;  // velocities
    double *px = nullptr, *py = nullptr;  // projected positions (the collide tree's columns)
    const uint32_t *lo = nullptr, *hi = nullptr;  // simple edges, lo < hi
    const uint32_t *roff = nullptr;  // n + 1: each node's row of slots (its edges, ascending)
    const uint32_t *slot = nullptr;  // 2m: edge e's slot in lo's row, then in hi's
    const double *ld = nullptr, *ls = nullptr, *lb = nullptr;  // link distance, strength, bias
    double *sh = nullptr;     // 2m slots x (x, y): each slot's share of its edge's force
    double *dl = nullptr;     // 2n: deltas of the collide pass
    double *shift = nullptr;  // 2: the centring shift of this tick
    uint32_t seed = 0;
    double decay = 0.58, center_strength = 1.0, reach2 = 1024.0;
    hipStream_t stream = nullptr;
};

// The collide tree as the walk reads it (an FP64 2D arena with exported bounds).
struct D3CollideTree {
    uint32_t C = 0, n = 0;
    const void *wc = nullptr;
    const double *cb = nullptr;
    const uint32_t *order = nullptr;
    const double *sx = nullptr, *sy = nullptr;  // projected positions in tree order
};

// link.rs: every edge's force from the displaced positions, then every node's shares added
// to its velocity (as sum + 0).
void d3sim_link(const D3SimDev &s, double alpha, uint32_t tick);
// step.rs merge: deltas (2n, node order) added to the velocities as d + 0.
void d3sim_merge(const D3SimDev &s, const double *deltas);
// sim.rs center_shift: the mean position (summed in node order) times the strength, into
// s.shift; reads only the positions, so it may run on `stream` beside the link and charge
// passes.
void d3sim_center_sum(const D3SimDev &s, hipStream_t stream);
// sim.rs center, then collide.rs prepare: positions -= shift; px = x + vx, py = y + vy.
void d3sim_center_project(const D3SimDev &s);
// collide.rs: every node's push from the projected positions within the reach, into s.dl.
void d3sim_collide(const D3SimDev &s, const D3CollideTree &t, uint32_t tick);
// sim.rs integrate: v *= decay; position += v.
void d3sim_integrate(const D3SimDev &s);

// ---- layout.force.particle_mesh's mesh passes (pm_d3.hip) ----

struct PmD3Frame {
    double h, ox, oy;
    int32_t step, cells, reach, valid;
};

struct PmD3Dev {
    uint32_t side = 0, bits = 0;      // mesh side (a power of two) and its log2
    uint32_t nparts = 0;              // bounds partials
    uint32_t bucket_bits = 0;         // the hashed grid: 2^bucket_bits buckets
    double dmin2 = 1.0, dmax2 = 0.0;  // the law's cutoffs, squared
    const double2 *tw_fwd = nullptr, *tw_inv = nullptr;  // side: fft.rs Plan tables
    double2 *a = nullptr, *b = nullptr;  // side * side: density / field, transform scratch
    double2 *ks = nullptr;               // side * side: the kernel's spectrum
    PmD3Frame *fr = nullptr;
    double *part = nullptr;    // nparts * 4 bounds partials
    double *origin = nullptr;  // 2: the hash origin
    uint32_t *order = nullptr, *slot = nullptr;  // n: the grid (slot -> node, node -> slot)
    uint32_t *cell = nullptr;  // n: each slot's stencil cell (side * side: none)
    double *fxy = nullptr;     // 2n: each slot's stencil fractions
    uint64_t *keys = nullptr, *keys_alt = nullptr;  // n: sort keys
    uint32_t *vals = nullptr, *vals_alt = nullptr;  // n
    void *sort_tmp = nullptr;
    uint32_t *cstart = nullptr;  // side * side + 1: first sorted slot of each stencil cell
    uint32_t *bstart = nullptr;  // 2^bucket_bits + 1: first slot of each bucket
    double2 *at = nullptr;       // n: projected positions in grid order
};

// The frame over the positions (bounds partials folded, frame.rs place) into m.fr.
void pmd3_frame(const D3SimDev &s, const PmD3Dev &m);
// The deposit into m.a's first `cells` rows, each cell in ascending grid slot.
void pmd3_deposit(const D3SimDev &s, const PmD3Dev &m, uint32_t cells);
// The kernel's spectrum at this rung and reach into m.ks (m.b as scratch).
void pmd3_spectrum(const PmD3Dev &m, double h, int reach, hipStream_t stream);
// Forward transform, spectrum product, inverse, and the read into s.dl (node order).
void pmd3_solve_read(const D3SimDev &s, const PmD3Dev &m, uint32_t cells, double strength);
// The hashed grid over s.px, s.py (m.order rebuilt) and every node's pushes into s.dl.
void pmd3_collide(const D3SimDev &s, const PmD3Dev &m, uint32_t tick);
// collide.rs's pushes alone over the grid order the last pmd3_collide sorted (writes s.dl)
void pmd3_collide_gather(const D3SimDev &s, const PmD3Dev &m, uint32_t tick);

}  // namespace gr
