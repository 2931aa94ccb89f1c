/*
 * CPU oracle for the ForceAtlas2 / Barnes-Hut hot path: TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's f64 Rust FA2 (graph-core), used by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * It is never linked into, loaded by or called from the product library.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math), so every f64
 * expression rounds exactly as the Rust source does.
 *
 * Parity pins (what checks each part of this restatement):
 *   gro_mulberry32_* / gro_initial_positions .. fa2-nx-reference.jsonl (starts feed every
 *                                               line), Mulberry32 published algorithm
 *   gro_fmix64 / gro_jiggle ..................  MurmurHash3 fmix64 closed-form vectors,
 *                                               jiggle range/antisymmetry invariants
 *   gro_gate_edges ...........................  fa2-nx-reference.jsonl (graph feeds line)
 *   gro_fa2_* dense 2D .......................  fa2-nx-reference.jsonl, <=1e-7 of extent
 *   gro_fa2_* dense 3D .......................  parity unpinned (no 3D vectors are held);
 *                                               pinned only by 2D/3D structural identity
 *   gro_tree_* quadtree/octree ...............  brute-force PR-tree definition on small
 *                                               inputs, reference quadtree invariants
 *   BH repulsion .............................  theta=0 equals dense within 1e-12, theta
 *                                               error bounds of the reference BH tests
 */
#ifndef GRAPH_RENDER_FA2_ORACLE_H
#define GRAPH_RENDER_FA2_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rng (crates/graph-core/src/synthetic.rs:37-50, rng.rs:23-84) ---- */
double gro_mulberry32_next(uint32_t *state);
void gro_initial_positions(uint32_t n, uint32_t seed, uint32_t dim, double *out);
uint64_t gro_fmix64(uint64_t k);
double gro_jiggle(uint32_t seed, uint32_t tick, uint32_t pass, uint32_t i, uint32_t j);

/* ---- gate model (synthetic.rs:92-170, stage/topology.rs:26-46) ----
 * Writes the raw edge list of seeded_model(*, count, *) in record order; returns its
 * length. `cap` must be >= 2*count + count/20 + 1. */
size_t gro_gate_edges(uint32_t count, uint32_t *source, uint32_t *target, size_t cap);
size_t gro_gate_model(uint32_t count, uint32_t *source, uint32_t *target, double *strength,
                      size_t cap);

/* ---- FA2 state (layout/forceatlas2/state.rs, state/schedule.rs, state/barnes_hut.rs) ---- */
typedef struct gro_fa2 gro_fa2;

/* repulsion: 0 = dense pair loop, 1 = Barnes-Hut tree with opening angle^2 theta2 */
gro_fa2 *gro_fa2_new(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m_raw,
                     const double *weights_raw, uint32_t dim, double jitter_tolerance,
                     double scaling_ratio, double gravity, uint32_t seed, const double *start,
                     int repulsion, double theta2);
void gro_fa2_free(gro_fa2 *s);
double gro_fa2_iterate(gro_fa2 *s);
uint32_t gro_fa2_run(gro_fa2 *s, uint32_t max_iter);
void gro_fa2_positions(const gro_fa2 *s, double *out);           /* n*dim row-major */
void gro_fa2_set_positions(gro_fa2 *s, const double *in);       /* n*dim row-major */
void gro_fa2_mass(const gro_fa2 *s, double *out);
/* one force evaluation on the current positions, no state change: bit 1 attraction,
 * bit 2 repulsion, bit 4 gravity. out is n*dim row-major. */
void gro_fa2_forces(gro_fa2 *s, uint32_t parts, double *out);
/* swing, traction, speed, speed_efficiency, iter_no */
void gro_fa2_scalars(const gro_fa2 *s, double *out5);
uint32_t gro_fa2_simple_edges(const gro_fa2 *s, uint32_t *lo, uint32_t *hi);
void gro_fa2_set_theta2(gro_fa2 *s, double theta2, int repulsion);
void gro_fa2_set_params(gro_fa2 *s, double jitter_tolerance, double scaling_ratio,
                        double gravity, uint32_t seed);

/* ---- quadtree / octree (layout/force/quadtree*.rs, octree*.rs) ---- */
typedef struct gro_tree gro_tree;
gro_tree *gro_tree_build(uint32_t n, uint32_t dim, const double *xs, const double *ys,
                         const double *zs);
void gro_tree_free(gro_tree *t);
uint32_t gro_tree_ncells(const gro_tree *t);
/* bounds: 2*dim per cell (lo axes then hi axes); cell arrays of ncells; order of n */
void gro_tree_cells(const gro_tree *t, double *bounds, uint32_t *skip, uint32_t *start,
                    uint32_t *end);
void gro_tree_order(const gro_tree *t, uint32_t *order);
/* each cell's node id: the shape slot the insertion created it as (quadtree.rs:64-88) */
void gro_tree_node_ids(const gro_tree *t, uint32_t *out);
/* bodies of the tree for per-point masses: mass then dim centre coordinates per cell */
/* The reference walk replayed per tree position and counted as a scan over groups of `group`
 * consecutive positions: steps (the union of each group's visit sets), visits, accepted
 * cells, leaf pairs (out[0..3]). */
void gro_walk_counts(uint32_t n, uint32_t dim, const double *xs, const double *ys,
                     const double *zs, const double *mass, double theta2, uint32_t group,
                     uint64_t *out);
void gro_tree_bodies(const gro_tree *t, const double *mass, const double *xs, const double *ys,
                     const double *zs, double *out);

/* ---- d3's many-body force over the quadtree (layout/force/barnes_hut/charge.rs) ----
 * parity: theta 0 equals the direct d3 pair sum (closed form); node ids are the insertion's
 * own shape slots */
/* ---- the frozen layout.force.barnes_hut (barnes_hut.rs, sim.rs tick) ----
 * p: charge, theta, distance_min, distance_max, link_distance, link_strength_scale,
 * collide_radius, center_strength, velocity_decay, alpha_decay, initial_alpha */
/* mesh_side 0: layout.force.barnes_hut; otherwise layout.force.particle_mesh on a mesh of
 * that side (its charge and collide passes from pm_d3_oracle.c) */
void gro_d3sim_run(uint32_t n, const uint32_t *src, const uint32_t *dst, const double *strength,
                   size_t m_raw, uint32_t ticks, const double *p, uint32_t seed,
                   uint32_t mesh_side, double *x_out, double *y_out, double *vx_out,
                   double *vy_out);

/* ---- layout.force.particle_mesh's mesh passes, d3's law (pm_d3_oracle.c) ---- */
typedef struct gro_pmd3 gro_pmd3;
gro_pmd3 *gro_pmd3_new(uint32_t n, int side);
void gro_pmd3_free(gro_pmd3 *m);
int gro_pmd3_charge(gro_pmd3 *m, const double *x, const double *y, double dmin, double dmax,
                    double strength, double *out, double *frame_out);
int gro_pmd3_collide(gro_pmd3 *m, const double *px, const double *py, double radius,
                     uint32_t seed, uint32_t tick, double *out);
void gro_pmd3_spectrum(const gro_pmd3 *m, double *re, double *im);
int gro_pmd3_charge_once(uint32_t n, const double *xy, int side, const uint32_t *order,
                         double charge, double alpha, double dmin, double dmax, double *out,
                         double *frame_out, double *spec_re, double *spec_im);
int gro_pmd3_collide_once(uint32_t n, const double *xy, double radius, uint32_t seed,
                          uint32_t tick, double *out, uint32_t *order_out);
/* libm 0.2.16's cos and sin (libm_trig.c); NaN past |x| = 2^20 pi/2 */
double gro_libm_cos(double x);
double gro_libm_sin(double x);
/* collide.rs node_delta of every node over the quadtree of (px, py): out n*2, node order */
void gro_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed, uint32_t tick,
                    double *out);
void gro_d3_charge_bh(uint32_t n, const double *xy, double theta, double charge, double alpha,
                      double dmin, double dmax, uint32_t seed, uint32_t tick, double *out);

/* ---- particle mesh (pm_oracle.c) ---- */
double gro_pm_h(int step);
void gro_fft2(double *re, double *im, int side, int sign);
int gro_pm_forces(uint32_t n, const double *xy, const double *mass, double k, int side,
                  double *out, double *frame_out);
int gro_pm_charge(uint32_t n, const double *xy, int side, double charge, double alpha,
                  double dmin, double dmax, double *out, double *frame_out);

#ifdef __cplusplus
}
#endif
#endif
