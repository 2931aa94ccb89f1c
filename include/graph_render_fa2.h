/*
 * graph_render_fa2 — the C-ABI boundary of the MI355X ForceAtlas2 / Barnes-Hut path.
 *
 * It replaces, behind the reference's layout registry, the three ForceAtlas2 entries of
 * graph-core (paths relative to crates/graph-core/src/ of the reference):
 *   "layout.forceatlas2"             layout/forceatlas2.rs:32-41   (dense, dim 2)
 *   "layout.forceatlas2.barnes_hut"  layout/forceatlas2.rs:49-61   (quadtree, dim 2)
 *   "layout.forceatlas2.3d"          layout/forceatlas2.rs:63-68   (dense, dim 3)
 * registered at registry/layouts.rs:98-103, :263-268 and registry/arms_3d.rs:294-299.
 *
 * Conventions (mirroring crates/graph-wasm's ABI, docs/contract/wasm-abi.md):
 *   - every call returns a gr_status; GR_OK is 0;
 *   - gr_last_error() reports the calling thread's last failure as (code, name, rule),
 *     the fields of StageError::Param / NonFinite (stage.rs:35-52);
 *   - input pointers are borrowed for the call only; outputs go to caller buffers;
 *   - device memory belongs to a session and is freed by gr_fa2_session_release();
 *   - one HIP stream per session; distinct sessions may run concurrently on one device,
 *     from different threads (nothing is issued on the legacy stream); a handle is used by
 *     one thread at a time.
 */
#ifndef GRAPH_RENDER_FA2_H
#define GRAPH_RENDER_FA2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: 1-4 are StageError's variants (stage.rs:35-52); 10-21 the wasm ABI's
 * (docs/contract/wasm-abi.md:355-380); 30+ belong to this device path. */
typedef enum gr_status {
    GR_OK = 0,
    GR_ERR_CAPACITY = 1,           /* StageError::Capacity: n or m past the u32 index space */
    GR_ERR_NON_FINITE = 2,         /* StageError::NonFinite{column} */
    GR_ERR_PARAM = 3,              /* StageError::Param{name, rule} */
    GR_ERR_SNAPSHOT = 4,           /* StageError::Snapshot (reserved) */
    GR_ERR_INVALID_HANDLE = 10,    /* wasm InvalidHandle */
    GR_ERR_UNKNOWN_LAYOUT = 11,    /* wasm UnknownLayoutId */
    GR_ERR_PARAM_OUT_OF_RANGE = 20,/* wasm ParamOutOfRange (ParamsError::OutOfRange) */
    GR_ERR_PARAMS_MALFORMED = 21,  /* wasm ParamsMalformed (ParamsError::Malformed) */
    GR_ERR_DEVICE = 30,            /* HIP runtime failure or no device */
    GR_ERR_UNSUPPORTED = 31        /* a combination this build does not run */
} gr_status;

/* Fa2Params (layout/forceatlas2/state.rs:20-54) field for field. Defaults: max_iter 100,
 * jitter_tolerance 1, scaling_ratio 2, gravity 1, seed 0, dim 2. */
typedef struct gr_fa2_params {
    uint32_t max_iter;
    double jitter_tolerance;
    double scaling_ratio;
    double gravity;
    uint32_t seed;
    uint32_t dim;
} gr_fa2_params;

typedef enum gr_precision {
    GR_PREC_FP64 = 0,       /* the reference's arithmetic, bit for bit: trees, per-node forces,
                               the global sums in node order, steps, trajectories (see
                               GR_EXT_BLOCK_SUMS for the faster, reassociated sums) */
    GR_PREC_FP32 = 1,       /* FP32 forces; FP64 centres of mass and global reductions */
    GR_PREC_FP32_KAHAN = 2  /* as FP32, with compensated per-body accumulation */
} gr_precision;

typedef enum gr_repulsion {
    GR_REP_DEFAULT = 0,  /* what the layout id runs: tree for barnes_hut, dense otherwise */
    GR_REP_TREE = 1,     /* Barnes-Hut over the reference's quadtree (octree at dim 3) */
    GR_REP_DENSE = 2,    /* exact pair sum, the reference's dense arm */
    GR_REP_PM = 3        /* particle mesh: CIC deposit, FFT convolution, CIC gradient read */
} gr_repulsion;

#define GR_EXT_EXPORT_TREE 1u   /* keep cell bounds for gr_fa2_session_tree() */
#define GR_EXT_DIRECT_LAUNCH 2u /* launch each iteration's kernels one by one instead of
                                   replaying the iteration as a captured HIP graph */
#define GR_EXT_BLOCK_SUMS 4u    /* FP64: form the global sums (centroid, swing, traction,
                                   moved) by block reductions instead of the reference's
                                   sequential node-order sums; faster on large graphs, and
                                   the trajectory then agrees with the reference to rounding
                                   (~1e-12 per step, which FA2 amplifies), not bit for bit.
                                   FP32 precisions always use block reductions. */

/* Extensions the reference does not have. A NULL gr_fa2_ext pointer means exactly the
 * reference's semantics: theta 0.8 (barnes_hut.rs:24), FP64, unit edge weights, Mulberry32
 * start positions (state.rs:272-283). */
typedef struct gr_fa2_ext {
    uint32_t struct_size;       /* sizeof(gr_fa2_ext) */
    uint32_t flags;             /* GR_EXT_* */
    double theta;               /* Barnes-Hut opening angle; < 0 means the reference's 0.8 */
    int32_t precision;          /* gr_precision */
    int32_t repulsion;          /* gr_repulsion */
    const double *edge_weights; /* one per raw edge, NULL = unit; a simple edge keeps the
                                   weight of its first raw edge */
    const double *start;        /* n*dim row-major start positions, NULL = Mulberry32 */
    uint32_t pm_grid;           /* PM mesh side (power of two, 64..4096), 0 = 2048 */
    uint32_t reserved;
} gr_fa2_ext;

/* Fills the reference defaults (Fa2Params::default, state.rs:43-54). */
void gr_fa2_params_default(gr_fa2_params *out);
void gr_fa2_ext_default(gr_fa2_ext *out);

/* The registry's wire form (registry/params.rs:144-202): one little-endian f64 per
 * published parameter in schema order max_iter, jitter_tolerance, scaling_ratio, gravity,
 * seed (registry/tunable.rs:229-240); an empty buffer is the defaults; a value outside the
 * schema's range, non-finite, or fractional for an Int is refused, never clamped. */
gr_status gr_fa2_params_decode(const uint8_t *bytes, size_t len, gr_fa2_params *out);

/* Stage::run for one of the registered ids (stage.rs:23-31); the FA2 ids below, and
 * "layout.forceatlas2.forcesim", "layout.force.barnes_hut" and "layout.force.particle_mesh",
 * which publish no parameters (params_len 0; for the force layouts ext->edge_weights carries
 * the edges' strength column and ext->theta, if >= 0, the opening angle). For
 * the three FA2 ids: dedups the raw edge list as
 * simple_graph does (layout/force/mod.rs:101-120), runs to max_iter or the 1e-10 early
 * exit, and writes f32 columns (x, y, and z for the 3D id) narrowed as settle() narrows
 * them (layout/forceatlas2.rs:82-91). On failure the output columns are zeroed. */
gr_status gr_layout_run(const char *layout_id, uint32_t n, const uint32_t *source,
                        const uint32_t *target, size_t m, const uint8_t *params,
                        size_t params_len, const gr_fa2_ext *ext, float *out_x, float *out_y,
                        float *out_z);

/* ---- resident session: the state of Fa2State (state.rs:60-80) kept in HBM ---- */
typedef struct gr_fa2_session gr_fa2_session;

/* Builds the simple graph, masses (degree + 1) and start positions, and uploads them.
 * repulsion GR_REP_DEFAULT is the tree. Returns NULL on failure (see gr_last_error). */
gr_fa2_session *gr_fa2_session_create(uint32_t n, const uint32_t *source,
                                      const uint32_t *target, size_t m,
                                      const gr_fa2_params *params, const gr_fa2_ext *ext);
/* Batched layout: `ngraphs` independent graphs in one session, run by the same launches.
 * Graph b has node_counts[b] nodes and edge_counts[b] raw edges; source/target hold every
 * graph's edges one graph after another, with graph-local endpoints. Edge weights and start
 * positions (gr_fa2_ext) are concatenated the same way. Each graph keeps its own swing,
 * traction, speed and early exit; positions() returns all graphs' nodes in order. No
 * reference counterpart (the reference lays out one graph per call). */
gr_fa2_session *gr_fa2_batch_create(uint32_t ngraphs, const uint32_t *node_counts,
                                    const size_t *edge_counts, const uint32_t *source,
                                    const uint32_t *target, const gr_fa2_params *params,
                                    const gr_fa2_ext *ext);
void gr_fa2_session_release(gr_fa2_session *s);

/* Runs up to `iters` iterations of iterate() (state.rs:130-148), stopping early once an
 * iteration moves less than 1e-10 in total (state.rs:122-128). *done_out (may be NULL)
 * receives the iterations actually applied by this call. */
gr_status gr_fa2_session_step(gr_fa2_session *s, uint32_t iters, uint32_t *done_out);

/* Changes a live session's parameters between steps (the studio's tick-by-tick drive of a
 * layout, layout/force/session.rs; the reference's FA2 takes them once, state.rs:20-54):
 * jitter_tolerance, scaling_ratio, gravity and seed take effect on the next iteration;
 * theta >= 0 replaces the opening angle (< 0 keeps it). Values outside the published
 * ranges (registry/tunable.rs:229-240) are refused, never clamped; dim must not change.
 * The swing/traction/speed state carries on. Drops any recording. */
gr_status gr_fa2_session_set_params(gr_fa2_session *s, const gr_fa2_params *params,
                                    double theta);

/* Current positions, n*dim row-major f64 (widened from FP32 in the FP32 precisions). */
gr_status gr_fa2_session_positions(gr_fa2_session *s, double *out);
gr_status gr_fa2_session_set_positions(gr_fa2_session *s, const double *in);

/* One force evaluation on the current positions with no state change: parts is a mask of
 * 1 attraction, 2 repulsion, 4 gravity, summed in iterate()'s order. out: n*dim f64. */
gr_status gr_fa2_session_forces(gr_fa2_session *s, uint32_t parts, double *out);

/* swing, traction, speed, speed_efficiency, iter_no, iterations applied, early-exit flag
 * (graph 0; gr_fa2_session_graph_scalars for graph b of a batch) */
gr_status gr_fa2_session_scalars(gr_fa2_session *s, double *out7);
gr_status gr_fa2_session_graph_scalars(gr_fa2_session *s, uint32_t graph, double *out7);

/* Simple edges (lo < hi, first raw edge kept) and per-node mass, as the session holds them. */
gr_status gr_fa2_session_graph(gr_fa2_session *s, uint32_t *m_out, uint32_t *lo, uint32_t *hi,
                               double *mass);

/* The tree of the last force evaluation or iteration (GR_REP_TREE only). Any output pointer
 * may be NULL. bounds: 2*dim f64 per cell (lo axes, then hi axes; needs GR_EXT_EXPORT_TREE);
 * bodies: (mass, centre...) f64 per cell; order: n tree positions -> point index. Two calls:
 * the first with cap 0 and no cell output gives *ncells_out; a call whose cell outputs hold
 * fewer than *ncells_out entries (cap) is refused with GR_ERR_CAPACITY. */
gr_status gr_fa2_session_tree(gr_fa2_session *s, uint32_t *ncells_out, uint32_t cap,
                              double *bounds, uint32_t *skip, uint32_t *start, uint32_t *end,
                              double *bodies, uint32_t *order);

/* ---- reverse mode: gradients through unrolled iterations (no reference counterpart; the
 * reference has no derivatives) ---- */

/* Runs up to `iters` iterations as gr_fa2_session_step does, keeping a device checkpoint of
 * the state at the start of every `checkpoint_every` iterations for gr_fa2_session_adjoint.
 * One graph, tree or dense repulsion. *done_out (may be NULL): iterations applied. */
gr_status gr_fa2_session_record(gr_fa2_session *s, uint32_t iters, uint32_t checkpoint_every,
                                uint32_t *done_out);

/* Reverse sweep of the last recording: given d loss / d final positions (n*dim f64,
 * row-major), writes d loss / d edge weight per raw edge (a repeated or self-loop raw edge,
 * which the simple graph drops, gets 0; unweighted sessions are differentiated at weight 1),
 * d loss / d (scaling_ratio, gravity), and d loss / d start positions (n*dim). Any output
 * may be NULL. Each segment is recomputed from its checkpoint and reversed iteration by
 * iteration; the tree's shape and the accept/open decisions are held as the forward made
 * them. The session is left in the state the recording ended in. */
gr_status gr_fa2_session_adjoint(gr_fa2_session *s, const double *pos_bar, double *weight_bar,
                                 double *param_bar2, double *start_bar);

/* The Barnes-Hut walk measured (no reference counterpart), on the tree of the last force
 * evaluation or iteration (as gr_fa2_session_tree; the tree of the current positions after
 * gr_fa2_session_forces): runs a counting variant of the walk once and writes out4 = {wave
 * scan steps, lanes acting summed over the steps, cells accepted, leaf pairs taken}; accepted
 * + pairs are the interactions. Each wave scans 64 consecutive tree positions and visits the
 * union of their reference walks, so in FP64 these are the oracle's gro_walk_counts. */
gr_status gr_fa2_session_walk_stats(gr_fa2_session *s, uint64_t *out4);

/* ---- schedules and the autotuner (no reference counterpart) ----
 * A schedule is how the kernels are launched, never what they compute: every schedule gives
 * the same result bits. It is chosen per session: from the in-process cache of autotuned
 * schedules for the session's size class when there is one, else by gr_fa2_schedule_select
 * (a pure function of the sizes over a table measured on MI355X, as the reference's
 * exec/select.rs picks tiers from measured thresholds). */
typedef struct gr_fa2_schedule {
    uint32_t walk_block;   /* threads per Barnes-Hut walk workgroup: 64, 128 or 256 */
    uint32_t walk_xcd;     /* 1: consecutive walk workgroups (Morton-adjacent bodies) share an
                              XCD and its L2; 0: the hardware's round-robin */
    uint32_t sort_items;   /* radix-sort keys per thread: 4, 8 or 16 (a tile is 256x) */
    uint32_t finish_split; /* force-finish workgroups per LDS round of rows: 1, 2, 4 or 8 */
} gr_fa2_schedule;

/* The measured table's schedule for a session of these sizes (n nodes over ngraphs graphs,
 * m raw edges, dim, gr_precision). Reads no clock and no device. */
gr_status gr_fa2_schedule_select(uint32_t n, size_t m, uint32_t dim, int32_t precision,
                                 uint32_t ngraphs, gr_fa2_schedule *out);
gr_status gr_fa2_session_get_schedule(gr_fa2_session *s, gr_fa2_schedule *out);
/* Installs a schedule (values outside the listed sets are refused). */
gr_status gr_fa2_session_set_schedule(gr_fa2_session *s, const gr_fa2_schedule *sched);
/* Times each schedule knob's candidates on this session's current state (HIP events, the
 * median of three runs of `reps` launches of the stage the knob drives: walk, tree build,
 * force finish), installs the fastest, stores it in the process's cache under the session's
 * size class (later sessions of that class start with it) and writes it to *out. Tree
 * sessions only. Positions and the adaptive-speed state are left as they were. */
gr_status gr_fa2_session_autotune(gr_fa2_session *s, uint32_t reps, gr_fa2_schedule *out);

/* Times `reps` launches of one stage on the session's stream with HIP events: "walk",
 * "tree", "iteration", "pm". Writes the mean milliseconds per launch. */
gr_status gr_fa2_session_time(gr_fa2_session *s, const char *stage, uint32_t reps,
                              double *ms_out);

/* The graph-motor snapshot's binary face (GMSN; the reference's docs/contract/binary-layout.md,
 * written by graph-contract's Snapshot::to_bytes, binary.rs:170-197, from the geometry
 * layout::snapshot builds, layout/mod.rs:205-235): Point nodes, Line edges, the node and edge
 * id tables, endpoints as node positions, f32 x, y (and z, which makes a 0.4 3D snapshot;
 * NULL for 2D, labelled 0.3) and `notes` (code, index) pairs in ascending order. Writes the
 * bytes to `out` when `cap` is large enough; *len_out always receives the size. Refuses what
 * Snapshot::new refuses: duplicate ids, endpoints outside the node table, non-finite
 * coordinates, unknown or unsorted notes. */
gr_status gr_snapshot_encode(uint32_t n, const char *const *node_ids, uint32_t m,
                             const char *const *edge_ids, const uint32_t *source,
                             const uint32_t *target, const float *x, const float *y,
                             const float *z, uint32_t notes, const uint32_t *note_code,
                             const uint32_t *note_index, uint8_t *out, size_t cap,
                             size_t *len_out);

/* ---- the particle mesh under d3's many-body law (layout.force's PM arm) ---- */

/* mesh.rs:31-34: the mesh side for n nodes, ceil(sqrt n) to a power of two in 128..1024. */
uint32_t gr_pm_side_for(uint32_t n);

/* One charge pass of the reference's particle mesh (particle_mesh/charge.rs:29-47) on the
 * device: frame (frame.rs:50-89, reach ceil(dmax/h)), unit-charge CIC deposit, the law
 * G(r) = -r/l (l = |r|^2, raised to sqrt(dmin^2 l) below dmin^2, zero at l >= dmax^2,
 * kernel.rs:80-92) convolved on a side x side mesh (0 = gr_pm_side_for(n)), CIC read, times
 * charge * alpha: the velocity each node gains (out, n*2 f64). xy: n*2 row-major f64.
 * frame_out (may be NULL): step, h, origin x, origin y, cells, reach, 1 if a frame was
 * placed. The mesh transforms in FP32; positions, frame and deposit weights are FP64. */
gr_status gr_pm_charge(uint32_t n, const double *xy, uint32_t side, double charge, double alpha,
                       double distance_min, double distance_max, double *out,
                       double *frame_out);

/* gr_pm_charge without device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, hosts that forbid hipMalloc). Same arguments, checks and result bits as
 * gr_pm_charge, but every device buffer lives in one array inside the library (64 MiB per
 * GPU, held until the process ends), so the sizes are capped: n <= GR_NOALLOC_MAX_NODES and
 * a mesh side <= GR_NOALLOC_MAX_SIDE; larger calls are refused with GR_ERR_CAPACITY. Calls
 * on one GPU share that array: they are serialised inside one process, and calls that
 * overlap on one GPU from several processes race. Runs on the null stream. */
#define GR_NOALLOC_MAX_NODES 262144u
#define GR_NOALLOC_MAX_SIDE 1024u
gr_status gr_pm_charge_noalloc(uint32_t n, const double *xy, uint32_t side, double charge,
                               double alpha, double distance_min, double distance_max,
                               double *out, double *frame_out);

/* ---- d3's many-body force over the quadtree (layout.force's Barnes-Hut arm) ---- */

/* One charge pass of the reference's Barnes-Hut force simulation (barnes_hut/charge.rs:41-53,
 * 120-294: prepare_with + node_delta over every node) on the device: the PR quadtree of the
 * positions (quadtree.rs:46-196), d3 bodies in reverse preorder (a leaf at its chain head,
 * an internal cell at the count-weighted mean of its children's centres, opening threshold
 * w^2/theta^2 per charge/threshold.rs), then every node's stackless walk: open when
 * threshold >= l, a leaf's chain summed pair by pair, a cell taken whole otherwise; gaps past
 * distance_max contribute nothing, an exactly zero axis is jiggled keyed on (node, the
 * cell's insertion node id or the chain head), l is raised to sqrt(dmin^2 l) below dmin^2.
 * out: n*2 f64, the velocity each node gains (NaN for a non-finite node); xy: n*2 row-major
 * f64. seed and tick key the jiggle as the simulation's seed and tick number do. theta >= 0
 * (0 is the exact pair sum); charge and alpha finite; distances >= 0, distance_max may be
 * infinite. All FP64, the reference's operations in its order: the same bits. */
gr_status gr_d3_charge_bh(uint32_t n, const double *xy, double theta, double charge,
                          double alpha, double distance_min, double distance_max, uint32_t seed,
                          uint32_t tick, double *out);

/* gr_d3_charge_bh, then the pass timed on the device-resident positions (no reference
 * counterpart; the timing face gr_fa2_session_time gives the FA2 stages): ms_out[0] the whole
 * pass (tree build, bodies, node ids, walk), ms_out[1] the walk alone, each the median of
 * three timings of `reps` runs with HIP events on the pass's stream, and ms_out[2] the cell
 * count. Every timed run recomputes the deltas in `out`; a run that gives other bits is
 * reported as a device failure. */
gr_status gr_d3_charge_bh_time(uint32_t n, const double *xy, double theta, double charge,
                               double alpha, double distance_min, double distance_max,
                               uint32_t seed, uint32_t tick, uint32_t reps, double *out,
                               double *ms_out);

/* d3's collision pass over the quadtree of the given positions (collide.rs:37-127, the
 * projected positions px = x + vx of a tick): every node's push from the others within
 * 2 * radius, each leaf's points in chain order, a cell skipped when its square lies beyond
 * the reach (outside()), an exactly zero axis jiggled (passes 4 and 5 keyed on (node, other)).
 * out: n*2 f64 deltas in node order (NaN where a node is NaN). FP64, bit for bit. */
gr_status gr_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed,
                        uint32_t tick, double *out);

/* ---- layout.force.barnes_hut: d3's force simulation (the reference's Barnes-Hut layout) ---- */

/* ForceParams (layout/force/params.rs:1-77) field for field; defaults: charge -90, theta 0.9,
 * distance_min 1, distance_max 520, link_distance 60, link_strength_scale 0.15,
 * collide_radius 16, center_strength 1, velocity_decay 0.58, alpha_decay 0.06,
 * alpha_min 0.001, initial_alpha 1. */
typedef struct gr_d3_params {
    double charge_strength;
    double theta;
    double distance_min;
    double distance_max;
    double link_distance;
    double link_strength_scale;
    double collide_radius;
    double center_strength;
    double velocity_decay;
    double alpha_decay;
    double alpha_min;
    double initial_alpha;
} gr_d3_params;

void gr_d3_params_default(gr_d3_params *out);

/* BarnesHut::run (layout/force/barnes_hut.rs:139-200) on the device: the simple graph (a
 * pair keeps its first raw edge's strength; strength NULL = 1 per edge), the golden-spiral
 * start (seed.rs), then `ticks` ticks (the stage runs 112) of the tick of sim.rs: alpha decay,
 * link springs (link.rs), many-body over the quadtree (as gr_d3_charge_bh), centering,
 * collision over the quadtree of the projected positions (collide.rs), integration. FP64,
 * the reference's operations and orders, jiggle seed 0; x, y narrowed to f32
 * (planar_points, NonFinite{node.x} refused). params NULL = the frozen ForceParams.
 * ms_ticks_out (may be NULL): the tick loop's time on the device (HIP events). */
gr_status gr_d3_layout_run(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, float *out_x, float *out_y, double *ms_ticks_out);

/* ParticleMesh::run (layout/force/particle_mesh.rs:96-131): gr_d3_layout_run with the
 * many-body pass on the mesh (d3's law on a side_for(n) mesh: frame, slot-ordered CIC
 * deposit, the reference's radix-2 transforms, kernel spectrum, CIC read) and collision on
 * the hashed grid (particle_mesh/collide.rs), FP64 in the reference's operations and orders.
 * "layout.force.particle_mesh" in gr_layout_run. */
gr_status gr_pm_layout_run(uint32_t n, const uint32_t *source, const uint32_t *target,
                           const double *strength, size_t m, const gr_d3_params *params,
                           uint32_t ticks, float *out_x, float *out_y, double *ms_ticks_out);

/* The mesh's two passes alone over given positions (a fresh mesh: the identity grid order):
 * particle_mesh/charge.rs apply from rest (deltas = field times charge * alpha, node order;
 * zero where the mesh places no frame; frame_out: step, h, origin x, origin y, cells, reach,
 * side) and particle_mesh/collide.rs apply (the hashed grid over xy as the projected
 * positions; order_out, may be NULL: the grid order it sorted). FP64, bit for bit. */
gr_status gr_pm_d3_charge(uint32_t n, const double *xy, double charge, double alpha,
                          double distance_min, double distance_max, double *out,
                          double *frame_out);
gr_status gr_pm_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed,
                           uint32_t tick, double *out, uint32_t *order_out);

/* ---- the resident simulation: ForceSession (layout/force/session.rs:118-270) in HBM ---- */
typedef struct gr_d3_session gr_d3_session;

/* ForceSession::from_frozen (session.rs:132-139): the simple graph with its strengths
 * (strength NULL = 1 per raw edge), the golden-spiral start, alpha = initial_alpha, target 0;
 * parameters checked for finiteness (theta and distances also >= 0). NULL on failure. */
gr_d3_session *gr_d3_session_create(uint32_t n, const uint32_t *source, const uint32_t *target,
                                    const double *strength, size_t m,
                                    const gr_d3_params *params);
/* ForceSession::step (session.rs:165-196): `ticks` ticks; step(112), 112 x step(1) and
 * 7 x step(16) give the same bytes. alpha_out / settled_out (may be NULL): StepReport's
 * alpha and settled (alpha < alpha_min with target 0). */
gr_status gr_d3_session_step(gr_d3_session *s, uint32_t ticks, double *alpha_out,
                             int32_t *settled_out);
/* xs(), ys() (session.rs:223-231) and, with no reference counterpart, the velocity columns:
 * f64, node order; any pointer may be NULL. */
gr_status gr_d3_session_state(gr_d3_session *s, double *x, double *y, double *vx, double *vy);
/* with_particle_mesh (session.rs:156-159): later ticks take the many-body pass on the mesh
 * and collision on the hashed grid (a fresh mesh: the identity grid order). */
gr_status gr_d3_session_with_particle_mesh(gr_d3_session *s);
/* Measurement, no reference counterpart: the last tick's two walks timed again on their own
 * trees (HIP events; medians of three timings of `reps` launches): ms_out[0] the many-body
 * walk, ms_out[1] the collide walk; cells_out[0..1] the two trees' cell counts. On a mesh
 * session: ms_out[0] the node-order centring sums, ms_out[1] the hashed grid's pushes over the
 * last tick's grid order; cells_out[0] the mesh's cells, cells_out[1] the grid's buckets
 * (refused at zero collide reach). The state is left as the last tick left it. Refused before
 * the first step. */
gr_status gr_d3_session_time(gr_d3_session *s, uint32_t reps, double *ms_out,
                             uint32_t *cells_out);
/* alpha() (session.rs:238); NaN for a NULL handle. */
double gr_d3_session_alpha(const gr_d3_session *s);
/* reheat (session.rs:247-252): alpha taken exactly, refused outside 0..=1. */
gr_status gr_d3_session_reheat(gr_d3_session *s, double alpha);
/* set_alpha_target (session.rs:263-267): refused outside 0 <= target < 1. */
gr_status gr_d3_session_set_alpha_target(gr_d3_session *s, double target);
void gr_d3_session_release(gr_d3_session *s);

/* The node id of every cell of the quadtree of xy, in preorder (Quadtree::node_id,
 * quadtree.rs:64-88 via preorder.rs's open): the shape slot the reference's insertion
 * (quadtree/build.rs:25-115) pushed the cell as, derived here from the arena a bulk build
 * makes. *ncells_out receives the cell count; node_id (cap entries) is written when
 * cap >= that count, a size query when node_id is NULL. */
gr_status gr_quadtree_node_ids(uint32_t n, const double *xy, uint32_t cap, uint32_t *ncells_out,
                               uint32_t *node_id);

/* ---- layout.forceatlas2.forcesim (SciGraphs' ForceSim as the reference ports it) ---- */

/* Fa2ForceSimParams (layout/force/forcesim.rs:104-153) and its Default: iterations 50 (at
 * least one step always runs), scale 5 (the rescale's width only), scaling_ratio 2, gravity 1
 * (times 0.1 inside), jitter_tolerance 1, seed 1767573729 (numpy default_rng's seed). */
typedef struct gr_forcesim_params {
    uint32_t iterations;
    double scale;
    double scaling_ratio;
    double gravity;
    double jitter_tolerance;
    uint64_t seed;
} gr_forcesim_params;

void gr_forcesim_params_default(gr_forcesim_params *out);

/* ForceAtlas2ForceSim::run (forcesim.rs:155-163) on the device: the simple graph of the raw
 * (source, target) columns, f32 start positions from numpy's default_rng(seed), the FA2
 * force field with direct repulsion (f64 coefficient rows), attraction and gravity, the
 * adaptive speed and capped move, recentring, then the f64 rescale to `scale`; x, y, z as
 * f32, bit for bit the reference's arithmetic and reduction orders. params NULL = defaults.
 * Refuses a non-finite coordinate as GR_ERR_NON_FINITE (column node.x), outputs zeroed. */
gr_status gr_forcesim_run(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m,
                          const gr_forcesim_params *params, float *out_x, float *out_y,
                          float *out_z);

/* Many graphs in one launch (one workgroup lays out one graph; no reference counterpart, the
 * reference runs one graph per call): graph g has node_counts[g] nodes and edge_counts[g]
 * raw edges (graph-local endpoints, concatenated in graph order); params holds ngraphs
 * entries (NULL = defaults for all). Outputs concatenated in graph order; status[g] is 0, or
 * 2 where the graph's coordinates are not finite (its columns then hold them unrefused). */
gr_status gr_forcesim_batch(uint32_t ngraphs, const uint32_t *node_counts,
                            const size_t *edge_counts, const uint32_t *source,
                            const uint32_t *target, const gr_forcesim_params *params,
                            float *out_x, float *out_y, float *out_z, int32_t *status);

/* gr_forcesim_batch, then the whole batch timed from its device-resident inputs: *ms_out is
 * the median of three timings of `reps` layouts (HIP events; each run copies the start
 * positions in and clears the previous field). */
gr_status gr_forcesim_batch_time(uint32_t ngraphs, const uint32_t *node_counts,
                                 const size_t *edge_counts, const uint32_t *source,
                                 const uint32_t *target, const gr_forcesim_params *params,
                                 float *out_x, float *out_y, float *out_z, int32_t *status,
                                 uint32_t reps, double *ms_out);

/* Last failure of the calling thread: code, parameter/column name, rule text. */
gr_status gr_last_error(const char **name, const char **rule);

/* Library version string. */
const char *gr_version(void);

#ifdef __cplusplus
}
#endif
#endif
