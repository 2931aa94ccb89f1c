/*
 * CPU oracle: TEST INFRASTRUCTURE ONLY (see fa2_oracle.h for what pins each part).
 *
 * A restatement, in C, of graph-core's f64 ForceAtlas2 and its Barnes-Hut arm. Every
 * floating-point expression keeps the operand order of the Rust source it cites, and the
 * file is compiled with -ffp-contract=off, so results are the reference's bit for bit.
 * Reference paths are relative to crates/graph-core/src/.
 */
#include "fa2_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXD 3

/* ------------------------------------------------------------------ rng */

/* synthetic.rs:44-50: integer steps, then / 2^32. */
double gro_mulberry32_next(uint32_t *state) {
    *state += 0x6D2B79F5u;
    uint32_t a = *state;
    uint32_t t = (a ^ (a >> 15)) * (1u | a);
    t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
    return (double)(t ^ (t >> 14)) / 4294967296.0;
}

static uint32_t mulberry_pick(uint32_t *state, size_t len) {
    /* synthetic.rs:60-63: floor(next * len) */
    return (uint32_t)floor(gro_mulberry32_next(state) * (double)len);
}

/* layout/forceatlas2/state.rs:272-283: row-major, dim draws per node. */
void gro_initial_positions(uint32_t n, uint32_t seed, uint32_t dim, double *out) {
    uint32_t st = seed;
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t a = 0; a < dim; a++) out[(size_t)i * dim + a] = gro_mulberry32_next(&st);
}

/* rng.rs:27-34: MurmurHash3 fmix64. */
uint64_t gro_fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

/* rng.rs:39-44 */
static uint64_t jig_fold(uint32_t seed, uint32_t tick, uint32_t pass, uint32_t lo, uint32_t hi) {
    uint64_t a = gro_fmix64((uint64_t)seed ^ ((uint64_t)tick << 32));
    uint64_t b = gro_fmix64(a ^ ((uint64_t)pass << 32) ^ (uint64_t)lo);
    return gro_fmix64(b ^ ((uint64_t)hi << 32) ^ a);
}

/* rng.rs:61-84 */
double gro_jiggle(uint32_t seed, uint32_t tick, uint32_t pass, uint32_t i, uint32_t j) {
    uint32_t lo = i <= j ? i : j, hi = i <= j ? j : i;
    uint64_t w = jig_fold(seed, tick, pass, lo, hi) >> 11;
    if (w == (1ull << 52)) w = (1ull << 52) + 1;
    double u = (double)w * (1.0 / (double)(1ull << 53));
    return (u - 0.5) * 1e-6;
}

/* ------------------------------------------------------------ gate model */

/* synthetic.rs:92-170: the node draws (roll, then one pick for 80% of nodes) precede the
 * edges; an edge's strength is drawn after its endpoints and only when it is kept. The
 * seed remix (stage/topology.rs:37-46) redraws kinds only, so endpoints are seed-free. */
size_t gro_gate_edges(uint32_t count, uint32_t *source, uint32_t *target, size_t cap) {
    return gro_gate_model(count, source, target, NULL, cap);
}

/* synthetic.rs:92-170 with each kept edge's strength (0.4 + 0.4 * a draw, drawn after the
 * endpoints; a self-edge draws none) */
size_t gro_gate_model(uint32_t count, uint32_t *source, uint32_t *target, double *strength,
                      size_t cap) {
    uint32_t st = 0x051042u;
    for (uint32_t i = 0; i < count; i++) {
        double roll = gro_mulberry32_next(&st);
        if (roll < 0.6)
            (void)mulberry_pick(&st, 12);
        else if (roll < 0.8)
            (void)mulberry_pick(&st, 4);
    }
    size_t m = 0;
#define PUSH(A, B)                                                                          \
    do {                                                                                    \
        uint32_t a_ = (A), b_ = (B);                                                        \
        if (a_ != b_ && m < cap) {                                                          \
            source[m] = a_;                                                                 \
            target[m] = b_;                                                                 \
            const double s_ = 0.4 + gro_mulberry32_next(&st) * 0.4;                         \
            if (strength) strength[m] = s_;                                                 \
            m++;                                                                            \
        }                                                                                   \
    } while (0)
    for (uint32_t i = 1; i < count; i++) {
        double x = gro_mulberry32_next(&st), y = gro_mulberry32_next(&st);
        uint32_t t = (uint32_t)floor(x * y * (double)i);
        PUSH(i, t);
        if (gro_mulberry32_next(&st) < 0.5) {
            x = gro_mulberry32_next(&st);
            y = gro_mulberry32_next(&st);
            t = (uint32_t)floor(x * y * (double)i);
            PUSH(i, t);
        }
    }
    uint32_t extras = (uint32_t)floor((double)count * 0.05);
    for (uint32_t e = 0; e < extras; e++) {
        uint32_t a = mulberry_pick(&st, count);
        uint32_t b = mulberry_pick(&st, count);
        PUSH(a, b);
    }
#undef PUSH
    return m;
}

/* ----------------------------------------------------------- tree build */

typedef struct {
    double lo[MAXD], hi[MAXD];
} box_t;

typedef struct {
    int leaf;
    int32_t head;        /* leaf: chain head */
    int32_t child[8];    /* internal: children by slot, -1 empty */
} shape_t;

struct gro_tree {
    uint32_t n, dim;
    const double *pts[MAXD];
    box_t root_bounds;
    int root_set;
    int32_t root;
    shape_t *shape;
    size_t nshape, capshape;
    int32_t *chain_next;
    /* preorder arena */
    box_t *cbox;
    uint32_t *skip, *start, *end;
    uint32_t *node_id; /* the shape slot each cell was flattened from (quadtree.rs:64-88) */
    size_t ncells, capcells;
    uint32_t *order;
    size_t norder;
};

static double pt(const gro_tree *t, uint32_t i, uint32_t a) { return t->pts[a][i]; }

static int finite_pt(const gro_tree *t, uint32_t i) {
    for (uint32_t a = 0; a < t->dim; a++)
        if (!isfinite(pt(t, i, a))) return 0;
    return 1;
}

static int32_t push_shape(gro_tree *t, int leaf, int32_t head) {
    if (t->nshape == t->capshape) {
        t->capshape = t->capshape ? t->capshape * 2 : 64;
        t->shape = (shape_t *)realloc(t->shape, t->capshape * sizeof(shape_t));
    }
    shape_t *s = &t->shape[t->nshape];
    s->leaf = leaf;
    s->head = head;
    for (int c = 0; c < 8; c++) s->child[c] = -1;
    return (int32_t)t->nshape++;
}

/* quadtree.rs:157-195 / octree.rs:96-112,220-254: root square from floor, then doubling. */
static void cover(gro_tree *t, const double *p) {
    box_t *b = &t->root_bounds;
    if (!t->root_set) {
        for (uint32_t a = 0; a < t->dim; a++) {
            double h = floor(p[a]) + 1.0;
            b->lo[a] = h - 1.0;
            b->hi[a] = h;
        }
        t->root_set = 1;
        return;
    }
    double z = (b->hi[0] - b->lo[0] != 0.0) ? b->hi[0] - b->lo[0] : 1.0;
    for (;;) {
        int out = 0;
        for (uint32_t a = 0; a < t->dim; a++)
            if (b->lo[a] > p[a] || p[a] >= b->hi[a]) out = 1;
        if (!out) break;
        z *= 2.0;
        for (uint32_t a = 0; a < t->dim; a++) {
            if (p[a] < b->lo[a])
                b->lo[a] = b->hi[a] - z;
            else
                b->hi[a] = b->lo[a] + z;
        }
        if (!(isfinite(z) && isfinite(b->hi[0] - b->lo[0]))) return;
    }
}

/* bounds.rs:34-50: narrow toward p, slot bit a set when p[a] >= mid[a]. */
static int narrow(box_t *b, uint32_t dim, const double *p) {
    int slot = 0;
    for (uint32_t a = 0; a < dim; a++) {
        double m = (b->lo[a] + b->hi[a]) / 2.0;
        if (p[a] >= m) {
            b->lo[a] = m;
            slot |= 1 << a;
        } else {
            b->hi[a] = m;
        }
    }
    return slot;
}

static int slot_of(const box_t *b, uint32_t dim, const double *p) {
    int slot = 0;
    for (uint32_t a = 0; a < dim; a++) {
        double m = (b->lo[a] + b->hi[a]) / 2.0;
        if (p[a] >= m) slot |= 1 << a;
    }
    return slot;
}

static void set_child(gro_tree *t, int32_t parent, int slot, int32_t child) {
    if (!t->shape[parent].leaf) t->shape[parent].child[slot] = child;
}

static void attach(gro_tree *t, int32_t parent, int slot, int32_t child) {
    if (parent >= 0)
        set_child(t, parent, slot, child);
    else
        t->root = child;
}

static void chain(gro_tree *t, uint32_t point, int32_t node) {
    t->chain_next[point] = t->shape[node].head;
    t->shape[node].head = (int32_t)point;
}

/* quadtree/build.rs:60-96 */
static void insert_leaf(gro_tree *t, uint32_t point, int32_t node, int32_t parent, int pslot,
                        box_t bounds) {
    double p[MAXD], q[MAXD];
    int32_t head = t->shape[node].head;
    int same = 1;
    for (uint32_t a = 0; a < t->dim; a++) {
        p[a] = pt(t, point, a);
        q[a] = pt(t, (uint32_t)head, a);
        if (p[a] != q[a]) same = 0;
    }
    if (same) {
        chain(t, point, node);
        return;
    }
    int32_t at = parent;
    int slot = parent >= 0 ? pslot : 0;
    for (;;) {
        double before = bounds.hi[0] - bounds.lo[0];
        box_t pre = bounds;
        int i = narrow(&bounds, t->dim, p);
        int j = slot_of(&pre, t->dim, q);
        double after = bounds.hi[0] - bounds.lo[0];
        if (!(after < before)) {
            attach(t, at, slot, node);
            chain(t, point, node);
            return;
        }
        int32_t internal = push_shape(t, 0, -1);
        attach(t, at, slot, internal);
        if (i != j) {
            set_child(t, internal, j, node);
            int32_t leaf = push_shape(t, 1, (int32_t)point);
            set_child(t, internal, i, leaf);
            return;
        }
        at = internal;
        slot = i;
    }
}

/* quadtree/build.rs:26-55 */
static void add(gro_tree *t, uint32_t point) {
    if (!finite_pt(t, point)) return;
    if (t->root < 0) {
        t->root = push_shape(t, 1, (int32_t)point);
        return;
    }
    double p[MAXD];
    for (uint32_t a = 0; a < t->dim; a++) p[a] = pt(t, point, a);
    int32_t node = t->root, parent = -1;
    int pslot = 0;
    box_t bounds = t->root_bounds;
    for (;;) {
        if (t->shape[node].leaf) {
            insert_leaf(t, point, node, parent, pslot, bounds);
            return;
        }
        int slot = narrow(&bounds, t->dim, p);
        int32_t child = t->shape[node].child[slot];
        if (child >= 0) {
            parent = node;
            pslot = slot;
            node = child;
        } else {
            int32_t leaf = push_shape(t, 1, (int32_t)point);
            set_child(t, node, slot, leaf);
            return;
        }
    }
}

static box_t quadrant(box_t b, uint32_t dim, int slot) {
    box_t r = b;
    for (uint32_t a = 0; a < dim; a++) {
        double m = (b.lo[a] + b.hi[a]) / 2.0;
        if (slot & (1 << a))
            r.lo[a] = m;
        else
            r.hi[a] = m;
    }
    return r;
}

static void push_cell(gro_tree *t, box_t b, uint32_t skip, uint32_t start, uint32_t end,
                      int32_t node) {
    if (t->ncells == t->capcells) {
        t->capcells = t->capcells ? t->capcells * 2 : 64;
        t->cbox = (box_t *)realloc(t->cbox, t->capcells * sizeof(box_t));
        t->skip = (uint32_t *)realloc(t->skip, t->capcells * sizeof(uint32_t));
        t->start = (uint32_t *)realloc(t->start, t->capcells * sizeof(uint32_t));
        t->end = (uint32_t *)realloc(t->end, t->capcells * sizeof(uint32_t));
        t->node_id = (uint32_t *)realloc(t->node_id, t->capcells * sizeof(uint32_t));
    }
    t->cbox[t->ncells] = b;
    t->skip[t->ncells] = skip;
    t->start[t->ncells] = start;
    t->end[t->ncells] = end;
    t->node_id[t->ncells] = (uint32_t)node;
    t->ncells++;
}

typedef struct {
    int close;
    int32_t node;
    uint32_t k;
    box_t b;
} visit_t;

/* quadtree/preorder.rs:41-99: preorder arena, children in slot order, non-finite points
 * appended ascending. */
static void flatten(gro_tree *t) {
    size_t cap = 64, top = 0;
    visit_t *stack = (visit_t *)malloc(cap * sizeof(visit_t));
    if (t->root >= 0) {
        stack[top].close = 0;
        stack[top].node = t->root;
        stack[top].b = t->root_bounds;
        top++;
    }
    uint32_t nslots = 1u << t->dim;
    while (top > 0) {
        visit_t v = stack[--top];
        if (v.close) {
            t->skip[v.k] = (uint32_t)t->ncells;
            t->end[v.k] = (uint32_t)t->norder;
            continue;
        }
        uint32_t k = (uint32_t)t->ncells;
        uint32_t start = (uint32_t)t->norder;
        shape_t *s = &t->shape[v.node];
        if (s->leaf) {
            int32_t p = s->head;
            while (p >= 0) {
                t->order[t->norder++] = (uint32_t)p;
                p = t->chain_next[p];
            }
            push_cell(t, v.b, k + 1, start, (uint32_t)t->norder, v.node);
        } else {
            push_cell(t, v.b, 0, start, 0, v.node);
            if (top + nslots + 1 >= cap) {
                cap = cap * 2 + nslots + 1;
                stack = (visit_t *)realloc(stack, cap * sizeof(visit_t));
            }
            stack[top].close = 1;
            stack[top].k = k;
            top++;
            for (int slot = (int)nslots - 1; slot >= 0; slot--) {
                int32_t c = t->shape[v.node].child[slot];
                if (c >= 0) {
                    stack[top].close = 0;
                    stack[top].node = c;
                    stack[top].b = quadrant(v.b, t->dim, slot);
                    top++;
                }
            }
        }
    }
    free(stack);
    for (uint32_t i = 0; i < t->n; i++)
        if (!finite_pt(t, i)) t->order[t->norder++] = i;
}

/* quadtree.rs:116-140 */
gro_tree *gro_tree_build(uint32_t n, uint32_t dim, const double *xs, const double *ys,
                         const double *zs) {
    gro_tree *t = (gro_tree *)calloc(1, sizeof(gro_tree));
    t->n = n;
    t->dim = dim;
    t->pts[0] = xs;
    t->pts[1] = ys;
    t->pts[2] = zs;
    t->root = -1;
    t->chain_next = (int32_t *)malloc((n ? n : 1) * sizeof(int32_t));
    for (uint32_t i = 0; i < n; i++) t->chain_next[i] = -1;
    t->order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
    /* bounds_of: the finite extent (quadtree.rs:46-53) */
    int any = 0;
    double lo[MAXD], hi[MAXD];
    for (uint32_t i = 0; i < n; i++) {
        if (!finite_pt(t, i)) continue;
        for (uint32_t a = 0; a < dim; a++) {
            double v = pt(t, i, a);
            if (!any) {
                lo[a] = v;
                hi[a] = v;
            } else {
                if (v < lo[a]) lo[a] = v;
                if (v > hi[a]) hi[a] = v;
            }
        }
        any = 1;
    }
    if (any) {
        cover(t, lo);
        cover(t, hi);
        for (uint32_t i = 0; i < n; i++) add(t, i);
    }
    flatten(t);
    return t;
}

void gro_tree_free(gro_tree *t) {
    if (!t) return;
    free(t->shape);
    free(t->chain_next);
    free(t->cbox);
    free(t->skip);
    free(t->start);
    free(t->end);
    free(t->node_id);
    free(t->order);
    free(t);
}

uint32_t gro_tree_ncells(const gro_tree *t) { return (uint32_t)t->ncells; }

void gro_tree_cells(const gro_tree *t, double *bounds, uint32_t *skip, uint32_t *start,
                    uint32_t *end) {
    uint32_t d = t->dim;
    for (size_t k = 0; k < t->ncells; k++) {
        for (uint32_t a = 0; a < d; a++) {
            bounds[k * 2 * d + a] = t->cbox[k].lo[a];
            bounds[k * 2 * d + d + a] = t->cbox[k].hi[a];
        }
        skip[k] = t->skip[k];
        start[k] = t->start[k];
        end[k] = t->end[k];
    }
}

void gro_tree_order(const gro_tree *t, uint32_t *order) {
    memcpy(order, t->order, t->n * sizeof(uint32_t));
}

void gro_tree_node_ids(const gro_tree *t, uint32_t *out) {
    if (t->ncells) memcpy(out, t->node_id, t->ncells * sizeof(uint32_t));
}

/* ------------------------------------------- d3's many-body force over the quadtree */

/* barnes_hut/charge/threshold.rs:20-41: w^2 / theta^2, descend everything at theta 0 */
static double opening_threshold(double w, double theta, double theta2) {
    if (theta == 0.0) return INFINITY;
    double naive = w * w / theta2;
    if (!isnan(naive)) return naive;
    if (!isfinite(theta)) return INFINITY;
    double ratio = w / theta;
    return isfinite(ratio) ? ratio * ratio : INFINITY;
}

typedef struct {
    double comx, comy, open;
    uint32_t count, skip, start;
} d3body_t;

typedef struct {
    const gro_tree *t;
    const d3body_t *bodies;
    const double *x, *y;
    double dmin2, dmax2, charge, alpha;
    uint32_t seed, tick;
} d3ctx_t;

/* barnes_hut/charge.rs:262-294 (settle): 0 past distanceMax, else the gap with an exactly
 * zero axis jiggled (keyed on (i, key)) and l raised to distanceMin */
static int d3_settle(const d3ctx_t *c, uint32_t i, uint32_t key, double *dx, double *dy,
                     double *l) {
    if (*l >= c->dmax2) return 0;
    if (*dx == 0.0) {
        *dx = gro_jiggle(c->seed, c->tick, 2, i, key);
        *l += *dx * *dx;
    }
    if (*dy == 0.0) {
        *dy = gro_jiggle(c->seed, c->tick, 3, i, key);
        *l += *dy * *dy;
    }
    if (*l < c->dmin2) *l = sqrt(c->dmin2 * *l);
    if (*l == 0.0) *l = c->dmin2 > DBL_MIN ? c->dmin2 : DBL_MIN;
    return 1;
}

/* barnes_hut/charge.rs:196-240 (node_delta, approx, direct) */
static void d3_node_delta(const d3ctx_t *c, uint32_t i, double *ox, double *oy) {
    const double xi = c->x[i], yi = c->y[i];
    double outx = 0.0, outy = 0.0;
    uint32_t k = 0;
    const uint32_t nc = (uint32_t)c->t->ncells;
    while (k < nc) {
        const d3body_t *b = &c->bodies[k];
        double dx = b->comx - xi, dy = b->comy - yi;
        double l = dx * dx + dy * dy;
        double ddx = 0.0, ddy = 0.0;
        uint32_t next;
        if (b->open >= l) {
            if (b->skip != k + 1) {
                k++;
                continue;
            }
            const uint32_t *chain = c->t->order + b->start;
            if (d3_settle(c, i, chain[0], &dx, &dy, &l)) {
                for (uint32_t q = 0; q < b->count; q++) {
                    if (chain[q] == i) continue;
                    double f = c->charge * c->alpha / l;
                    ddx += dx * f;
                    ddy += dy * f;
                }
            }
            next = k + 1;
        } else {
            if (d3_settle(c, i, c->t->node_id[k], &dx, &dy, &l)) {
                double m = (double)b->count;
                double f = m * c->charge * c->alpha / l;
                ddx = dx * f;
                ddy = dy * f;
            }
            next = b->skip;
        }
        outx += ddx;
        outy += ddy;
        k = next;
    }
    *ox = outx;
    *oy = outy;
}

/* The velocity each node gains from d3's many-body force over the quadtree of its
 * positions (barnes_hut/charge.rs:41-53 + 120-140): bodies in reverse preorder (a leaf at
 * its chain head's position, an internal cell at the count-weighted mean of its children's
 * centres in slot order), then every node's walk, non-finite nodes included. xy: n*2
 * row-major; out: n*2. */
static void d3_charge_core(uint32_t n, const double *xs, const double *ys, double theta,
                           double charge, double alpha, double dmin, double dmax, uint32_t seed,
                           uint32_t tick, double *out);

void gro_d3_charge_bh(uint32_t n, const double *xy, double theta, double charge, double alpha,
                      double dmin, double dmax, uint32_t seed, uint32_t tick, double *out) {
    double *xs = (double *)malloc((n ? n : 1) * sizeof(double));
    double *ys = (double *)malloc((n ? n : 1) * sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        xs[i] = xy[2 * (size_t)i];
        ys[i] = xy[2 * (size_t)i + 1];
    }
    d3_charge_core(n, xs, ys, theta, charge, alpha, dmin, dmax, seed, tick, out);
    free(xs);
    free(ys);
}

/* out: n*2, the delta of every node */
static void d3_charge_core(uint32_t n, const double *xs, const double *ys, double theta,
                           double charge, double alpha, double dmin, double dmax, uint32_t seed,
                           uint32_t tick, double *out) {
    gro_tree *t = gro_tree_build(n, 2, xs, ys, NULL);
    const size_t nc = t->ncells;
    d3body_t *bodies = (d3body_t *)calloc(nc ? nc : 1, sizeof(d3body_t));
    const double theta2 = theta * theta;
    for (size_t kk = nc; kk-- > 0;) {
        double cx, cy;
        if (t->skip[kk] == kk + 1) {
            uint32_t head = t->order[t->start[kk]];
            cx = xs[head];
            cy = ys[head];
        } else {  /* charge/threshold.rs:54-68 */
            double m = 0.0, sx = 0.0, sy = 0.0;
            size_t ch = kk + 1;
            while (ch < t->skip[kk]) {
                const d3body_t *cb = &bodies[ch];
                double cm = (double)cb->count;
                m += cm;
                sx += cm * cb->comx;
                sy += cm * cb->comy;
                ch = cb->skip;
            }
            cx = sx / m;
            cy = sy / m;
        }
        double w = t->cbox[kk].hi[0] - t->cbox[kk].lo[0];
        d3body_t *b = &bodies[kk];
        b->comx = cx;
        b->comy = cy;
        b->open = opening_threshold(w, theta, theta2);
        b->count = t->end[kk] - t->start[kk];
        b->skip = t->skip[kk];
        b->start = t->start[kk];
    }
    d3ctx_t c = {t, bodies, xs, ys, dmin * dmin, dmax * dmax, charge, alpha, seed, tick};
    /* every node walks, a non-finite one too (it is in no leaf, so its walk reads every
     * cell as it would any far query: NaN gaps approximate, infinite ones fall past
     * distanceMax) */
    for (uint32_t i = 0; i < n; i++)
        d3_node_delta(&c, i, &out[2 * (size_t)i], &out[2 * (size_t)i + 1]);
    free(bodies);
    gro_tree_free(t);
}

/* ------------------------------------------------- d3's Barnes-Hut force simulation */

/* collide.rs: the node's walk over the quadtree of the projected positions; a leaf's
 * points are resolved pair by pair, an internal cell is skipped when its square lies
 * beyond the reach on some axis */
static void d3_collide_delta(const gro_tree *t, uint32_t i, const double *px, const double *py,
                             double d2, uint32_t seed, uint32_t tick, double *ox, double *oy) {
    const double xi = px[i], yi = py[i], reach = sqrt(d2);
    double outx = 0.0, outy = 0.0;
    uint32_t k = 0;
    while (k < t->ncells) {
        if (t->skip[k] == k + 1) {
            for (uint32_t q = t->start[k]; q < t->end[k]; q++) {
                const uint32_t p = t->order[q];
                if (p == i) continue;
                double dx = xi - px[p], dy = yi - py[p];
                double l = dx * dx + dy * dy;
                if (l >= d2) continue;
                if (dx == 0.0) {
                    dx = gro_jiggle(seed, tick, 4, i, p);
                    l += dx * dx;
                }
                if (dy == 0.0) {
                    dy = gro_jiggle(seed, tick, 5, i, p);
                    l += dy * dy;
                }
                const double dist = sqrt(l);
                const double push = (reach - dist) / dist * 0.5;
                outx += dx * push;
                outy += dy * push;
            }
            k = k + 1;
        } else {
            const box_t *b = &t->cbox[k];
            const int outside = b->lo[0] > xi + reach || b->hi[0] < xi - reach ||
                                b->lo[1] > yi + reach || b->hi[1] < yi - reach;
            k = outside ? t->skip[k] : k + 1;
        }
    }
    *ox = outx;
    *oy = outy;
}

void gro_d3_collide(uint32_t n, const double *xy, double radius, uint32_t seed, uint32_t tick,
                    double *out) {
    const size_t nn = n ? n : 1;
    double *px = (double *)calloc(nn, sizeof(double)), *py = (double *)calloc(nn, sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        px[i] = xy[2 * (size_t)i];
        py[i] = xy[2 * (size_t)i + 1];
    }
    gro_tree *t = gro_tree_build(n, 2, px, py, NULL);
    const double d2 = (2.0 * radius) * (2.0 * radius);
    for (uint32_t i = 0; i < n; i++)
        d3_collide_delta(t, i, px, py, d2, seed, tick, &out[2 * (size_t)i], &out[2 * (size_t)i + 1]);
    gro_tree_free(t);
    free(px);
    free(py);
}

/* step.rs:200-214 (merge): every delta added to its own node's velocity, as d + 0 */
static void d3_merge(uint32_t n, double *vx, double *vy, const double *d) {
    for (uint32_t i = 0; i < n; i++) {
        vx[i] += d[2 * (size_t)i] + 0.0;
        vy[i] += d[2 * (size_t)i + 1] + 0.0;
    }
}

/* The frozen layout.force.barnes_hut (barnes_hut.rs run_under + session.rs step_under +
 * sim.rs tick): the simple graph with strengths, link geometry (distance / max(0.4, s),
 * min(0.7, scale * s), degree bias), golden-spiral start, then `ticks` ticks of alpha decay,
 * link, charge, center, collide, integrate. p: charge, theta, distance_min, distance_max,
 * link_distance, link_strength_scale, collide_radius, center_strength, velocity_decay,
 * alpha_decay, initial_alpha. x_out, y_out: n each. */
void gro_d3sim_run(uint32_t n, const uint32_t *src, const uint32_t *dst, const double *strength,
                   size_t m_raw, uint32_t ticks, const double *p, uint32_t seed,
                   uint32_t mesh_side, double *x_out, double *y_out, double *vx_out,
                   double *vy_out) {
    const double charge = p[0], theta = p[1], dmin = p[2], dmax = p[3], ldist = p[4];
    const double lscale = p[5], radius = p[6], cstrength = p[7], decay = p[8];
    const double adecay = p[9];
    double alpha = p[10];
    /* simple graph (layout/force/mod.rs:101-120): first occurrence of a pair keeps its
     * strength */
    uint32_t *lo = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    uint32_t *hi = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    double *str = (double *)malloc((m_raw ? m_raw : 1) * sizeof(double));
    uint32_t m = 0;
    /* seen pairs in an open-addressed table (linear probing, at most half full) */
    size_t cap = 16;
    while (cap < 2 * m_raw + 2) cap <<= 1;
    uint64_t *seen = (uint64_t *)malloc(cap * sizeof(uint64_t));
    for (size_t q = 0; q < cap; q++) seen[q] = ~0ull;
    for (size_t e = 0; e < m_raw; e++) {
        uint32_t a = src[e], b = dst[e];
        if (a == b) continue;
        const uint32_t l = a < b ? a : b, h = a < b ? b : a;
        const uint64_t key = ((uint64_t)l << 32) | h;
        size_t q = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 20) & (cap - 1);
        int dup = 0;
        while (seen[q] != ~0ull && !dup) {
            dup = seen[q] == key;
            q = (q + 1) & (cap - 1);
        }
        if (dup) continue;
        seen[q] = key;
        lo[m] = l;
        hi[m] = h;
        str[m] = strength ? strength[e] : 1.0;
        m++;
    }
    free(seen);
    const size_t nn = n ? n : 1;
    uint32_t *deg = (uint32_t *)calloc(nn, sizeof(uint32_t));
    for (uint32_t e = 0; e < m; e++) {
        deg[lo[e]]++;
        deg[hi[e]]++;
    }
    /* rows: each node's simple edges in ascending edge index */
    uint32_t *roff = (uint32_t *)calloc(nn + 1, sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) roff[v + 1] = roff[v] + deg[v];
    uint32_t *rows = (uint32_t *)malloc((size_t)(m ? 2 * m : 1) * sizeof(uint32_t));
    uint32_t *fill = (uint32_t *)malloc(nn * sizeof(uint32_t));
    for (uint32_t v = 0; v < n; v++) fill[v] = roff[v];
    for (uint32_t e = 0; e < m; e++) {
        rows[fill[lo[e]]++] = e;
        rows[fill[hi[e]]++] = e;
    }
    /* link.rs geometry */
    double *ld = (double *)malloc((m ? m : 1) * sizeof(double));
    double *ls = (double *)malloc((m ? m : 1) * sizeof(double));
    double *lb = (double *)malloc((m ? m : 1) * sizeof(double));
    for (uint32_t e = 0; e < m; e++) {
        const double sv = str[e], dlo = (double)deg[lo[e]], dhi = (double)deg[hi[e]];
        ld[e] = ldist / fmax(0.4, sv);
        ls[e] = fmin(0.7, lscale * sv);
        lb[e] = dlo / (dlo + dhi);
    }
    /* seed.rs: node i at 12 sqrt(i + 1) along the golden-angle spiral, libm's cos and sin */
    double *x = (double *)malloc(nn * sizeof(double)), *y = (double *)malloc(nn * sizeof(double));
    double *vx = (double *)calloc(nn, sizeof(double)), *vy = (double *)calloc(nn, sizeof(double));
    double *px = (double *)malloc(nn * sizeof(double)), *py = (double *)malloc(nn * sizeof(double));
    double *d = (double *)malloc(nn * 2 * sizeof(double));
    double *lf = (double *)malloc((size_t)(m ? m : 1) * 2 * sizeof(double));
    for (uint32_t i = 0; i < n; i++) {
        const double r = 12.0 * sqrt((double)i + 1.0);
        const double a = (double)i * 2.399963229728653;
        x[i] = gro_libm_cos(a) * r;
        y[i] = gro_libm_sin(a) * r;
    }
    const double reach2 = (2.0 * radius) * (2.0 * radius);
    gro_pmd3 *mesh = mesh_side ? gro_pmd3_new(n, (int)mesh_side) : NULL;
    for (uint32_t tick = 0; tick < ticks; tick++) {
        alpha += (0.0 - alpha) * adecay;
        /* link (link.rs force + step.rs LinkPass): per edge, then each node's share */
        for (uint32_t e = 0; e < m; e++) {
            const uint32_t a = lo[e], b = hi[e];
            double dx = (x[b] + vx[b]) - (x[a] + vx[a]);
            double dy = (y[b] + vy[b]) - (y[a] + vy[a]);
            if (dx == 0.0) dx = gro_jiggle(seed, tick, 0, a, b);
            if (dy == 0.0) dy = gro_jiggle(seed, tick, 1, a, b);
            const double l = sqrt(dx * dx + dy * dy);
            if (l == 0.0) {
                lf[2 * e] = 0.0;
                lf[2 * e + 1] = 0.0;
                continue;
            }
            const double factor = (l - ld[e]) / l * alpha * ls[e];
            lf[2 * e] = dx * factor;
            lf[2 * e + 1] = dy * factor;
        }
        for (uint32_t v = 0; v < n; v++) {
            double sx = 0.0, sy = 0.0;
            for (uint32_t q = roff[v]; q < roff[v + 1]; q++) {
                const uint32_t e = rows[q];
                const double b = lb[e];
                if (hi[e] == v) {
                    sx += -lf[2 * e] * b;
                    sy += -lf[2 * e + 1] * b;
                } else {
                    sx += lf[2 * e] * (1.0 - b);
                    sy += lf[2 * e + 1] * (1.0 - b);
                }
            }
            d[2 * v] = sx;
            d[2 * v + 1] = sy;
        }
        d3_merge(n, vx, vy, d);
        /* charge: the quadtree's (barnes_hut/charge.rs), or the mesh's (particle_mesh/
         * charge.rs: no merge when the mesh places no frame) */
        if (!mesh) {
            d3_charge_core(n, x, y, theta, charge, alpha, dmin, dmax, seed, tick, d);
            d3_merge(n, vx, vy, d);
        } else if (gro_pmd3_charge(mesh, x, y, dmin, dmax, charge * alpha, d, NULL)) {
            d3_merge(n, vx, vy, d);
        }
        /* center (sim.rs center_shift): the mean in node order, times the strength */
        if (n > 0) {
            double sx = 0.0, sy = 0.0;
            for (uint32_t i = 0; i < n; i++) {
                sx = sx + x[i];
                sy = sy + y[i];
            }
            const double shx = (sx / (double)n) * cstrength, shy = (sy / (double)n) * cstrength;
            for (uint32_t i = 0; i < n; i++) {
                x[i] -= shx;
                y[i] -= shy;
            }
        }
        /* collide over the projected positions */
        for (uint32_t i = 0; i < n; i++) {
            px[i] = x[i] + vx[i];
            py[i] = y[i] + vy[i];
        }
        if (!mesh) {
            gro_tree *ct = gro_tree_build(n, 2, px, py, NULL);
            for (uint32_t i = 0; i < n; i++)
                d3_collide_delta(ct, i, px, py, reach2, seed, tick, &d[2 * i], &d[2 * i + 1]);
            gro_tree_free(ct);
            d3_merge(n, vx, vy, d);
        } else if (gro_pmd3_collide(mesh, px, py, radius, seed, tick, d)) {
            /* particle_mesh/collide.rs: the hashed grid (its order kept for the next
             * deposit); no pass at zero reach */
            d3_merge(n, vx, vy, d);
        }
        /* integrate (sim.rs): velocity decays, position moves */
        for (uint32_t i = 0; i < n; i++) {
            vx[i] *= decay;
            x[i] += vx[i];
            vy[i] *= decay;
            y[i] += vy[i];
        }
    }
    memcpy(x_out, x, (size_t)n * sizeof(double));
    memcpy(y_out, y, (size_t)n * sizeof(double));
    if (vx_out) memcpy(vx_out, vx, (size_t)n * sizeof(double));
    if (vy_out) memcpy(vy_out, vy, (size_t)n * sizeof(double));
    gro_pmd3_free(mesh);
    free(lo);
    free(hi);
    free(str);
    free(deg);
    free(roff);
    free(rows);
    free(fill);
    free(ld);
    free(ls);
    free(lb);
    free(x);
    free(y);
    free(vx);
    free(vy);
    free(px);
    free(py);
    free(d);
    free(lf);
}

/* barnes_hut.rs:62-110: reverse preorder; a leaf sums its chain, an internal cell its
 * children's unnormalised sums in slot order; then every sum becomes a centre. */
static void aggregate(const gro_tree *t, const double *mass, const double *const *pos,
                      double *bodies /* (1+dim) per cell */) {
    uint32_t d = t->dim, w = 1 + d;
    for (size_t kk = t->ncells; kk-- > 0;) {
        double s[1 + MAXD] = {0.0, 0.0, 0.0, 0.0};
        if (t->skip[kk] == kk + 1) {
            for (uint32_t q = t->start[kk]; q < t->end[kk]; q++) {
                uint32_t p = t->order[q];
                double m = mass[p];
                s[0] += m;
                for (uint32_t a = 0; a < d; a++) s[1 + a] += m * pos[a][p];
            }
        } else {
            size_t c = kk + 1;
            while (c < t->skip[kk]) {
                for (uint32_t a = 0; a < w; a++) s[a] += bodies[c * w + a];
                c = t->skip[c];
            }
        }
        for (uint32_t a = 0; a < w; a++) bodies[kk * w + a] = s[a];
    }
    for (size_t k = 0; k < t->ncells; k++)
        for (uint32_t a = 0; a < d; a++) bodies[k * w + 1 + a] = bodies[k * w + 1 + a] / bodies[k * w];
}

void gro_tree_bodies(const gro_tree *t, const double *mass, const double *xs, const double *ys,
                     const double *zs, double *out) {
    const double *pos[MAXD] = {xs, ys, zs};
    aggregate(t, mass, pos, out);
}

/* ------------------------------------------------------------- FA2 state */

struct gro_fa2 {
    uint32_t n, dim, m;
    uint32_t *lo, *hi;
    double *w; /* per simple edge weight, NULL = unit (the reference) */
    double jitter_tolerance, scaling_ratio, gravity;
    uint32_t seed;
    uint32_t iter_no;
    double swing, traction, speed, speed_efficiency;
    double *p, *mass, *u; /* p,u: n*MAXD rows */
    int repulsion;
    double theta2;
};

/* layout/force/mod.rs:101-120: drop self-loops and repeated unordered pairs, first kept. */
static void simple_graph(gro_fa2 *s, const uint32_t *src, const uint32_t *dst, size_t m_raw,
                         const double *weights) {
    size_t cap = 16;
    while (cap < 2 * m_raw + 2) cap <<= 1;
    uint64_t *table = (uint64_t *)malloc(cap * sizeof(uint64_t));
    for (size_t i = 0; i < cap; i++) table[i] = ~0ull;
    s->lo = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    s->hi = (uint32_t *)malloc((m_raw ? m_raw : 1) * sizeof(uint32_t));
    s->w = weights ? (double *)malloc((m_raw ? m_raw : 1) * sizeof(double)) : NULL;
    uint32_t m = 0;
    for (size_t e = 0; e < m_raw; e++) {
        uint32_t a = src[e], b = dst[e];
        if (a == b) continue;
        uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
        uint64_t key = ((uint64_t)lo << 32) | hi;
        size_t h = (size_t)(gro_fmix64(key) & (cap - 1));
        int dup = 0;
        while (table[h] != ~0ull) {
            if (table[h] == key) {
                dup = 1;
                break;
            }
            h = (h + 1) & (cap - 1);
        }
        if (dup) continue;
        table[h] = key;
        s->lo[m] = lo;
        s->hi[m] = hi;
        if (s->w) s->w[m] = weights[e];
        m++;
    }
    free(table);
    s->m = m;
}

gro_fa2 *gro_fa2_new(uint32_t n, const uint32_t *source, const uint32_t *target, size_t m_raw,
                     const double *weights_raw, uint32_t dim, double jitter_tolerance,
                     double scaling_ratio, double gravity, uint32_t seed, const double *start,
                     int repulsion, double theta2) {
    gro_fa2 *s = (gro_fa2 *)calloc(1, sizeof(gro_fa2));
    s->n = n;
    s->dim = dim;
    s->jitter_tolerance = jitter_tolerance;
    s->scaling_ratio = scaling_ratio;
    s->gravity = gravity;
    s->seed = seed;
    s->swing = s->traction = s->speed = s->speed_efficiency = 1.0;
    s->repulsion = repulsion;
    s->theta2 = theta2;
    simple_graph(s, source, target, m_raw, weights_raw);
    size_t nn = n ? n : 1;
    s->p = (double *)calloc(nn * MAXD, sizeof(double));
    s->u = (double *)calloc(nn * MAXD, sizeof(double));
    s->mass = (double *)calloc(nn, sizeof(double));
    uint32_t *deg = (uint32_t *)calloc(nn, sizeof(uint32_t));
    for (uint32_t e = 0; e < s->m; e++) {
        deg[s->lo[e]]++;
        deg[s->hi[e]]++;
    }
    for (uint32_t v = 0; v < n; v++) s->mass[v] = (double)deg[v] + 1.0; /* state.rs:86 */
    free(deg);
    if (start) {
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t a = 0; a < dim; a++) s->p[(size_t)i * MAXD + a] = start[(size_t)i * dim + a];
    } else {
        uint32_t st = seed;
        for (uint32_t i = 0; i < n; i++)
            for (uint32_t a = 0; a < dim; a++) s->p[(size_t)i * MAXD + a] = gro_mulberry32_next(&st);
    }
    return s;
}

void gro_fa2_free(gro_fa2 *s) {
    if (!s) return;
    free(s->lo);
    free(s->hi);
    free(s->w);
    free(s->p);
    free(s->u);
    free(s->mass);
    free(s);
}

/* A live session's parameter change between iterations: the next iterate() uses them. */
void gro_fa2_set_params(gro_fa2 *s, double jitter_tolerance, double scaling_ratio,
                        double gravity, uint32_t seed) {
    s->jitter_tolerance = jitter_tolerance;
    s->scaling_ratio = scaling_ratio;
    s->gravity = gravity;
    s->seed = seed;
}

void gro_fa2_set_theta2(gro_fa2 *s, double theta2, int repulsion) {
    s->theta2 = theta2;
    s->repulsion = repulsion;
}

#define P(s, i, a) ((s)->p[(size_t)(i) * MAXD + (a)])
#define U(s, i, a) ((s)->u[(size_t)(i) * MAXD + (a)])

static double norm2(const double *v, uint32_t dim) {
    double sum = 0.0;
    for (uint32_t a = 0; a < dim; a++) sum += v[a] * v[a];
    return sum;
}

/* state.rs:152-162 (weights: the extension multiplies the pull per simple edge) */
static void attraction(gro_fa2 *s) {
    for (uint32_t e = 0; e < s->m; e++) {
        uint32_t lo = s->lo[e], hi = s->hi[e];
        for (uint32_t a = 0; a < s->dim; a++) {
            double d = P(s, hi, a) - P(s, lo, a);
            if (s->w) d = d * s->w[e];
            U(s, lo, a) += d;
            U(s, hi, a) -= d;
        }
    }
}

/* state.rs:185-203 */
static void repel_delta(const gro_fa2 *s, uint32_t i, uint32_t j, double *d) {
    int all_zero = 1;
    for (uint32_t a = 0; a < s->dim; a++) {
        d[a] = P(s, i, a) - P(s, j, a);
        if (d[a] != 0.0) all_zero = 0;
    }
    if (all_zero)
        for (uint32_t a = 0; a < s->dim; a++) d[a] = gro_jiggle(s->seed, s->iter_no, a, i, j);
}

/* state.rs:169-181 */
static void repulsion_dense(gro_fa2 *s) {
    double k = s->scaling_ratio;
    for (uint32_t i = 0; i < s->n; i++) {
        for (uint32_t j = i + 1; j < s->n; j++) {
            double d[MAXD];
            repel_delta(s, i, j, d);
            double d2 = norm2(d, s->dim);
            double f = s->mass[i] * s->mass[j] / d2 * k;
            for (uint32_t a = 0; a < s->dim; a++) {
                U(s, i, a) += d[a] * f;
                U(s, j, a) -= d[a] * f;
            }
        }
    }
}

/* barnes_hut.rs:163-183 */
static void pairs(const gro_fa2 *s, uint32_t i, const uint32_t *leaf, uint32_t len, double *f) {
    double kr = s->scaling_ratio;
    for (uint32_t q = 0; q < len; q++) {
        uint32_t j = leaf[q];
        if (j == i) continue;
        double d[MAXD];
        if (i < j) {
            repel_delta(s, i, j, d);
        } else {
            repel_delta(s, j, i, d);
            for (uint32_t a = 0; a < s->dim; a++) d[a] = -d[a];
        }
        double d2 = d[0] * d[0] + d[1] * d[1];
        if (s->dim == 3) d2 = d2 + d[2] * d[2];
        double ff = s->mass[i] * s->mass[j] / d2 * kr;
        for (uint32_t a = 0; a < s->dim; a++) f[a] = f[a] + d[a] * ff;
    }
}

/* barnes_hut.rs:112-159: one stackless preorder walk per tree position. */
static void repulsion_tree(gro_fa2 *s) {
    uint32_t n = s->n, d = s->dim, w = 1 + d;
    double *cols[MAXD];
    for (uint32_t a = 0; a < d; a++) {
        cols[a] = (double *)malloc((n ? n : 1) * sizeof(double));
        for (uint32_t i = 0; i < n; i++) cols[a][i] = P(s, i, a);
    }
    gro_tree *t = gro_tree_build(n, d, cols[0], cols[1], d == 3 ? cols[2] : NULL);
    double *bodies = (double *)malloc((t->ncells ? t->ncells : 1) * w * sizeof(double));
    aggregate(t, s->mass, (const double *const *)cols, bodies);
    double kr = s->scaling_ratio;
    for (uint32_t p = 0; p < n; p++) {
        uint32_t i = t->order[p];
        double xi[MAXD], f[MAXD] = {0.0, 0.0, 0.0};
        for (uint32_t a = 0; a < d; a++) xi[a] = P(s, i, a);
        double mi = s->mass[i];
        size_t k = 0;
        while (k < t->ncells) {
            int holds = t->start[k] <= p && p < t->end[k];
            const double *b = &bodies[k * w];
            double dx[MAXD] = {0.0, 0.0, 0.0};
            for (uint32_t a = 0; a < d; a++) dx[a] = xi[a] - b[1 + a];
            double d2 = dx[0] * dx[0] + dx[1] * dx[1];
            if (d == 3) d2 = d2 + dx[2] * dx[2];
            double wd = t->cbox[k].hi[0] - t->cbox[k].lo[0];
            if (!holds && d2 > 0.0 && wd * wd < s->theta2 * d2) {
                double ff = mi * b[0] / d2 * kr;
                for (uint32_t a = 0; a < d; a++) f[a] = f[a] + dx[a] * ff;
                k = t->skip[k];
            } else if (t->skip[k] == k + 1) {
                double lf[MAXD] = {0.0, 0.0, 0.0};
                pairs(s, i, &t->order[t->start[k]], t->end[k] - t->start[k], lf);
                for (uint32_t a = 0; a < d; a++) f[a] = f[a] + lf[a];
                k += 1;
            } else {
                k += 1;
            }
        }
        for (uint32_t a = 0; a < d; a++) U(s, i, a) += f[a];
    }
    free(bodies);
    gro_tree_free(t);
    for (uint32_t a = 0; a < d; a++) free(cols[a]);
}

/* The walk of barnes_hut.rs:112-159 replayed per tree position, counted as a wave-uniform scan
 * over groups of `group` consecutive positions visits it: out[0] the steps, the size of the
 * union of the group's visit sets summed over groups; out[1] the visits (each position's
 * cells); out[2] the cells accepted; out[3] the leaf pairs taken (a leaf's other points).
 * Points given in node order; the tree is the reference's over them. */
void gro_walk_counts(uint32_t n, uint32_t dim, const double *xs, const double *ys,
                     const double *zs, const double *mass, double theta2, uint32_t group,
                     uint64_t *out) {
    const uint32_t d = dim, w = 1 + d;
    const double *cols[MAXD] = {xs, ys, zs};
    for (int q = 0; q < 4; q++) out[q] = 0;
    if (n == 0 || group == 0) return;
    gro_tree *t = gro_tree_build(n, d, xs, ys, d == 3 ? zs : NULL);
    double *bodies = (double *)malloc((t->ncells ? t->ncells : 1) * w * sizeof(double));
    aggregate(t, mass, cols, bodies);
    uint32_t *mark = (uint32_t *)malloc((t->ncells ? t->ncells : 1) * sizeof(uint32_t));
    for (size_t c = 0; c < t->ncells; c++) mark[c] = UINT32_MAX;
    for (uint32_t p = 0; p < n; p++) {
        const uint32_t i = t->order[p], g = p / group;
        double xi[MAXD];
        for (uint32_t a = 0; a < d; a++) xi[a] = cols[a][i];
        size_t k = 0;
        while (k < t->ncells) {
            out[1]++;
            if (mark[k] != g) {
                mark[k] = g;
                out[0]++;
            }
            int holds = t->start[k] <= p && p < t->end[k];
            const double *b = &bodies[k * w];
            double dx[MAXD] = {0.0, 0.0, 0.0};
            for (uint32_t a = 0; a < d; a++) dx[a] = xi[a] - b[1 + a];
            double d2 = dx[0] * dx[0] + dx[1] * dx[1];
            if (d == 3) d2 = d2 + dx[2] * dx[2];
            double wd = t->cbox[k].hi[0] - t->cbox[k].lo[0];
            if (!holds && d2 > 0.0 && wd * wd < theta2 * d2) {
                out[2]++;
                k = t->skip[k];
            } else if (t->skip[k] == k + 1) {
                for (uint32_t q = t->start[k]; q < t->end[k]; q++)
                    if (t->order[q] != i) out[3]++;
                k += 1;
            } else {
                k += 1;
            }
        }
    }
    free(mark);
    free(bodies);
    gro_tree_free(t);
}

/* state.rs:208-251: sequential centroid, then the unit vector scaled by gravity*mass. */
static void gravity_pull(gro_fa2 *s) {
    uint32_t n = s->n, d = s->dim;
    if (n == 0) return;
    double mean[MAXD] = {0.0, 0.0, 0.0};
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t a = 0; a < d; a++) mean[a] += P(s, i, a);
    for (uint32_t a = 0; a < d; a++) mean[a] /= (double)n;
    for (uint32_t i = 0; i < n; i++) {
        double c[MAXD];
        for (uint32_t a = 0; a < d; a++) c[a] = P(s, i, a) - mean[a];
        double norm = sqrt(norm2(c, d));
        double g = s->gravity * s->mass[i];
        for (uint32_t a = 0; a < d; a++) {
            double unit = norm > 0.0 ? c[a] / norm : 0.0;
            U(s, i, a) -= g * unit;
        }
    }
}

/* schedule.rs:16-30 */
static void swing_and_traction(const gro_fa2 *s, double *sw, double *tr) {
    double swing = 0.0, traction = 0.0;
    for (uint32_t i = 0; i < s->n; i++) {
        double back[MAXD], fwd[MAXD];
        for (uint32_t a = 0; a < s->dim; a++) {
            back[a] = P(s, i, a) - U(s, i, a);
            fwd[a] = P(s, i, a) + U(s, i, a);
        }
        swing += s->mass[i] * sqrt(norm2(back, s->dim));
        traction += 0.5 * s->mass[i] * sqrt(norm2(fwd, s->dim));
    }
    *sw = swing;
    *tr = traction;
}

/* schedule.rs:33-60: networkx's estimate_factor. */
static void estimate_factor(gro_fa2 *s, double swing, double traction) {
    double n = (double)s->n;
    double jt = s->jitter_tolerance;
    double opt_jitter = 0.05 * sqrt(n);
    double min_jitter = sqrt(opt_jitter);
    double min_speed_efficiency = 0.05;
    double other = fmin(10.0, opt_jitter * traction / (n * n));
    double jitter = jt * fmax(min_jitter, other);
    if (swing / traction > 2.0) {
        if (s->speed_efficiency > min_speed_efficiency) s->speed_efficiency *= 0.5;
        jitter = fmax(jitter, jt);
    }
    double target_speed = swing == 0.0 ? INFINITY : jitter * s->speed_efficiency * traction / swing;
    if (swing > jitter * traction) {
        if (s->speed_efficiency > min_speed_efficiency) s->speed_efficiency *= 0.7;
    } else if (s->speed < 1000.0) {
        s->speed_efficiency *= 1.3;
    }
    s->speed += fmin(target_speed - s->speed, 0.5 * s->speed);
}

/* schedule.rs:62-77 */
static double apply_update(gro_fa2 *s) {
    double moved = 0.0;
    for (uint32_t i = 0; i < s->n; i++) {
        double norm = sqrt(norm2(&s->u[(size_t)i * MAXD], s->dim));
        double factor = s->speed / (1.0 + sqrt(s->speed * s->mass[i] * norm));
        double step = 0.0;
        for (uint32_t a = 0; a < s->dim; a++) {
            double dd = U(s, i, a) * factor;
            P(s, i, a) += dd;
            step += fabs(dd);
        }
        moved += step;
    }
    return moved;
}

static void forces(gro_fa2 *s, uint32_t parts) {
    memset(s->u, 0, (size_t)(s->n ? s->n : 1) * MAXD * sizeof(double));
    if (parts & 1) attraction(s);
    if (parts & 2) {
        if (s->repulsion == 1)
            repulsion_tree(s);
        else
            repulsion_dense(s);
    }
    if (parts & 4) gravity_pull(s);
}

/* state.rs:130-148 */
double gro_fa2_iterate(gro_fa2 *s) {
    forces(s, 7);
    double sw, tr;
    swing_and_traction(s, &sw, &tr);
    s->swing += sw;
    s->traction += tr;
    estimate_factor(s, s->swing, s->traction);
    s->iter_no += 1;
    return apply_update(s);
}

/* state.rs:122-128 */
uint32_t gro_fa2_run(gro_fa2 *s, uint32_t max_iter) {
    uint32_t it = 0;
    for (; it < max_iter;) {
        double moved = gro_fa2_iterate(s);
        it++;
        if (moved < 1e-10) break;
    }
    return it;
}

void gro_fa2_positions(const gro_fa2 *s, double *out) {
    for (uint32_t i = 0; i < s->n; i++)
        for (uint32_t a = 0; a < s->dim; a++) out[(size_t)i * s->dim + a] = P(s, i, a);
}

void gro_fa2_set_positions(gro_fa2 *s, const double *in) {
    for (uint32_t i = 0; i < s->n; i++)
        for (uint32_t a = 0; a < s->dim; a++) P(s, i, a) = in[(size_t)i * s->dim + a];
}

void gro_fa2_mass(const gro_fa2 *s, double *out) { memcpy(out, s->mass, s->n * sizeof(double)); }

void gro_fa2_forces(gro_fa2 *s, uint32_t parts, double *out) {
    forces(s, parts);
    for (uint32_t i = 0; i < s->n; i++)
        for (uint32_t a = 0; a < s->dim; a++) out[(size_t)i * s->dim + a] = U(s, i, a);
}

void gro_fa2_scalars(const gro_fa2 *s, double *out5) {
    out5[0] = s->swing;
    out5[1] = s->traction;
    out5[2] = s->speed;
    out5[3] = s->speed_efficiency;
    out5[4] = (double)s->iter_no;
}

uint32_t gro_fa2_simple_edges(const gro_fa2 *s, uint32_t *lo, uint32_t *hi) {
    if (lo) memcpy(lo, s->lo, s->m * sizeof(uint32_t));
    if (hi) memcpy(hi, s->hi, s->m * sizeof(uint32_t));
    return s->m;
}
