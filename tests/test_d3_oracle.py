"""Pins of the oracle's d3 many-body pass over the quadtree (barnes_hut/charge.rs) and of
the cell node ids it keys its jiggle on.

- Node ids: a brute-force insertion written here from the definition (add.js's order: the
  root leaf, then per point an empty-slot leaf, or the split's internal cells top down
  followed by the new leaf) gives every cell's shape slot; the oracle's arena carries the
  same ids. The same ids are the rank of each cell by (the insertion that created it, its
  depth, internal before leaf), the rule the device computes them by.
- theta 0 opens every cell, so the pass is the direct d3 pair sum (distanceMin raised,
  distanceMax cut), summed here in numpy."""
import math

import numpy as np
import pytest

from tests.test_oracle import tricky_points


def insertion_ids(pts):
    """(cell -> shape slot) in preorder, by replaying the insertion: shapes are pushed in
    creation order; each holds its region and children by slot."""
    n = len(pts)
    fin = [i for i in range(n) if all(math.isfinite(v) for v in pts[i])]
    if not fin:
        return []
    lo = [min(pts[i][a] for i in fin) for a in range(2)]
    hi = [max(pts[i][a] for i in fin) for a in range(2)]
    # the root square: floor of the min corner, doubled until it holds the max
    blo = [math.floor(v) for v in lo]
    bhi = [v + 1.0 for v in blo]
    z = 1.0
    while any(blo[a] > hi[a] or hi[a] >= bhi[a] for a in range(2)):
        z *= 2.0
        for a in range(2):
            if hi[a] < blo[a]:
                blo[a] = bhi[a] - z
            else:
                bhi[a] = blo[a] + z
    shapes = []  # [kind, head or children]

    def push(kind, val):
        shapes.append([kind, val])
        return len(shapes) - 1

    def slot(b, p):
        xm, ym = (b[0] + b[2]) / 2, (b[1] + b[3]) / 2
        return (1 if p[0] >= xm else 0) | (2 if p[1] >= ym else 0)

    def narrow(b, s):
        xm, ym = (b[0] + b[2]) / 2, (b[1] + b[3]) / 2
        x0, y0, x1, y1 = b
        x0, x1 = (xm, x1) if s & 1 else (x0, xm)
        y0, y1 = (ym, y1) if s & 2 else (y0, ym)
        return [x0, y0, x1, y1]

    root = None
    rb = [blo[0], blo[1], bhi[0], bhi[1]]
    for i in fin:
        p = pts[i]
        if root is None:
            root = push("leaf", i)
            continue
        node, b, parent = root, list(rb), None
        while shapes[node][0] == "int":
            s = slot(b, p)
            ch = shapes[node][1][s]
            if ch is None:
                leaf = push("leaf", i)
                shapes[node][1][s] = leaf
                node = None
                break
            parent, node, b = (node, s), ch, narrow(b, s)
        if node is None:
            continue
        head = shapes[node][1]
        q = pts[head]
        if q[0] == p[0] and q[1] == p[1]:
            continue  # chained
        at = parent
        while True:
            before = b[2] - b[0]
            si, sj = slot(b, p), slot(b, q)
            nb = narrow(b, si)
            if not (nb[2] - nb[0] < before) or (si == sj and not (narrow(b, sj)[2] - narrow(b, sj)[0] < before)):
                break  # bail: left to the oracle's tricky cases below
            internal = push("int", [None] * 4)
            if at is None:
                root = internal
            else:
                shapes[at[0]][1][at[1]] = internal
            if si != sj:
                shapes[internal][1][sj] = node
                shapes[internal][1][si] = push("leaf", i)
                break
            at, b = (internal, si), nb
    out = []

    def visit(v):
        out.append(v)
        if shapes[v][0] == "int":
            for c in shapes[v][1]:
                if c is not None:
                    visit(c)
    visit(root)
    return out


def pushers(tree):
    """Per cell: the point whose insertion pushed it, its depth and whether it is a leaf."""
    C = tree.ncells
    skip, start, end, order = tree.skip, tree.start, tree.end, tree.order
    depth = np.zeros(C, np.int64)
    for k in range(C):
        if skip[k] != k + 1:
            c = k + 1
            while c < skip[k]:
                depth[c] = depth[k] + 1
                c = skip[c]
    a = np.zeros(C, np.int64)
    b = np.full(C, 1 << 40, np.int64)
    for k in range(C - 1, -1, -1):
        if skip[k] == k + 1:
            pts = order[start[k]:end[k]]
            a[k] = pts.min()
            others = [int(q) for q in pts if not np.array_equal(tree.pos[q], tree.pos[a[k]])]
            b[k] = min(others) if others else 1 << 40
        else:
            kids = []
            c = k + 1
            while c < skip[k]:
                kids.append(c)
                c = skip[c]
            a[k] = min(a[c] for c in kids)
            first = next(c for c in kids if a[c] == a[k])
            b[k] = min([b[first]] + [a[c] for c in kids if c != first])
    leaf = np.array([skip[k] == k + 1 for k in range(C)])
    return np.where(leaf, a, b), depth, leaf


def rank_ids(tree):
    """Node ids as the rank of (creating insertion, depth, leaf flag), from the arena alone."""
    C = tree.ncells
    t, depth, leaf = pushers(tree)
    key = sorted(range(C), key=lambda k: (t[k], depth[k], leaf[k]))
    ids = np.zeros(C, np.int64)
    for r, k in enumerate(key):
        ids[k] = r
    return ids


def test_node_ids_replay_the_insertion(oracle):
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(1, 120))
        pts = rng.normal(0.0, 5.0, (n, 2))
        pts[rng.integers(0, n, n // 5)] = pts[rng.integers(0, n, n // 5)]  # coincidences
        t = oracle.Tree(pts)
        assert t.node_ids().tolist() == insertion_ids(pts.tolist()), trial
    # the reference's own example: four points on a diagonal (quadtree/tests.rs:189-215)
    t = oracle.Tree(np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [3.0, 3.0]]))
    ids = t.node_ids()
    assert sorted(ids.tolist()) == list(range(t.ncells)) and ids.max() >= 4


def count_ids(tree, n):
    """The same ids as the device forms them, without a sort: cells pushed by earlier points,
    plus the rank inside the point's own group (its internal cells sit at consecutive depths;
    its leaf, which later splits may have moved down, comes last)."""
    t, depth, leaf = pushers(tree)
    cnt = np.bincount(t, minlength=n)
    base = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    low = np.full(n, 1 << 40)
    np.minimum.at(low, t, depth)
    return np.where(leaf, base[t] + cnt[t] - 1, base[t] + depth - low[t])


@pytest.mark.parametrize("seed", range(8))
def test_node_ids_are_the_creation_rank(oracle, seed):
    pts = tricky_points(seed, 150 + 31 * seed, 2)
    t = oracle.Tree(pts)
    t.pos = pts
    assert np.array_equal(t.node_ids(), rank_ids(t))
    assert np.array_equal(t.node_ids(), count_ids(t, len(pts)))


def test_a_moved_leaf_keeps_the_last_rank_of_its_group(oracle):
    # a leaf pushed below its point's internal cells and then moved down by later splits:
    # its depth is no longer next to theirs, so a rank by depth offset would skip ids
    pts = tricky_points(0, 60, 2)
    t = oracle.Tree(pts)
    t.pos = pts
    tt, depth, leaf = pushers(t)
    moved = [k for k in range(t.ncells) if leaf[k] and
             any(tt[j] == tt[k] and not leaf[j] and depth[k] - depth[j] > 1
                 and depth[j] == max(depth[q] for q in range(t.ncells)
                                     if tt[q] == tt[k] and not leaf[q])
                 for j in range(t.ncells))]
    assert moved
    assert np.array_equal(t.node_ids(), count_ids(t, len(pts)))


def d3_direct(xy, charge, alpha, dmin, dmax, oracle, seed=0, tick=0):
    n = xy.shape[0]
    out = np.zeros((n, 2))
    for i in range(n):
        for j in range(n):
            if j == i:
                continue
            dx, dy = xy[j, 0] - xy[i, 0], xy[j, 1] - xy[i, 1]
            l = dx * dx + dy * dy
            if l >= dmax * dmax:
                continue
            if dx == 0.0:
                dx = oracle.jiggle(seed, tick, 2, i, j)
                l += dx * dx
            if dy == 0.0:
                dy = oracle.jiggle(seed, tick, 3, i, j)
                l += dy * dy
            if l < dmin * dmin:
                l = math.sqrt(dmin * dmin * l)
            f = charge * alpha / l
            out[i] += (dx * f, dy * f)
    return out


@pytest.mark.parametrize("dmax", [float("inf"), 30.0])
def test_theta_zero_is_the_direct_pair_sum(oracle, dmax):
    rng = np.random.default_rng(5)
    xy = rng.normal(0.0, 20.0, (300, 2))
    got = oracle.d3_charge_bh(xy, theta=0.0, charge=-30.0, alpha=0.7, distance_min=1.0,
                              distance_max=dmax)
    want = d3_direct(xy, -30.0, 0.7, 1.0, dmax, oracle)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_theta_bounds_the_approximation(oracle):
    rng = np.random.default_rng(6)
    xy = rng.normal(0.0, 20.0, (2000, 2))
    exact = oracle.d3_charge_bh(xy, theta=0.0)
    for theta, bar in ((0.5, 0.01), (0.9, 0.05)):
        got = oracle.d3_charge_bh(xy, theta=theta)
        rel = np.linalg.norm(got - exact) / np.linalg.norm(exact)
        assert 0 < rel < bar, (theta, rel)
