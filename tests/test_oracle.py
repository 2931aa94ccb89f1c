"""Pins the CPU oracle (oracle/) before anything is checked against it.

- dense 2D FA2 + Mulberry32 starts + gate model + simple graph: the reference's pinned
  networkx 3.6 vectors (crates/graph-cli/tests/fixtures/fa2-nx-reference.jsonl), with the
  reference's own negative control (crates/graph-cli/tests/cli_fa2.rs:176-197);
- the quadtree/octree: a brute-force recursive PR-tree written from the definition, on
  small inputs that include coincident, near-coincident, stalled and non-finite points;
- BH repulsion: theta = 0 equals the dense sum within 1e-12 and the reference's error
  bounds at theta = 0.8 (forceatlas2/state/barnes_hut/tests.rs:61-88);
- gravity's association and the coincident-pair guard (forceatlas2/state/tests.rs);
- fmix64 constants via an independently derived inverse; jiggle's invariants
  (rng/tests.rs:97-135).
"""
import json
import math
import os

import numpy as np
import pytest

from tests.refgraphs import bh_graph, errors, line

HERE = os.path.dirname(os.path.abspath(__file__))
NX = os.path.join(HERE, "golden", "fa2-nx-reference.jsonl")


def nx_cases():
    with open(NX) as fh:
        return [json.loads(l) for l in fh]


def gap(ours, case):
    theirs = np.column_stack([case["x"], case["y"]])
    ext = float((theirs.max(0) - theirs.min(0)).max())
    return float(np.abs(ours - theirs).max()) / ext


def test_dense_fa2_matches_the_pinned_networkx_vectors(oracle):
    cases = nx_cases()
    assert len(cases) == 64
    worst = 0.0
    for c in cases:
        n = c["n"]
        assert n == 2 + c["seed"]
        s, t = oracle.gate_edges(n)
        st = oracle.Fa2(n, s, t)
        st.run(c["max_iter"])
        ours = st.positions().astype(np.float32).astype(np.float64)  # settle()'s narrowing
        worst = max(worst, gap(ours, c))
    # cli_fa2.rs:21-28: ceiling 1e-7; the reference port measured 3.156e-08
    assert worst <= 1e-7, worst


def test_the_negative_control_goes_red(oracle):
    worst = 0.0
    for c in nx_cases():
        n = c["n"]
        s, t = oracle.gate_edges(n)
        st = oracle.Fa2(n, s, t, scaling_ratio=3.0)
        st.run(c["max_iter"])
        worst = max(worst, gap(st.positions(), c))
    assert worst > 1e-7 * 1e5, worst


# ------------------------------------------------------------------ tree definition


def cover(lo, hi):
    """The root square: floor of the min corner, doubled until it holds the max."""
    dim = len(lo)
    blo = [math.floor(v) + 1.0 - 1.0 for v in lo]
    bhi = [math.floor(v) + 1.0 for v in lo]
    z = bhi[0] - blo[0] if bhi[0] - blo[0] != 0.0 else 1.0
    while any(blo[a] > hi[a] or hi[a] >= bhi[a] for a in range(dim)):
        z *= 2.0
        for a in range(dim):
            if hi[a] < blo[a]:
                blo[a] = bhi[a] - z
            else:
                bhi[a] = blo[a] + z
        if not (math.isfinite(z) and math.isfinite(bhi[0] - blo[0])):
            break
    return blo, bhi


def pr_tree(pts):
    """Preorder PR-tree by definition: a cell holding one distinct position is a leaf (its
    points most recent first); one holding several is split into its non-empty children in
    slot order, unless halving no longer narrows its x-span, when it is a leaf of them all.
    Non-finite points follow in ascending index."""
    n, dim = pts.shape
    fin = [i for i in range(n) if all(math.isfinite(v) for v in pts[i])]
    cells, order = [], []
    if fin:
        lo = [min(pts[i][a] for i in fin) for a in range(dim)]
        hi = [max(pts[i][a] for i in fin) for a in range(dim)]
        blo, bhi = cover(lo, hi)

        def rec(blo, bhi, members):
            k = len(cells)
            cells.append([list(blo) + list(bhi), 0, len(order), 0])
            distinct = {tuple(pts[i]) for i in members}
            mids = [(blo[a] + bhi[a]) / 2.0 for a in range(dim)]
            stalled = False
            if len(distinct) > 1:
                p = pts[members[0]]
                right = p[0] >= mids[0]
                span = (bhi[0] - mids[0]) if right else (mids[0] - blo[0])
                stalled = not (span < bhi[0] - blo[0])
            if len(distinct) == 1 or stalled:
                order.extend(sorted(members, reverse=True))
                cells[k][1] = k + 1
            else:
                for slot in range(1 << dim):
                    sub = [i for i in members
                           if all((pts[i][a] >= mids[a]) == bool(slot >> a & 1)
                                  for a in range(dim))]
                    if not sub:
                        continue
                    clo = [mids[a] if slot >> a & 1 else blo[a] for a in range(dim)]
                    chi = [bhi[a] if slot >> a & 1 else mids[a] for a in range(dim)]
                    rec(clo, chi, sub)
                cells[k][1] = len(cells)
            cells[k][3] = len(order)

        rec(blo, bhi, fin)
    order.extend(i for i in range(n) if i not in set(fin))
    return cells, order


def tricky_points(seed, n, dim):
    rng = np.random.default_rng(seed)
    kind = seed % 6
    if kind == 0:
        p = rng.uniform(-5, 5, (n, dim))
    elif kind == 1:  # clusters with exact duplicates
        p = rng.integers(0, 6, (n, dim)).astype(float) * 0.5
    elif kind == 2:  # near-coincident: one-ulp neighbours
        base = rng.uniform(0, 1, (n // 2 + 1, dim))
        p = np.concatenate([base, np.nextafter(base, 2.0)])[:n]
    elif kind == 3:  # far from the origin with a tiny spread: the midpoint stops halving
        p = 1e15 + rng.integers(0, 4, (n, dim)).astype(float) * 0.125
    elif kind == 4:  # tiny numbers near zero: deep trees
        p = rng.uniform(-1e-9, 1e-9, (n, dim))
    else:  # non-finite points mixed in
        p = rng.uniform(-3, 3, (n, dim))
        p[rng.integers(0, n, 3), rng.integers(0, dim, 3)] = [np.nan, np.inf, -np.inf]
    return np.ascontiguousarray(p)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", range(18))
def test_the_insertion_tree_is_the_pr_tree(oracle, seed, dim):
    pts = tricky_points(seed, 40 + seed, dim)
    cells, order = pr_tree(pts)
    t = oracle.Tree(pts)
    assert t.ncells == len(cells)
    assert list(t.order) == order
    for k, (b, skip, start, end) in enumerate(cells):
        assert list(t.bounds[k]) == b, k
        assert (t.skip[k], t.start[k], t.end[k]) == (skip, start, end), k


def test_tree_invariants(oracle):
    """quadtree/tests.rs: every internal cell has a child; a subtree is one run of cells
    and one run of points; the order is a permutation."""
    pts = tricky_points(1, 300, 2)
    t = oracle.Tree(pts)
    assert sorted(t.order) == list(range(300))
    for k in range(t.ncells):
        if t.skip[k] != k + 1:
            c = k + 1
            assert c < t.skip[k]
            while c < t.skip[k]:
                assert t.start[k] <= t.start[c] and t.end[c] <= t.end[k]
                c = t.skip[c]
            assert c == t.skip[k]


# ------------------------------------------------------------------ repulsion


def settling(oracle, n, iterations, tree=False):
    s, t = bh_graph(n)
    st = oracle.Fa2(n, s, t, tree=tree)
    st.run(iterations)
    return st


def test_theta_zero_is_the_dense_sum(oracle):
    """barnes_hut/tests.rs:61-70."""
    st = settling(oracle, 400, 5)
    exact = st.forces(2)
    st.set_repulsion(True, 0.0)
    tree = st.forces(2)
    node, total = errors(tree, exact)
    assert node < 1e-12 and total < 1e-12


@pytest.mark.parametrize("n,its", [(1500, 0), (1500, 30), (1500, 100)])
def test_theta_0_8_error_bounds(oracle, n, its):
    """barnes_hut/tests.rs:72-88 (the 5000-node cell runs in the slow GPU suite)."""
    st = settling(oracle, n, its)
    exact = st.forces(2)
    st.set_repulsion(True, oracle.THETA2_REFERENCE)
    tree = st.forces(2)
    node, total = errors(tree, exact)
    assert node < 0.02 and total < 0.05, (node, total)


def test_the_tree_run_is_repeatable_and_finite(oracle):
    """barnes_hut/tests.rs:90-100."""
    a = settling(oracle, 300, 100, tree=True).positions()
    b = settling(oracle, 300, 100, tree=True).positions()
    assert np.array_equal(a, b) and np.isfinite(a).all()


def test_gravity_scales_the_unit_vector(oracle):
    """forceatlas2/state/tests.rs:27-62: -g*(c/norm), not -g*c/norm, bit for bit."""
    s, t = line(20)
    st = oracle.Fa2(20, s, t)
    u = st.forces(4)
    p = st.positions()
    mass = st.mass()
    mean = np.array([sum(p[:, 0].tolist()) / 20, sum(p[:, 1].tolist()) / 20])
    differ = 0
    for i in range(20):
        c = [p[i, 0] - mean[0], p[i, 1] - mean[1]]
        norm = math.sqrt(c[0] * c[0] + c[1] * c[1])
        g = 1.0 * mass[i]
        for a in range(2):
            unit_first = -g * (c[a] / norm)
            offset_first = -g * c[a] / norm
            assert u[i, a] == unit_first
            differ += unit_first != offset_first
    assert differ > 0


def test_a_coincident_pair_is_nudged_not_divided(oracle):
    """forceatlas2/state/tests.rs:69-94, through the dense repulsion."""
    s, t = line(2)
    start = np.array([[0.25, 0.75], [0.25, 0.75]])
    st = oracle.Fa2(2, s, t, start=start)
    f = st.forces(2)
    assert np.isfinite(f).all() and np.abs(f).sum() > 0
    assert np.array_equal(f[0], -f[1])


# ------------------------------------------------------------------ rng


def test_fmix64_constants_invert(oracle):
    """The two multipliers are checked through their modular inverses (a typo in either
    breaks the round trip), and the xorshift-33 steps through their inverse."""
    c1, c2 = 0xFF51AFD7ED558CCD, 0xC4CEB9FE1A85EC53
    M = 1 << 64
    i1, i2 = pow(c1, -1, M), pow(c2, -1, M)

    def unxs(k):
        return k ^ (k >> 33)

    rng = np.random.default_rng(5)
    assert oracle.fmix64(0) == 0
    for _ in range(200):
        x = int(rng.integers(0, 2**63)) * 2 + int(rng.integers(0, 2))
        y = oracle.fmix64(x)
        y = unxs(y)
        y = (y * i2) % M
        y = unxs(y)
        y = (y * i1) % M
        y = unxs(y)
        assert y == x


def test_jiggle_invariants(oracle):
    """rng/tests.rs:97-120."""
    for i in range(50):
        for j in range(3):
            assert oracle.jiggle(7, 3, 0, i, j) == oracle.jiggle(7, 3, 0, j, i)
    a = oracle.jiggle(1, 2, 3, 4, 5)
    assert a != oracle.jiggle(1, 2, 3, 4, 6) and a != oracle.jiggle(1, 2, 4, 4, 5)
    assert a != oracle.jiggle(1, 3, 3, 4, 5) and a != oracle.jiggle(2, 2, 3, 4, 5)
    for n in range(2000):
        v = oracle.jiggle(0x9E3779B9, n // 37, n % 5, n, (n * 31) & 0xFFFFFFFF)
        assert math.isfinite(v) and -5e-7 <= v < 5e-7 and v != 0.0


def test_mulberry32_draws_are_in_range_and_distinct(oracle):
    """synthetic/tests.rs:38-39 and rng/tests.rs:86-87 (range, no repeats)."""
    d = oracle.mulberry32(0x051042, 20000)
    assert ((d >= 0) & (d < 1)).all() and len(set(d.tolist())) == d.size
    assert (d != 0.5).all()
