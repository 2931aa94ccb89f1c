"""More pins of the CPU oracle, for the parts the reference's own vectors do not reach.

- Mulberry32 at seeds other than 0: draws of an independent JavaScript implementation of
  the published generator (tests/golden/make_mulberry32_vectors.js), bit for bit.
- The 3D arm (dense run_3d and the octree walk): a 3D state whose z coordinates are all
  equal IS the 2D state (the networkx-pinned arm) bit for bit, forces and trajectories;
  permuting the axes permutes the forces (which catches a z term read from the wrong
  axis); the octree at theta 0 equals the dense 3D sum.
- Edge weights: unit weights give the unweighted bits; a power-of-two scale of every weight
  scales the attraction exactly; one edge's weight moves only its two endpoints' rows; a
  repeated raw edge's weight is ignored (the first occurrence is kept); masses do not
  depend on weights; the attraction is linear in the weights.
- The particle mesh's rung ladder h = 2^(q/4) against the C library's exp2 (what the
  reference calls, particle_mesh/frame.rs:75) for every quarter step.
"""
import json
import math
import os

import numpy as np
import pytest

from tests.refgraphs import bh_graph, errors

HERE = os.path.dirname(os.path.abspath(__file__))


# ------------------------------------------------------------------ Mulberry32


def test_mulberry32_matches_the_published_generator_at_many_seeds(oracle):
    vec = json.load(open(os.path.join(HERE, "golden", "mulberry32_vectors.json")))
    assert len(vec) >= 8
    for seed, draws in vec.items():
        got = oracle.mulberry32(int(seed), len(draws))
        assert np.array_equal(got, np.array(draws)), seed


def test_mulberry32_negative_control(oracle):
    vec = json.load(open(os.path.join(HERE, "golden", "mulberry32_vectors.json")))
    got = oracle.mulberry32(8, 64)  # a neighbouring seed reproduces none of seed 7's draws
    assert not np.isin(got, np.array(vec["7"])).any()


def test_initial_positions_are_row_major_mulberry32_draws(oracle):
    vec = json.load(open(os.path.join(HERE, "golden", "mulberry32_vectors.json")))
    for seed in (7, 123456789):
        for dim in (2, 3):
            p = oracle.initial_positions(64 // dim, seed, dim)
            assert np.array_equal(p.reshape(-1), np.array(vec[str(seed)][: p.size]))


# ------------------------------------------------------------------ 3D arm


def lifted(p2):
    return np.column_stack([p2, np.zeros(p2.shape[0])])


@pytest.mark.parametrize("tree", [False, True])
def test_3d_with_equal_z_is_the_2d_state(oracle, tree):
    n = 1200
    s, t = oracle.gate_edges(n)
    two = oracle.Fa2(n, s, t, tree=tree)
    three = oracle.Fa2(n, s, t, dim=3, start=lifted(two.positions()), tree=tree)
    for it in range(12):
        f2, f3 = two.forces(7), three.forces(7)
        assert np.array_equal(f3[:, :2], f2), it
        assert not f3[:, 2].any()
        two.iterate()
        three.iterate()
        p2, p3 = two.positions(), three.positions()
        assert np.array_equal(p3[:, :2], p2) and not p3[:, 2].any(), it
    assert np.array_equal(three.scalars(), two.scalars())


def test_3d_equal_z_tree_is_the_quadtree(oracle):
    rng = np.random.default_rng(4)
    p2 = rng.normal(0.0, 20.0, (3000, 2))
    q, o = oracle.Tree(p2), oracle.Tree(lifted(p2))
    assert q.ncells == o.ncells
    for a in ("order", "skip", "start", "end"):
        assert np.array_equal(getattr(q, a), getattr(o, a)), a
    assert np.array_equal(o.bounds[:, [0, 1, 3, 4]], q.bounds)
    mass = rng.integers(1, 9, 3000).astype(np.float64)
    assert np.array_equal(o.bodies(mass)[:, :3], q.bodies(mass))


@pytest.mark.parametrize("tree", [False, True])
def test_3d_axis_permutation_permutes_the_forces(oracle, tree):
    n = 900
    s, t = bh_graph(n)
    p = np.random.default_rng(9).uniform(-50.0, 50.0, (n, 3))
    perm = [1, 2, 0]
    a = oracle.Fa2(n, s, t, dim=3, start=p, tree=tree, theta2=0.0 if tree else 0.64)
    b = oracle.Fa2(n, s, t, dim=3, start=p[:, perm], tree=tree, theta2=0.0 if tree else 0.64)
    fa, fb = a.forces(7), b.forces(7)
    node, field = errors(fb, fa[:, perm])
    assert node < 1e-12 and field < 1e-12, (node, field)
    # a force that ignored z, or took it from the wrong axis, fails this by orders of magnitude
    assert np.abs(fa[:, 2]).max() > 1e-3 * np.abs(fa).max()


def test_3d_octree_at_theta_zero_is_the_dense_sum(oracle):
    n = 700
    s, t = bh_graph(n)
    p = np.random.default_rng(3).normal(0.0, 10.0, (n, 3))
    dense = oracle.Fa2(n, s, t, dim=3, start=p)
    octree = oracle.Fa2(n, s, t, dim=3, start=p, tree=True, theta2=0.0)
    node, field = errors(octree.forces(2), dense.forces(2))
    assert node < 1e-12 and field < 1e-12


# ------------------------------------------------------------------ edge weights


def weighted_case(n=800, seed=2):
    s, t = bh_graph(n)
    rng = np.random.default_rng(seed)
    return n, s, t, rng


def test_unit_weights_give_the_unweighted_bits(oracle):
    n, s, t, _ = weighted_case()
    a = oracle.Fa2(n, s, t, tree=True)
    b = oracle.Fa2(n, s, t, tree=True, weights=np.ones(s.size))
    assert np.array_equal(a.forces(7), b.forces(7))
    a.run(20)
    b.run(20)
    assert np.array_equal(a.positions(), b.positions())


def test_power_of_two_weights_scale_the_attraction_exactly(oracle):
    n, s, t, rng = weighted_case()
    w = rng.lognormal(0.0, 0.5, s.size)
    base = oracle.Fa2(n, s, t, weights=w).forces(1)
    for k in (-3, 1, 5):
        f = oracle.Fa2(n, s, t, weights=w * 2.0 ** k).forces(1)
        assert np.array_equal(f, base * 2.0 ** k), k


def test_one_weight_moves_only_its_endpoints(oracle):
    n, s, t, rng = weighted_case()
    w = rng.lognormal(0.0, 0.5, s.size)
    base = oracle.Fa2(n, s, t, weights=w).forces(1)
    for e in rng.choice(s.size, 5, replace=False):
        w2 = w.copy()
        w2[e] *= 3.0
        f = oracle.Fa2(n, s, t, weights=w2).forces(1)
        moved = np.nonzero((f != base).any(1))[0]
        assert set(moved.tolist()) == {int(s[e]), int(t[e])}, e


def test_a_repeated_edge_keeps_its_first_weight(oracle):
    n, s, t, rng = weighted_case()
    w = rng.lognormal(0.0, 0.5, s.size)
    s2 = np.concatenate([s, t[:10], s[:3]])  # reversed and repeated copies of early edges
    t2 = np.concatenate([t, s[:10], t[:3]])
    w2 = np.concatenate([w, rng.lognormal(0.0, 0.5, 13)])
    a = oracle.Fa2(n, s, t, weights=w)
    b = oracle.Fa2(n, s2, t2, weights=w2)
    assert np.array_equal(a.forces(1), b.forces(1))
    assert np.array_equal(a.mass(), b.mass())


def test_mass_ignores_weights_and_attraction_is_linear_in_them(oracle):
    n, s, t, rng = weighted_case()
    w1 = rng.lognormal(0.0, 0.5, s.size)
    w2 = rng.lognormal(0.0, 0.5, s.size)
    a, b = oracle.Fa2(n, s, t, weights=w1), oracle.Fa2(n, s, t, weights=w2)
    c, u = oracle.Fa2(n, s, t, weights=w1 + w2), oracle.Fa2(n, s, t)
    for x in (a, b, c):
        assert np.array_equal(x.mass(), u.mass())
    lo, hi = np.minimum(s, t).astype(np.int64), np.maximum(s, t).astype(np.int64)
    pairs = np.unique(lo[lo != hi] * n + hi[lo != hi])  # simple edges: loops and repeats out
    deg = np.bincount(np.concatenate([pairs // n, pairs % n]), minlength=n)
    assert np.array_equal(u.mass(), deg + 1.0)
    f = a.forces(1) + b.forces(1)
    g = c.forces(1)
    assert np.abs(f - g).max() <= 1e-12 * np.abs(g).max()


# ------------------------------------------------------------------ particle-mesh rungs


def test_rung_ladder_is_the_c_library_exp2(oracle):
    for q in range(-96, 97):
        assert oracle.lib().gro_pm_h(q) == np.exp2(q / 4.0), q
        # and each quarter step is what it claims: four of them make one octave
        assert math.isclose(oracle.lib().gro_pm_h(q) ** 4, 2.0 ** q, rel_tol=1e-15)


def _walk_brute(pts, mass, theta2, group):
    """The reference walk in plain Python over the oracle's tree, each visit recorded."""
    from oracle import oracle as O
    t = O.Tree(pts)
    bodies = t.bodies(mass)
    n = len(pts)
    visits = [[] for _ in range(n)]
    acc = pairs = 0
    for p in range(n):
        i = int(t.order[p])
        k = 0
        while k < t.ncells:
            visits[p].append(k)
            holds = t.start[k] <= p < t.end[k]
            dx = pts[i] - bodies[k, 1:]
            d2 = float(dx[0] * dx[0] + dx[1] * dx[1])
            w = t.bounds[k, 2] - t.bounds[k, 0]
            if not holds and d2 > 0 and w * w < theta2 * d2:
                acc += 1
                k = int(t.skip[k])
            else:
                if t.skip[k] == k + 1:
                    pairs += sum(1 for q in range(t.start[k], t.end[k]) if t.order[q] != i)
                k += 1
    steps = sum(len(set().union(*visits[g:g + group])) for g in range(0, n, group))
    return dict(steps=steps, acting_lanes=sum(map(len, visits)), accepted=acc, pairs=pairs)


@pytest.mark.parametrize("theta2", [0.64, 1.44])
def test_walk_counts_are_the_replayed_walks_unions(oracle, theta2):
    # the scan of 64 consecutive tree positions visits exactly the union of their walks
    rng = np.random.default_rng(4)
    pts = np.concatenate([rng.normal(0, 30, (500, 2)), rng.uniform(-200, 200, (300, 2))])
    pts[7] = pts[8]  # a coincident pair
    mass = rng.integers(1, 6, len(pts)).astype(np.float64)
    assert oracle.walk_counts(pts, mass, theta2, 64) == _walk_brute(pts, mass, theta2, 64)


def test_walk_counts_invariants(oracle):
    rng = np.random.default_rng(9)
    pts = rng.uniform(-100, 100, (2000, 2))
    mass = np.ones(2000)
    c1 = oracle.walk_counts(pts, mass, 1.44, 1)
    assert c1["steps"] == c1["acting_lanes"]  # one position per group: every visit a step
    prev = c1["steps"]
    for g in (2, 8, 64, 256, 2000):
        c = oracle.walk_counts(pts, mass, 1.44, g)
        assert {k: c[k] for k in ("acting_lanes", "accepted", "pairs")} == \
            {k: c1[k] for k in ("acting_lanes", "accepted", "pairs")}
        assert c["steps"] <= prev and c["steps"] * g >= c["acting_lanes"]  # merging groups
        prev = c["steps"]
    # theta 0 accepts nothing: every position walks the whole arena and meets every other
    ncells = oracle.Tree(pts).ncells
    c0 = oracle.walk_counts(pts, mass, 0.0, 64)
    assert c0["accepted"] == 0 and c0["pairs"] == 2000 * 1999
    assert c0["acting_lanes"] == 2000 * ncells and c0["steps"] == 32 * ncells
