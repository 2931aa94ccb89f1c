"""d3's many-body pass over the quadtree on the device (gr_d3_charge_bh, gr_quadtree_node_ids)
against the oracle's restatement of barnes_hut/charge.rs, through the C ABI.

Bar: bit for bit. Node ids are integers; the bodies and the walk are FP64 in the reference's
operations and order, so every delta must carry the oracle's exact bits (NaN where a node is
not finite). The inputs include lattices, where exactly zero axes make the walk jiggle keyed
on cell node ids, and the reference's own charge tests (charge/tests.rs)."""
import numpy as np
import pytest

from graph_render_amd import barnes_hut
from tests.test_oracle import tricky_points

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    # every finite or infinite value bit for bit, NaN where the oracle has NaN (a NaN's sign
    # and payload are the hardware's: x86 and the GPU make different default NaNs)
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    if a.shape != b.shape or not np.array_equal(np.isnan(a), np.isnan(b)):
        return False
    a, b = np.where(np.isnan(a), 0.0, a), np.where(np.isnan(b), 0.0, b)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    a, b = np.where(np.isnan(a), 0.0, a), np.where(np.isnan(b), 0.0, b)
    bad = np.nonzero(a.view(np.uint64) != b.view(np.uint64))
    return None if len(bad[0]) == 0 else (bad[0][0], a[bad][0], b[bad][0], len(bad[0]))


@pytest.mark.parametrize("seed", range(18))
def test_node_ids_match_the_insertion(oracle, seed):
    pts = tricky_points(seed, 60 + 23 * seed, 2)
    print(f"d3 node ids seed={seed} n={len(pts)}", flush=True)
    got = barnes_hut.node_ids(pts)
    want = oracle.Tree(pts).node_ids()
    assert np.array_equal(got, want), (seed, got[:20], want[:20])


def test_node_ids_of_the_reference_diagonal():
    # quadtree/tests.rs:189-215: one id per cell, ids below the shape count, some at or past
    # the point count (an internal cell's slot), and each leaf holds its chain head
    pts = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [3.0, 3.0]])
    ids = barnes_hut.node_ids(pts)
    assert sorted(ids.tolist()) == list(range(len(ids)))
    assert ids.max() >= 4


def test_node_ids_at_scale(oracle):
    rng = np.random.default_rng(11)
    pts = rng.normal(0.0, 100.0, (200000, 2))
    pts[::7] = np.round(pts[::7])  # lattice points and exact coincidences among them
    print("d3 node ids n=200000", flush=True)
    assert np.array_equal(barnes_hut.node_ids(pts), oracle.Tree(pts).node_ids())


def lattice(n, side, rng):
    return rng.integers(0, side, (n, 2)).astype(np.float64)


CASES = {
    "normal": lambda r: r.normal(0.0, 80.0, (3000, 2)),
    "lattice": lambda r: lattice(4000, 40, r),          # many exact zero axes and chains
    "fine lattice": lambda r: lattice(2000, 400, r) * 0.25,
    "clusters": lambda r: np.concatenate([r.normal(0, 0.5, (1500, 2)),
                                          r.normal(300, 2.0, (1500, 2))]),
    "line": lambda r: np.stack([r.uniform(0, 1000, 2000), np.full(2000, 7.0)], 1),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("theta", [0.9, 0.0, 1.5])
def test_charge_pass_is_the_oracle_bit_for_bit(oracle, case, theta):
    rng = np.random.default_rng(len(case))
    xy = CASES[case](rng)
    print(f"d3 charge case={case} theta={theta} n={len(xy)}", flush=True)
    kw = dict(theta=theta, charge=-90.0, alpha=0.37, distance_min=1.0, distance_max=520.0,
              seed=0x9E3779B9, tick=5)
    got = barnes_hut.charge_pass(xy, **kw)
    want = oracle.d3_charge_bh(xy, **kw)
    assert same_bits(got, want), first_diff(got, want)
    if case == "line":
        # every gap's y is exactly zero: each term's y is a jiggle keyed on the node and the
        # cell's node id (or the chain head), so these bits pin the node ids too
        assert (got[:, 1] != 0.0).all()


@pytest.mark.parametrize("seed", range(12))
def test_charge_pass_on_tricky_points(oracle, seed):
    pts = tricky_points(seed, 80 + 29 * seed, 2)
    print(f"d3 charge tricky seed={seed} n={len(pts)}", flush=True)
    kw = dict(theta=0.9, charge=-30.0, alpha=1.0, distance_min=1e-3,
              distance_max=float("inf"), seed=seed, tick=seed)
    got = barnes_hut.charge_pass(pts, **kw)
    want = oracle.d3_charge_bh(pts, **kw)
    assert same_bits(got, want), first_diff(got, want)


def test_charge_pass_at_scale(oracle):
    rng = np.random.default_rng(12)
    xy = rng.normal(0.0, 300.0, (300000, 2))
    xy[::5] = np.round(xy[::5])
    print("d3 charge n=300000", flush=True)
    got = barnes_hut.charge_pass(xy, seed=3, tick=40)
    want = oracle.d3_charge_bh(xy, theta=0.9, charge=-90.0, alpha=1.0, distance_min=1.0,
                               distance_max=520.0, seed=3, tick=40)
    assert same_bits(got, want), first_diff(got, want)


def test_a_leaf_past_distance_max_contributes_nothing():
    # charge/tests.rs:14-50: a point 1000 away with distanceMax 520 adds nothing
    got = barnes_hut.charge_pass(np.array([[1000.0, 0.0], [0.0, 0.0]]), charge=-90.0,
                                 distance_max=520.0)
    assert (got == 0.0).all()


def test_zero_theta_with_coincident_nodes_stays_finite(oracle):
    # charge/tests.rs:91-123
    xy = np.array([[1.0, 2.0], [1.0, 2.0], [1.0, 2.0], [40.0, 3.0], [41.0, 9.0]])
    kw = dict(theta=0.0, charge=-30.0, distance_min=1.0, distance_max=1e9, seed=0x9E3779B9,
              tick=3)
    got = barnes_hut.charge_pass(xy, **kw)
    assert np.isfinite(got).all()
    assert same_bits(got, oracle.d3_charge_bh(xy, **kw))


def test_extreme_magnitudes_get_a_usable_threshold(oracle):
    # charge/tests.rs:125-161: theta 1e200 over cells 1e200 wide
    xy = np.array([[1e200, 1e200], [-1e200, 1e200], [1e200, -1e200], [0.5, -0.25]])
    kw = dict(theta=1e200, charge=-30.0, distance_min=1.0, distance_max=float("inf"),
              seed=0x9E3779B9, tick=3)
    got = barnes_hut.charge_pass(xy, **kw)
    assert np.isfinite(got).all()
    assert same_bits(got, oracle.d3_charge_bh(xy, **kw))


def test_a_coincident_pair_gives_finite_deltas(oracle):
    # charge/tests.rs:163-205
    xy = np.array([[1.0, 2.0], [1.0, 2.0]])
    kw = dict(theta=0.9, charge=-30.0, distance_min=1.0, distance_max=1e9)
    got = barnes_hut.charge_pass(xy, **kw)
    assert np.isfinite(got).all()
    assert same_bits(got, oracle.d3_charge_bh(xy, **kw))


def test_a_single_node_and_non_finite_nodes(oracle):
    assert (barnes_hut.charge_pass(np.array([[3.0, 4.0]])) == 0.0).all()
    # no finite point: an empty tree, so every node's walk adds nothing (node_delta's loop
    # reads no body)
    got = barnes_hut.charge_pass(np.array([[np.nan, 0.0], [np.inf, 1.0]]))
    assert (got == 0.0).all()
    # a NaN query approximates every cell it meets (open >= NaN is false) and comes out NaN;
    # an infinite one falls past distanceMax everywhere and comes out zero
    xy = np.array([[0.0, 0.0], [5.0, 1.0], [np.nan, 2.0], [np.inf, 3.0], [2.0, -np.inf]])
    got = barnes_hut.charge_pass(xy)
    assert np.isnan(got[2]).all() and (got[3] == 0.0).all() and (got[4] == 0.0).all()
    assert same_bits(got, oracle.d3_charge_bh(xy, charge=-90.0, distance_max=520.0))


def test_the_pass_is_deterministic():
    rng = np.random.default_rng(4)
    xy = lattice(50000, 300, rng)
    a = barnes_hut.charge_pass(xy, seed=1, tick=2)
    b = barnes_hut.charge_pass(xy, seed=1, tick=2)
    assert same_bits(a, b)
    c = barnes_hut.charge_pass(xy, seed=1, tick=3)  # another tick, other jiggles
    assert not same_bits(a, c)
