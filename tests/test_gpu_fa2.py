"""GPU parity of the HIP FA2 path against the CPU oracle, through the C ABI.

Bars: bit-exact for the tree (keys, order, topology, bounds) and for every FP64 force part
whose reduction order the reference fixes per body (attraction, tree repulsion, dense
repulsion, centres of mass); within 1e-12 where a global sum (centroid, swing, traction)
enters; FP32 forces within 1e-5 of FP64 (both of barnes_hut/tests.rs:34-47's metrics).
"""
import json
import os

import numpy as np
import pytest

import graph_render_amd as G
from tests.refgraphs import bh_graph, errors, line
from tests.test_oracle import nx_cases, gap, tricky_points

pytestmark = pytest.mark.gpu

THETA_REF = 0.8


def session(n, s, t, dim=2, start=None, theta=None, precision="fp64", repulsion="tree",
            weights=None, export=True, **params):
    p = G.Fa2Params(dim=dim, **params)
    x = G.Fa2Ext(theta=theta, precision=precision, repulsion=repulsion, start=start,
                 edge_weights=weights, export_tree=export)
    return G.Session(G.Topology(n, s, t), p, x)


def oracle_state(oracle, n, s, t, dim=2, start=None, tree=True, theta=THETA_REF,
                 weights=None, **params):
    return oracle.Fa2(n, s, t, dim=dim, start=start, tree=tree, theta2=theta * theta,
                      weights=weights, **params)


# ------------------------------------------------------------------ tree


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("seed", range(12))
def test_tree_is_bit_exact(oracle, seed, dim):
    pts = tricky_points(seed, 200 + 37 * seed, dim)
    n = pts.shape[0]
    print(f"tree case seed={seed} dim={dim} n={n}", flush=True)
    s, t = np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    with session(n, s, t, dim=dim, start=pts) as sess:
        sess.forces(2)
        tr = sess.tree()
    ref = oracle.Tree(pts)
    assert tr["ncells"] == ref.ncells
    assert np.array_equal(tr["order"], ref.order)
    assert np.array_equal(tr["skip"], ref.skip)
    assert np.array_equal(tr["start"], ref.start)
    assert np.array_equal(tr["end"], ref.end)
    assert np.array_equal(tr["bounds"], ref.bounds)
    mass = np.ones(n)
    rb = ref.bodies(mass)
    assert np.array_equal(tr["bodies"], rb, equal_nan=True)


@pytest.mark.parametrize("n", [20000, 200000])
def test_tree_on_a_layout_is_bit_exact(oracle, n):
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t)
    ref.run(3)
    p = ref.positions()
    with session(n, s, t, start=p) as sess:
        sess.forces(2)
        tr = sess.tree()
    rt = oracle.Tree(p)
    assert tr["ncells"] == rt.ncells
    assert np.array_equal(tr["order"], rt.order)
    assert np.array_equal(tr["skip"], rt.skip)
    assert np.array_equal(tr["bounds"], rt.bounds)
    assert np.array_equal(tr["bodies"], rt.bodies(ref.mass()))


@pytest.mark.parametrize("offset", [0.0, 3.0e9, 1.0e13])
def test_cell_widths_and_keys_on_and_off_dyadic_roots(oracle, offset):
    # Without exported bounds a session takes cell widths from the level and keys from one
    # product per axis when the root square is dyadic (offsets 0 and 3e9); at 1e13 the root
    # is too far out for that and both fall back to the reference's descent. Either way the
    # repulsion must be the oracle's, bit for bit, with and without exported bounds.
    n = 5000
    s, t = bh_graph(n)
    rng = np.random.default_rng(7)
    pts = rng.normal(0.0, 30.0, (n, 2)) + offset
    print(f"dyadic case offset={offset}", flush=True)
    with session(n, s, t, start=pts, export=False) as a:
        fa = a.forces(2)
    with session(n, s, t, start=pts, export=True) as b:
        fb = b.forces(2)
    ref = oracle_state(oracle, n, s, t, start=pts)
    fr = ref.forces(2)
    assert np.array_equal(fa, fb)
    assert np.array_equal(fa, fr)


# ------------------------------------------------------------------ forces


@pytest.mark.parametrize("n", [2, 3, 61, 600, 5000])
def test_force_parts_are_bit_exact(oracle, n):
    s, t = oracle.gate_edges(n)
    ref = oracle_state(oracle, n, s, t)
    ref.run(4)
    p = ref.positions()
    with session(n, s, t, start=p) as sess:
        for parts in (1, 2, 3):
            g = sess.forces(parts)
            r = ref.forces(parts)
            assert np.array_equal(g, r), (n, parts, np.abs(g - r).max())
        g = sess.forces(7)
    r = ref.forces(7)
    scale = np.abs(r).max()
    assert np.abs(g - r).max() <= 1e-12 * scale


def test_dense_repulsion_is_bit_exact(oracle):
    n = 700
    s, t = oracle.gate_edges(n)
    ref = oracle_state(oracle, n, s, t, tree=False)
    ref.run(3)
    p = ref.positions()
    with session(n, s, t, start=p, repulsion="dense") as sess:
        for parts in (1, 2, 3):
            assert np.array_equal(sess.forces(parts), ref.forces(parts)), parts


def test_weighted_attraction_is_bit_exact(oracle):
    n = 3000
    s, t = bh_graph(n)
    w = np.random.default_rng(3).lognormal(0.0, 0.5, s.size)
    ref = oracle_state(oracle, n, s, t, weights=w)
    with session(n, s, t, start=ref.positions(), weights=w) as sess:
        assert np.array_equal(sess.forces(1), ref.forces(1))
        assert np.array_equal(sess.forces(3), ref.forces(3))


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("precision", ["fp64", "fp32_kahan"])
def test_hub_rows_attraction(oracle, weighted, precision):
    # hubs of degree 2999, 700 and 129..131 take the long-row path (more than 128 entries),
    # the rest the per-lane path; FP64 must stay bit-exact, both precisions one sequence
    n = 3000
    rng = np.random.default_rng(21)
    hub = [(0, j) for j in range(1, n)]
    mid = [(5, j) for j in rng.choice(np.arange(6, n), 700, replace=False)]
    edge = [(7 + k, j) for k in range(3) for j in range(100, 100 + 129 + k)]
    rest = [tuple(e) for e in rng.integers(0, n, (2 * n, 2))]
    s, t = np.array(hub + mid + edge + rest, np.uint32).T
    w = rng.lognormal(0.0, 0.5, s.size) if weighted else None
    ref = oracle_state(oracle, n, s, t, weights=w)
    ref.run(2)
    print(f"hub rows weighted={weighted} precision={precision}", flush=True)
    with session(n, s, t, start=ref.positions(), weights=w, precision=precision) as sess:
        got = sess.forces(1)
    want = ref.forces(1)
    if precision == "fp64":
        assert np.array_equal(got, want)
    else:
        p32 = ref.positions().astype(np.float32).astype(np.float64)
        with session(n, s, t, start=p32, weights=w) as s64:
            want = s64.forces(1)
        _, field = errors(got, want)
        assert field <= 1e-5


def test_theta_zero_equals_dense(oracle):
    """barnes_hut/tests.rs:61-70 on the device."""
    n = 400
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t, tree=False)
    ref.run(5)
    exact = ref.forces(2)
    with session(n, s, t, start=ref.positions(), theta=0.0) as sess:
        tree = sess.forces(2)
    node, total = errors(tree, exact)
    assert node < 1e-12 and total < 1e-12


@pytest.mark.parametrize("n,its", [(1500, 0), (1500, 30), (1500, 100), (5000, 100)])
def test_theta_bounds_on_device(oracle, n, its):
    """barnes_hut/tests.rs:72-88: dense state, tree repulsion within 2% / 5%."""
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t, tree=False)
    ref.run(its)
    exact = ref.forces(2)
    with session(n, s, t, start=ref.positions(), repulsion="dense") as dense:
        assert np.array_equal(dense.forces(2), exact)
    with session(n, s, t, start=ref.positions()) as sess:
        node, total = errors(sess.forces(2), exact)
    assert node < 0.02 and total < 0.05


# ------------------------------------------------------------------ trajectories


@pytest.mark.parametrize("tree", [True, False])
def test_two_iterations_track_the_oracle(oracle, tree):
    n = 2000
    s, t = oracle.gate_edges(n)
    ref = oracle_state(oracle, n, s, t, tree=tree)
    with session(n, s, t, repulsion="tree" if tree else "dense") as sess:
        assert np.array_equal(sess.positions(), ref.positions())  # Mulberry32 starts
        for it in range(2):
            sess.step(1)
            ref.iterate()
            g, r = sess.positions(), ref.positions()
            ext = (r.max(0) - r.min(0)).max()
            assert np.abs(g - r).max() <= 1e-12 * ext, it
        sc = sess.scalars()
    rs = ref.scalars()
    for k, v in zip(("swing", "traction", "speed", "speed_efficiency"), rs[:4]):
        assert abs(sc[k] - v) <= 1e-12 * abs(v), k
    assert sc["iter_no"] == rs[4] == 2


def test_dense_stage_matches_networkx_vectors(oracle):
    """cli_fa2.rs:21-28 through gr_layout_run("layout.forceatlas2")."""
    worst = 0.0
    for c in nx_cases():
        n = c["n"]
        s, t = oracle.gate_edges(n)
        geo = G.ForceAtlas2.run(G.Topology(n, s, t), G.Fa2Params(max_iter=c["max_iter"]))
        ours = np.column_stack([geo.x, geo.y]).astype(np.float64)
        worst = max(worst, gap(ours, c))
    assert worst <= 1e-7, worst


def test_barnes_hut_stage_is_repeatable_and_finite():
    s, t = bh_graph(300)
    topo = G.Topology(300, s, t)
    a = G.ForceAtlas2BarnesHut.run(topo)
    b = G.ForceAtlas2BarnesHut.run(topo)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)
    assert np.isfinite(a.x).all() and np.isfinite(a.y).all()


def test_barnes_hut_stage_tracks_the_oracle_over_100_iterations(oracle):
    """Short graphs stay on the oracle's trajectory to within float32 narrowing."""
    s, t = line(20)
    geo = G.ForceAtlas2BarnesHut.run(G.Topology(20, s, t))
    ref = oracle_state(oracle, 20, s, t)
    ref.run(100)
    r = ref.positions().astype(np.float32)
    ext = (r.max(0) - r.min(0)).max()
    assert np.abs(np.column_stack([geo.x, geo.y]) - r).max() <= 1e-5 * ext


def test_early_exit_matches(oracle):
    s, t = line(2)
    with session(2, s, t) as sess:
        done = sess.step(100)
    ref = oracle_state(oracle, 2, s, t)
    assert done == ref.run(100)


def test_3d_dense_is_bit_exact(oracle):
    n = 500
    s, t = oracle.gate_edges(n)
    ref = oracle_state(oracle, n, s, t, dim=3, tree=False)
    ref.run(2)
    with session(n, s, t, dim=3, start=ref.positions(), repulsion="dense") as sess:
        assert np.array_equal(sess.forces(3), ref.forces(3))


def test_3d_octree_forces_are_bit_exact(oracle):
    n = 4000
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t, dim=3)
    ref.run(2)
    with session(n, s, t, dim=3, start=ref.positions()) as sess:
        assert np.array_equal(sess.forces(2), ref.forces(2))


# ------------------------------------------------------------------ precision


@pytest.mark.parametrize("precision", ["fp32", "fp32_kahan"])
def test_fp32_forces_within_1e5_of_fp64(oracle, precision):
    n = 30000
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t)
    ref.run(10)
    p = ref.positions().astype(np.float32).astype(np.float64)
    with session(n, s, t, start=p) as s64:
        f64 = s64.forces(7)
    with session(n, s, t, start=p, precision=precision) as s32:
        f32 = s32.forces(7)
    node, total = errors(f32, f64)
    print(precision, "per-node rms", node, "whole-field", total)
    assert total <= 1e-5 and node <= 1e-5, (node, total)


# ------------------------------------------------------------------ edges


def test_empty_and_single_node():
    e = np.zeros(0, np.uint32)
    geo = G.ForceAtlas2BarnesHut.run(G.Topology(0, e, e))
    assert geo.x.size == 0
    geo = G.ForceAtlas2BarnesHut.run(G.Topology(1, e, e))
    assert np.isfinite(geo.x).all() and np.isfinite(geo.y).all()


def test_all_points_coincident(oracle):
    n = 50
    s, t = line(n)
    start = np.tile([[0.5, 0.25]], (n, 1))
    ref = oracle_state(oracle, n, s, t, start=start)
    with session(n, s, t, start=start) as sess:
        assert np.array_equal(sess.forces(2), ref.forces(2))
        tr = sess.tree()
    assert tr["ncells"] == 1


def test_params_are_refused_not_clamped():
    import struct
    s, t = line(5)
    topo = G.Topology(5, s, t)
    with pytest.raises(G.ParamError):
        G.run_params("layout.forceatlas2.barnes_hut", topo, struct.pack("<5d", 0, 1, 2, 1, 0))
    with pytest.raises(G.StageError):
        G.run_params("layout.nope", topo, b"")
    geo = G.run_params("layout.forceatlas2.barnes_hut", topo, b"")
    assert geo.x.size == 5


def test_layout_forceatlas2_matches_the_reference_snapshot_pins(oracle):
    """The registry entry layout.forceatlas2 on the device, default parameters (100
    iterations), encoded as the reference's snapshot: the sha256 the reference itself pins
    for seeds 0..7 (harness/golden-2d-before.txt)."""
    from graph_render_amd import snapshot as S
    from tests.refgraphs import fa2_snapshot_pins, gate_model
    for seed, want in fa2_snapshot_pins().items():
        n, nodes, edges, s, t = gate_model(oracle, seed)
        print(f"snapshot pin seed={seed} n={n}", flush=True)
        geo = G.ForceAtlas2.run(G.Topology(n, s, t), G.Fa2Params())
        assert S.sha256(S.of_geometry(nodes, edges, s, t, geo)) == want, seed


@pytest.mark.parametrize("repulsion", ["tree", "dense", "pm"])
def test_graph_replay_equals_direct_launches(repulsion):
    # iterations replayed from the captured HIP graph, with host reads, forces and timing
    # in between, give the same bits as the kernels launched one by one
    n = 3000 if repulsion != "dense" else 600
    s, t = bh_graph(n)
    p = G.Fa2Params(max_iter=200)
    out = []
    for direct in (False, True):
        x = G.Fa2Ext(repulsion=repulsion, direct_launch=direct,
                     precision="fp32" if repulsion == "pm" else "fp64",
                     pm_grid=256 if repulsion == "pm" else 0)
        with G.Session(G.Topology(n, s, t), p, x) as sess:
            trace = []
            for chunk in (1, 7, 33, 2):
                sess.step(chunk)
                trace.append(sess.positions())
                trace.append(sess.forces(7))
            out.append(trace)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_live_parameter_change_tracks_the_oracle(oracle):
    # a session drives FA2 tick by tick and changes its parameters between ticks; the
    # device (through a replayed graph) follows the oracle doing the same
    n = 1500
    s, t = oracle.gate_edges(n)
    ref = oracle_state(oracle, n, s, t)
    with session(n, s, t, max_iter=1000) as sess:
        sess.step(3)
        for _ in range(3):
            ref.iterate()
        p2 = G.Fa2Params(scaling_ratio=5.0, gravity=0.25, jitter_tolerance=0.5, seed=7)
        sess.set_params(p2, theta=1.1)
        ref.set_params(jitter_tolerance=0.5, scaling_ratio=5.0, gravity=0.25, seed=7)
        ref.set_repulsion(True, 1.1 * 1.1)
        sess.step(3)
        for _ in range(3):
            ref.iterate()
        g, r = sess.positions(), ref.positions()
    ext = (r.max(0) - r.min(0)).max()
    assert np.abs(g - r).max() <= 1e-10 * ext


def test_set_params_refuses_out_of_range_values():
    s, t = bh_graph(100)
    with session(100, s, t) as sess:
        for bad in (dict(gravity=-1.0), dict(scaling_ratio=float("nan")),
                    dict(jitter_tolerance=2e6)):
            with pytest.raises(G.StageError) as e:
                sess.set_params(G.Fa2Params(**bad))
            assert e.value.name == next(iter(bad))
        with pytest.raises(G.StageError):
            sess.set_params(G.Fa2Params(dim=3))
        sess.step(2)  # still usable


# ------------------------------------------------------------------ controls and errors


def test_negative_control_one_ulp_of_scaling_ratio_goes_red(oracle):
    """The bit-exact comparisons above would catch a one-ulp change of a parameter: with
    the scaling ratio one ulp above 2 the repulsion differs from the oracle's (by about one
    ulp), and with 2 it is equal."""
    n = 3000
    s, t = bh_graph(n)
    ref = oracle_state(oracle, n, s, t)
    ref.run(3)
    p = ref.positions()
    r = ref.forces(2)
    with session(n, s, t, start=p) as good:
        assert np.array_equal(good.forces(2), r)
    with session(n, s, t, start=p, scaling_ratio=float(np.nextafter(2.0, 3.0))) as bad:
        f = bad.forces(2)
    assert not np.array_equal(f, r)
    assert np.abs(f - r).max() <= 1e-14 * np.abs(r).max()


def test_device_starts_are_mulberry32_at_any_seed(oracle):
    e = np.zeros(0, np.uint32)
    for seed, dim in ((7, 2), (123456789, 3), (0xFFFFFFFF, 2)):
        with G.Session(G.Topology(40, e, e), G.Fa2Params(seed=seed, dim=dim)) as s:
            assert np.array_equal(s.positions(), oracle.initial_positions(40, seed, dim))


def test_non_finite_run_is_refused_as_node_x_and_clears_the_output(oracle):
    """forceatlas2.rs:86-88: a run whose positions end non-finite is StageError::NonFinite
    {column: "node.x"}, and gr_layout_run leaves no stale output (the columns are zeroed)."""
    import ctypes as C
    from graph_render_amd import _abi as A
    n = 50
    s, t = line(n)
    start = np.random.default_rng(1).uniform(-1.0, 1.0, (n, 2))
    start[0, 0] = np.nan  # settle() reads node 0's coordinates (the reference's quirk)
    ref = oracle_state(oracle, n, s, t, start=start)
    ref.run(3)
    assert not np.isfinite(ref.positions()[0]).all()
    with pytest.raises(G.NonFiniteError) as e:
        G.ForceAtlas2BarnesHut.run(G.Topology(n, s, t), G.Fa2Params(max_iter=3),
                                   G.Fa2Ext(start=start))
    assert e.value.code == A.GR_ERR_NON_FINITE and e.value.name == "node.x"
    x = np.ones(n, np.float32)
    y = np.ones(n, np.float32)
    ext = G.Fa2Ext(start=start)._c(n, s.size, 2)
    code = A.lib().gr_layout_run(b"layout.forceatlas2.barnes_hut", n, s.ctypes.data,
                                 t.ctypes.data, s.size, None, 0, C.addressof(ext),
                                 x.ctypes.data, y.ctypes.data, None)
    assert code == A.GR_ERR_NON_FINITE and not x.any() and not y.any()


def test_layout_forceatlas2_3d_stage_tracks_the_oracle(oracle):
    """The registry id layout.forceatlas2.3d (run_3d, dense, dim 3) through gr_layout_run,
    against the oracle's 3D arm (pinned to the 2D arm in tests/test_oracle_pins.py)."""
    n = 400
    s, t = oracle.gate_edges(n)
    geo = G.run_3d(G.Topology(n, s, t), G.Fa2Params(max_iter=2))
    assert geo.dim == 3 and geo.z is not None
    ref = oracle_state(oracle, n, s, t, dim=3, tree=False)
    ref.run(2)
    r = ref.positions()
    ours = np.column_stack([geo.x, geo.y, geo.z]).astype(np.float64)
    ext = (r.max(0) - r.min(0)).max()
    assert np.abs(ours - r).max() <= 1e-7 * ext
    assert np.array_equal(ours, r.astype(np.float32).astype(np.float64))


def test_tree_export_refuses_a_short_buffer():
    import ctypes as C
    from graph_render_amd import _abi as A
    n = 500
    s, t = bh_graph(n)
    with session(n, s, t) as sess:
        sess.forces(2)
        nc = C.c_uint32(0)
        assert A.lib().gr_fa2_session_tree(sess._h, C.byref(nc), 0, None, None, None, None,
                                           None, None) == A.GR_OK
        short = np.zeros(max(nc.value - 1, 1), np.uint32)
        code = A.lib().gr_fa2_session_tree(sess._h, C.byref(nc), nc.value - 1, None,
                                           short.ctypes.data, None, None, None, None)
        assert code == A.GR_ERR_CAPACITY


@pytest.mark.parametrize("theta", [0.8, 1.2])
def test_the_walks_scan_is_the_union_of_its_lanes_walks(oracle, theta):
    # the wave-uniform scan visits each cell its 64 lanes' reference walks visit, once: its
    # step, acting-lane, accepted-cell and leaf-pair counts equal the oracle's replay of the
    # reference walk per tree position (FP64: every decision is the reference's)
    n = 60000
    s, t = bh_graph(n)
    print(f"walk counts n={n} theta={theta}", flush=True)
    with session(n, s, t, theta=theta) as sess:
        sess.step(3)
        pos = sess.positions()
        sess.forces(7)  # the tree of these positions (the last iteration's predates its step)
        got = sess.walk_stats()
    mass = oracle.Fa2(n, s, t).mass()
    want = oracle.walk_counts(pos, mass, theta * theta, 64)
    assert got == want
