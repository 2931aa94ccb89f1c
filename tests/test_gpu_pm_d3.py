"""layout.force.particle_mesh on the device (gr_pm_layout_run, gr_pm_d3_charge,
gr_pm_d3_collide) against the reference's own fixtures/gpu/mesh-1k-{start,settled}.gmfx and
against the oracle (oracle/pm_d3_oracle.c), through the C ABI.

Bar: bit for bit. The mesh passes are FP64 in the reference's operations and orders: the
deposit summed per cell in ascending grid slot, fft.rs's radix-2 butterflies on libm's
twiddles, the hashed grid's read order."""
import os

import numpy as np
import pytest

from graph_render_amd import barnes_hut, fa2
from graph_render_amd.fa2 import Fa2Ext, Topology

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def mesh():
    return np.load(os.path.join(HERE, "golden", "pm_d3_mesh1k.npz"))


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    bad = np.nonzero(np.ascontiguousarray(a).view(np.uint64) !=
                     np.ascontiguousarray(b).view(np.uint64))
    return None if len(bad[0]) == 0 else (bad[0][0], a[bad][0], b[bad][0], len(bad[0]))


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_charge_pass_is_the_references_bit_for_bit(mesh, state):
    xy, fr = mesh[f"{state}_xy"], mesh[f"{state}_frame"]
    print(f"pm d3 charge fixture {state}", flush=True)
    got, frame = barnes_hut.pm_charge_pass(xy)
    assert (frame["step"], frame["h"], frame["ox"], frame["oy"], frame["cells"],
            frame["reach"], frame["side"]) == (int(fr[0]), fr[1], fr[2], fr[3], int(fr[4]),
                                               int(fr[5]), int(fr[6]))
    want = mesh[f"{state}_delta_charge"]
    assert same_bits(got, want), first_diff(got, want)


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_collide_pass_is_the_references_bit_for_bit(mesh, state):
    xy = mesh[f"{state}_xy"]
    print(f"pm d3 collide fixture {state}", flush=True)
    got, _ = barnes_hut.pm_collide_pass(xy, 16.0)
    want = mesh[f"{state}_delta_collide"]
    assert same_bits(got, want), first_diff(got, want)


POINTS = {
    "normal": lambda r: r.normal(0.0, 300.0, (3000, 2)),
    "clusters": lambda r: np.concatenate([r.normal(0, 5.0, (2000, 2)),
                                          r.normal(900, 40.0, (2000, 2))]),
    "lattice": lambda r: r.integers(0, 80, (4000, 2)).astype(np.float64) * 9.0,
    "wide": lambda r: r.normal(0.0, 5000.0, (30000, 2)),
    "nonfinite": lambda r: np.concatenate([r.normal(0, 200.0, (1000, 2)),
                                           [[np.nan, 1.0], [np.inf, 2.0], [3.0, -np.inf]]]),
}


@pytest.mark.parametrize("case", sorted(POINTS))
def test_the_charge_pass_is_the_oracle_bit_for_bit(oracle, case):
    xy = POINTS[case](np.random.default_rng(len(case)))
    print(f"pm d3 charge case={case} n={len(xy)}", flush=True)
    got, frame = barnes_hut.pm_charge_pass(xy, charge=-90.0, alpha=0.37)
    want, oframe, _ = oracle.pmd3_charge(xy, oracle.pm_side_for(len(xy)), charge=-90.0,
                                         alpha=0.37)
    assert (frame is None) == (oframe is None)
    if frame is not None:
        assert frame == oframe
    assert same_bits(got, want), first_diff(got, want)


@pytest.mark.parametrize("case", sorted(POINTS))
def test_the_collide_pass_is_the_oracle_bit_for_bit(oracle, case):
    xy = POINTS[case](np.random.default_rng(10 + len(case)))
    print(f"pm d3 collide case={case} n={len(xy)}", flush=True)
    got, order = barnes_hut.pm_collide_pass(xy, 16.0, seed=7, tick=3)
    want, oorder = oracle.pmd3_collide(xy, 16.0, seed=7, tick=3)
    assert np.array_equal(order, oorder)
    assert same_bits(got, want), first_diff(got, want)


def pm_device(n, s, t, w=None, **kw):
    return barnes_hut.ParticleMesh.run(Topology(n, s, t), w, **kw)


def same_f32(g, x, y):
    return (np.array_equal(g.x.view(np.uint32), np.asarray(x, np.float32).view(np.uint32)) and
            np.array_equal(g.y.view(np.uint32), np.asarray(y, np.float32).view(np.uint32)))


@pytest.mark.parametrize("n", [66, 700, 3000])
def test_the_mesh_layout_is_the_oracle_bit_for_bit(oracle, n):
    s, t, w = oracle.gate_model(n)
    print(f"pm layout gate model n={n}", flush=True)
    g = pm_device(n, s, t, w)
    x, y = oracle.d3sim(n, s, t, w, mesh_side=oracle.pm_side_for(n))
    assert same_f32(g, x, y)


def test_a_large_mesh_layout_over_twenty_ticks(oracle):
    n = 20000
    s, t, w = oracle.gate_model(n)
    print(f"pm layout gate model n={n} ticks=20", flush=True)
    with barnes_hut.ForceSession(Topology(n, s, t), w, particle_mesh=True) as S:
        S.step(20)
        dev = S.state()
    want = oracle.d3sim(n, s, t, w, ticks=20, mesh_side=oracle.pm_side_for(n), velocities=True)
    for a, b in zip(dev, want):
        assert same_bits(a, b), first_diff(a, b)


def test_the_mesh_layout_at_a_million_nodes(oracle):
    # the reference's own timed graph (the gate model at 1M nodes, 1 549 929 raw edges, a
    # 1024 x 1024 mesh), its spiral past the medium reduction included: four ticks
    n = 1_000_000
    s, t, w = oracle.gate_model(n)
    assert len(s) == 1_549_929
    print(f"pm layout gate model n={n} ticks=4", flush=True)
    with barnes_hut.ForceSession(Topology(n, s, t), w, particle_mesh=True) as S:
        S.step(4)
        dev = S.state()
    want = oracle.d3sim(n, s, t, w, ticks=4, mesh_side=oracle.pm_side_for(n), velocities=True)
    for a, b in zip(dev, want):
        assert np.isfinite(b).all()
        assert same_bits(a, b), first_diff(a, b)


def test_the_registry_id_and_the_session_agree(oracle):
    s, t, w = oracle.gate_model(500)
    top = Topology(500, s, t)
    got = fa2.find("layout.force.particle_mesh").run(top, None, Fa2Ext(edge_weights=w))
    want = pm_device(500, s, t, w)
    assert np.array_equal(got.x, want.x) and np.array_equal(got.y, want.y)
    with barnes_hut.ForceSession(top, w, particle_mesh=True) as S:
        for _ in range(4):
            S.step(28)
        x, y, _, _ = S.state()
    assert np.array_equal(x.astype(np.float32), want.x)
    assert np.array_equal(y.astype(np.float32), want.y)


def test_without_charge_or_collision_the_mesh_layout_is_the_tree_layout(oracle):
    s, t, w = oracle.gate_model(800)
    p = barnes_hut.ForceParams(charge_strength=0.0, collide_radius=0.0)
    a = pm_device(800, s, t, w, params=p, ticks=40)
    b = barnes_hut.BarnesHut.run(Topology(800, s, t), w, params=p, ticks=40)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.y, b.y)


def test_timing_the_mesh_passes_leaves_the_state(oracle):
    # the centring sums and the hashed grid's pushes timed again on a stepped mesh session:
    # the state is the last tick's, and the next ticks are an untimed session's
    s, t, w = oracle.gate_model(3000)
    top = Topology(3000, s, t)
    with barnes_hut.ForceSession(top, w, particle_mesh=True) as S:
        S.step(5)
        before = S.state()
        mc, mg, cells, buckets = S.time_walks(3)
        assert mc > 0 and mg > 0
        assert cells == oracle.pm_side_for(3000) ** 2 and buckets == 8192  # next_pow2(2n)
        for a, b in zip(before, S.state()):
            assert same_bits(a, b)
        S.step(3)
        after = S.state()
    with barnes_hut.ForceSession(top, w, particle_mesh=True) as S:
        S.step(8)
        for a, b in zip(after, S.state()):
            assert same_bits(a, b), first_diff(a, b)
