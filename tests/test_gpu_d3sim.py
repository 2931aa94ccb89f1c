"""The frozen layout.force.barnes_hut on the device (gr_d3_layout_run: golden-spiral start,
ticks of link, many-body over the quadtree, center, collide over the quadtree of the
projected positions, integration) against the reference's own 65 golden digests
(crates/graph-core/src/layout/force/session/tests/golden.rs) and against the oracle's
restatement, through the C ABI.

Bar: bit for bit. Every pass is FP64 in the reference's operations and orders (the centre's
sums in node order, each node's link shares in ascending edge order, the walks' cells in
preorder), so the f32 columns must carry the oracle's narrowed bits exactly."""
import numpy as np
import pytest

from graph_render_amd import barnes_hut, fa2
from graph_render_amd.fa2 import Fa2Ext, StageError, Topology
from tests.test_d3sim_oracle import digest, golden

pytestmark = pytest.mark.gpu


def device(n, s, t, w=None, **kw):
    return barnes_hut.BarnesHut.run(Topology(n, s, t), w, **kw)


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def same_f32(g, x, y):
    return (np.array_equal(g.x.view(np.uint32), np.asarray(x, np.float32).view(np.uint32)) and
            np.array_equal(g.y.view(np.uint32), np.asarray(y, np.float32).view(np.uint32)))


def test_the_device_reproduces_every_golden_digest(oracle):
    rows = golden()
    assert len(rows) == 65
    bad = []
    for seed, want in rows:
        n = 2 + seed % 600
        s, t, w = oracle.gate_model(n)
        g = device(n, s, t, w)
        if digest(g.x, g.y) != want:
            bad.append(seed)
    print("d3sim golden digests, misses:", bad, flush=True)
    assert not bad, bad


@pytest.mark.parametrize("n", [700, 5000])
def test_gate_model_is_the_oracle_bit_for_bit(oracle, n):
    s, t, w = oracle.gate_model(n)
    print(f"d3sim gate model n={n} m={len(s)}", flush=True)
    g = device(n, s, t, w)
    x, y = oracle.d3sim(n, s, t, w)
    assert same_f32(g, x, y)


def test_a_large_gate_model_over_twenty_ticks(oracle):
    n = 30000
    s, t, w = oracle.gate_model(n)
    print(f"d3sim gate model n={n} ticks=20", flush=True)
    g = device(n, s, t, w, ticks=20)
    x, y = oracle.d3sim(n, s, t, w, ticks=20)
    assert same_f32(g, x, y)


def ragged_graph(rng, n, m):
    # repeated pairs, both orientations and self-loops; the first raw edge of a pair keeps
    # its strength
    s = rng.integers(0, n, m).astype(np.uint32)
    t = rng.integers(0, n, m).astype(np.uint32)
    t[: m // 10] = s[: m // 10]
    s[m // 2: m // 2 + m // 8], t[m // 2: m // 2 + m // 8] = t[: m // 8], s[: m // 8]
    w = rng.choice([0.1, 0.4, 1.0, 2.5, 7.0], m)
    return s, t, w


@pytest.mark.parametrize("case", range(6))
def test_ragged_graphs_and_parameters(oracle, case):
    rng = np.random.default_rng(100 + case)
    n = int(rng.integers(3, 1500))
    s, t, w = ragged_graph(rng, n, int(rng.integers(0, 3 * n)))
    change = [dict(), dict(theta=0.0), dict(collide_radius=60.0), dict(charge=-400.0, theta=1.4),
              dict(center_strength=0.3, link_distance=10.0), dict(distance_max=40.0)][case]
    ticks = [112, 40, 112, 60, 112, 112][case]
    print(f"d3sim ragged case={case} n={n} m={len(s)} {change} ticks={ticks}", flush=True)
    p = barnes_hut.ForceParams()
    for k, v in change.items():
        setattr(p, "charge_strength" if k == "charge" else k, v)
    g = device(n, s, t, w, params=p, ticks=ticks)
    x, y = oracle.d3sim(n, s, t, w, ticks=ticks, **change)
    assert same_f32(g, x, y)


def test_edge_free_graphs_and_single_nodes(oracle):
    for n in (1, 2, 3, 64, 65, 500):
        e = np.zeros(0, np.uint32)
        print(f"d3sim edge-free n={n}", flush=True)
        g = device(n, e, e)
        x, y = oracle.d3sim(n, e, e, np.zeros(0))
        assert same_f32(g, x, y), n
    # unit strengths when none are given
    s, t, _ = oracle.gate_model(300)
    g = device(300, s, t)
    x, y = oracle.d3sim(300, s, t, np.ones(len(s)))
    assert same_f32(g, x, y)


def test_the_registry_id_reads_the_strength_column(oracle):
    # registry/layouts.rs:92-97: no published parameters; the strength column as edge weights
    s, t, w = oracle.gate_model(400)
    top = Topology(400, s, t)
    got = fa2.find("layout.force.barnes_hut").run(top, None, Fa2Ext(edge_weights=w))
    want = device(400, s, t, w)
    assert np.array_equal(got.x, want.x) and np.array_equal(got.y, want.y)
    with pytest.raises(StageError):
        fa2.run_params("layout.force.barnes_hut", top, b"\0" * 8)


def test_the_layout_is_deterministic(oracle):
    s, t, w = oracle.gate_model(20000)
    a = device(20000, s, t, w, ticks=30)
    b = device(20000, s, t, w, ticks=30)
    assert np.array_equal(a.x.view(np.uint32), b.x.view(np.uint32))
    assert np.array_equal(a.y.view(np.uint32), b.y.view(np.uint32))


@pytest.fixture(scope="module")
def mesh():
    import os
    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                "pm_d3_mesh1k.npz"))


def test_the_session_starts_on_the_references_spiral(oracle, mesh):
    # fixtures/gpu/mesh-1k-start.gmfx: the golden spiral of libm's cos and sin, every bit
    s, t, w = oracle.gate_model(1000)
    with barnes_hut.ForceSession(Topology(1000, s, t), w) as S:
        x, y, vx, vy = S.state()
    assert same_bits(np.stack([x, y], 1), mesh["start_xy"])
    assert not vx.any() and not vy.any()


def test_the_spiral_past_the_medium_reduction_is_the_oracles(oracle):
    # from node 686 302 on, the spiral's angles pass 2^20 pi/2: the device reduces them
    # exactly over 256 bits of 2/pi and pi/2, the oracle by libm's rem_pio2_large (24-bit
    # chunks of 2/pi); both must give libm's bits for every node up to a million
    n = 1_000_000
    e = np.zeros(0, np.uint32)
    print(f"spiral n={n}", flush=True)
    with barnes_hut.ForceSession(Topology(n, e, e), None) as S:
        x, y, _, _ = S.state()
    wx, wy = oracle.golden_spiral(n)
    assert np.isfinite(wx).all() and np.isfinite(wy).all()
    assert same_bits(x, wx) and same_bits(y, wy)


def test_the_link_pass_from_rest_is_the_references(oracle, mesh):
    # one tick with only the springs acting at alpha 1 (no charge, no collision reach, no
    # centring, no decay): the velocities are the fixture's link column, bit for bit
    s, t, w = oracle.gate_model(1000)
    p = barnes_hut.ForceParams(charge_strength=0.0, collide_radius=0.0, center_strength=0.0,
                               velocity_decay=1.0, alpha_decay=0.0)
    with barnes_hut.ForceSession(Topology(1000, s, t), w, p) as S:
        S.step(1)
        _, _, vx, vy = S.state()
    assert same_bits(np.stack([vx, vy], 1), mesh["start_delta_link"])


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_collide_pass_is_the_oracle_bit_for_bit(oracle, mesh, state):
    xy = mesh[f"{state}_xy"]
    print(f"d3 collide fixture state={state}", flush=True)
    got = barnes_hut.collide_pass(xy, 16.0, seed=0, tick=0)
    assert same_bits(got, oracle.d3_collide(xy, 16.0, seed=0, tick=0))


@pytest.mark.parametrize("case", ["lattice", "coincident", "line", "wide"])
def test_the_collide_pass_on_jiggled_and_coincident_points(oracle, case):
    rng = np.random.default_rng(len(case))
    xy = {"lattice": lambda: rng.integers(0, 60, (3000, 2)).astype(np.float64) * 8.0,
          "coincident": lambda: np.repeat(rng.normal(0, 40, (300, 2)), 5, axis=0),
          "line": lambda: np.stack([rng.uniform(0, 4000, 2000), np.zeros(2000)], 1),
          "wide": lambda: rng.normal(0, 3000, (20000, 2))}[case]()
    print(f"d3 collide case={case} n={len(xy)}", flush=True)
    got = barnes_hut.collide_pass(xy, 16.0, seed=7, tick=3)
    assert same_bits(got, oracle.d3_collide(xy, 16.0, seed=7, tick=3))


def test_step_counts_give_the_same_bytes(oracle):
    # session.rs step: step(12), 12 x step(1) and 3 x step(4) are the same bytes
    s, t, w = oracle.gate_model(3000)
    top = Topology(3000, s, t)
    out = []
    for chunks in ([12], [1] * 12, [4, 4, 4]):
        with barnes_hut.ForceSession(top, w) as S:
            for c in chunks:
                alpha, settled, _ = S.step(c)
            out.append(S.state())
    for a, b in zip(out[0], out[1]):
        assert same_bits(a, b)
    for a, b in zip(out[0], out[2]):
        assert same_bits(a, b)
    assert alpha == pytest.approx(0.94 ** 12) and not settled


def test_timing_the_walks_leaves_the_state(oracle):
    s, t, w = oracle.gate_model(2000)
    with barnes_hut.ForceSession(Topology(2000, s, t), w) as S:
        with pytest.raises(StageError):
            S.time_walks(2)  # nothing to time before the first tick
        S.step(5)
        before = S.state()
        mc, mk, cc, ck = S.time_walks(3)
        assert mc > 0 and mk > 0 and cc >= 1 and ck >= 1
        for a, b in zip(before, S.state()):
            assert same_bits(a, b)
        S.step(3)  # and the next ticks are the ones an untimed session takes
        after = S.state()
    with barnes_hut.ForceSession(Topology(2000, s, t), w) as S:
        S.step(8)
        for a, b in zip(after, S.state()):
            assert same_bits(a, b)


def test_a_diverging_layout_is_refused_as_non_finite(oracle):
    # a charge near the f64 range overflows the velocities: planar_points refuses the
    # non-finite coordinates (StageError::NonFinite{node.x}) and the columns come back zero
    s, t, w = oracle.gate_model(50)
    p = barnes_hut.ForceParams(charge_strength=-1e307)
    x, y = oracle.d3sim(50, s, t, w, ticks=6, charge=-1e307)
    assert not (np.isfinite(x).all() and np.isfinite(y).all())
    print("d3sim non-finite n=50", flush=True)
    with pytest.raises(StageError) as e:
        device(50, s, t, w, params=p, ticks=6)
    assert e.value.code == 2
