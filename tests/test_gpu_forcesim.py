"""layout.forceatlas2.forcesim on the device (gr_forcesim_run / gr_forcesim_batch) against the
reference's measured bit vectors (crates/graph-core/src/layout/force/forcesim/tests.rs) and
the oracle (oracle/forcesim_oracle.c), through the C ABI.

Bar: bit for bit. The layout is f32 state with f64 coefficient rows and numpy's reduction
orders; every f32 coordinate must carry the oracle's bits (the oracle itself carries the
reference's measured bits on its fixtures)."""
import numpy as np
import pytest

import graph_render_amd as G
from graph_render_amd import forcesim as FS
from tests.test_forcesim_oracle import DIAMOND, GATE04_RAW, TREE

pytestmark = pytest.mark.gpu


def topo(n, pairs):
    s = np.array([a for a, _ in pairs], np.uint32)
    t = np.array([b for _, b in pairs], np.uint32)
    return G.Topology(n, s, t)


def xyz_of(geo):
    return np.stack([geo.x, geo.y, geo.z], 1)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_the_reference_fixtures_bit_for_bit():
    print("forcesim fixtures", flush=True)
    d = FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND))
    assert bits(d.x).tolist() == [0x4082FC14, 0xBD9BF065, 0x3D9BF04D, 0xC082FC14]
    assert bits(d.y).tolist() == [0xBFD4B204, 0x40A00000, 0xC0A00000, 0x3FD4B204]
    assert bits(d.z).tolist() == [0x405BC46D, 0x4020A903, 0xC020A903, 0xC05BC46E]
    t = FS.ForceAtlas2ForceSim.run(topo(15, TREE))
    assert bits(t.x).tolist() == [0xBF31B6C9, 0xC00AF28C, 0xC03CE39C, 0xC0025F77, 0xBEDCAE74,
                                  0xBB4704E5, 0xBF2B75BB, 0x3F897283, 0xBE18BCE2, 0xBFAF023C,
                                  0xBDC676AF, 0x402E25C4, 0x40688EA7, 0x40371CAC, 0x3E8FF4C8]
    g = G.find("layout.forceatlas2.forcesim").run(topo(6, GATE04_RAW))
    assert bits(g.x).tolist() == [0xBBC1C7FB, 0x3E18749E, 0xBF4E0B13, 0xBF00F81B, 0x3F33F82B,
                                  0x3EECE2D6]
    assert g.dim == 3


def random_graph(rng, n, deg):
    m = int(n * deg / 2)
    s = rng.integers(0, max(n, 1), m).astype(np.uint32)
    t = rng.integers(0, max(n, 1), m).astype(np.uint32)
    return s, t


CASES = [(2, 1.0), (3, 2.0), (17, 3.0), (50, 4.0), (129, 3.0), (200, 6.0), (257, 2.0),
         (400, 8.0), (1000, 5.0), (2000, 3.0)]


@pytest.mark.parametrize("n,deg", CASES)
def test_the_device_is_the_oracle_bit_for_bit(oracle, n, deg):
    rng = np.random.default_rng(n)
    s, t = random_graph(rng, n, deg)
    p = FS.Fa2ForceSimParams(iterations=30 if n > 500 else 50, seed=int(rng.integers(0, 2**40)),
                             scaling_ratio=float(rng.uniform(0.5, 4)),
                             gravity=float(rng.uniform(0.2, 3)),
                             jitter_tolerance=float(rng.uniform(0.5, 2)), scale=7.5)
    print(f"forcesim n={n} deg={deg} iters={p.iterations}", flush=True)
    got = xyz_of(FS.ForceAtlas2ForceSim.run(G.Topology(n, s, t), p))
    want, st, _ = oracle.forcesim(n, s, t, iterations=p.iterations, scale=p.scale,
                                  scaling_ratio=p.scaling_ratio, gravity=p.gravity,
                                  jitter_tolerance=p.jitter_tolerance, seed=p.seed)
    assert st == 0
    bad = np.nonzero(bits(got) != bits(want))
    assert len(bad[0]) == 0, (len(bad[0]), got[bad][:4], want[bad][:4])


def test_a_batch_is_its_graphs_run_alone(oracle):
    rng = np.random.default_rng(77)
    topos, params = [], []
    for b in range(24):
        n = int(rng.integers(0, 300))
        s, t = random_graph(rng, n, float(rng.uniform(1, 6)))
        topos.append(G.Topology(n, s, t))
        params.append(FS.Fa2ForceSimParams(iterations=int(rng.integers(0, 60)), seed=b * 7919))
    print("forcesim batch of 24", flush=True)
    got = FS.batch(topos, params)
    for (xyz, st), tp, p in zip(got, topos, params):
        want, wst, _ = oracle.forcesim(tp.n, tp.source, tp.target, iterations=p.iterations,
                                       seed=p.seed)
        assert st == wst == 0
        assert np.array_equal(bits(xyz), bits(want))


def test_parallel_edges_self_loops_seed_and_iteration_floor():
    plain = xyz_of(FS.ForceAtlas2ForceSim.run(topo(3, [(0, 1), (1, 2), (2, 0)])))
    noisy = xyz_of(FS.ForceAtlas2ForceSim.run(
        topo(3, [(0, 1), (1, 0), (1, 2), (2, 2), (2, 0), (1, 2)])))
    assert np.array_equal(bits(plain), bits(noisy))
    d = xyz_of(FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND)))
    moved = xyz_of(FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND),
                                              FS.Fa2ForceSimParams(seed=1_767_573_728)))
    assert (d != moved).all()
    zero = xyz_of(FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND), FS.Fa2ForceSimParams(iterations=0)))
    one = xyz_of(FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND), FS.Fa2ForceSimParams(iterations=1)))
    assert np.array_equal(bits(zero), bits(one)) and not np.array_equal(bits(zero), bits(d))


def test_graphs_too_small_to_step():
    for n in (0, 1):
        g = FS.ForceAtlas2ForceSim.run(topo(n, []))
        assert g.x.shape == g.y.shape == g.z.shape == (n,)
        assert (g.x == 0).all()


def test_a_non_finite_answer_is_refused():
    with pytest.raises(G.NonFiniteError) as e:
        FS.ForceAtlas2ForceSim.run(topo(4, DIAMOND), FS.Fa2ForceSimParams(scale=float("inf")))
    assert e.value.name == "node.x"
    out = FS.batch([topo(4, DIAMOND), topo(4, DIAMOND)],
                   [FS.Fa2ForceSimParams(scale=float("inf")), FS.Fa2ForceSimParams()])
    assert [st for _, st in out] == [2, 0]
