"""GPU particle-mesh repulsion (FP32 mesh) against the FP64 PM oracle, through the C ABI.

Bar: the device solves the same frame, deposit, kernel and read as the oracle, in FP32 with
an exact fixed-point deposit, so the forces agree to FP32 FFT accuracy: whole-field relative
error <= 1e-4 and per-node rms <= 1e-3 (both metrics of barnes_hut/tests.rs:34-47)."""
import numpy as np
import pytest

import graph_render_amd as G
from tests.refgraphs import bh_graph, errors

pytestmark = pytest.mark.gpu


def pm_session(n, s, t, start, side, precision="fp32"):
    return G.Session(G.Topology(n, s, t), G.Fa2Params(),
                     G.Fa2Ext(repulsion="pm", pm_grid=side, start=start, precision=precision))


@pytest.mark.parametrize("side,n,spread", [(128, 300, 10.0), (256, 3000, 50.0),
                                          (1024, 20000, 200.0), (2048, 60000, 500.0)])
def test_pm_matches_the_oracle(oracle, side, n, spread):
    rng = np.random.default_rng(side)
    xy = np.concatenate([rng.normal(0, spread / 4, (n // 2, 2)),
                         rng.uniform(-spread, spread, (n - n // 2, 2))])
    s, t = bh_graph(n)
    print(f"pm case side={side} n={n}", flush=True)
    with pm_session(n, s, t, xy.astype(np.float32).astype(np.float64), side) as sess:
        f = sess.forces(2)
        _, _, mass = sess.graph()
    ref, frame = oracle.pm_forces(xy.astype(np.float32).astype(np.float64), mass, 2.0, side)
    assert frame is not None
    node, total = errors(f, ref)
    print("pm vs oracle", side, n, node, total)
    assert total <= 1e-4 and node <= 1e-3, (node, total)


def test_pm_far_field_follows_the_pair_sum():
    rng = np.random.default_rng(2)
    a = rng.normal(0, 1.0, (200, 2))
    b = rng.normal(0, 1.0, (200, 2)) + [60.0, 0.0]
    xy = np.concatenate([a, b])
    e = np.zeros(0, np.uint32)
    with pm_session(400, e, e, xy, 512) as sess:
        f = sess.forces(2)
    with G.Session(G.Topology(400, e, e), G.Fa2Params(),
                   G.Fa2Ext(repulsion="dense", start=xy)) as dense:
        ex = dense.forces(2)
    for sl in (slice(0, 200), slice(200, 400)):
        assert np.linalg.norm(f[sl].sum(0) - ex[sl].sum(0)) <= 2e-2 * np.linalg.norm(ex[sl].sum(0))


def test_pm_layout_runs_and_stays_finite():
    n = 5000
    s, t = bh_graph(n)
    with pm_session(n, s, t, None, 512) as sess:
        assert sess.step(30) == 30
        p = sess.positions()
    assert np.isfinite(p).all()


@pytest.mark.parametrize("state", ["start", "settled"])
def test_d3_charge_pass_matches_the_reference_fixture(state):
    """The device mesh under the reference's d3 law against its own fixture (f64): frame
    bit for bit, charge deltas within the FP32 transform's error."""
    import os
    from graph_render_amd import pm
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "pm_d3_mesh1k.npz"))
    xy, want, fr = f[f"{state}_xy"], f[f"{state}_delta_charge"], f[f"{state}_frame"]
    assert pm.side_for(xy.shape[0]) == int(fr[6]) == 128
    got, frame = pm.charge_pass(xy)
    print(f"d3 charge pass {state}: frame {frame}", flush=True)
    assert frame["step"] == int(fr[0]) and frame["h"] == fr[1]
    assert (frame["ox"], frame["oy"]) == (fr[2], fr[3])
    assert frame["cells"] == int(fr[4]) and frame["reach"] == int(fr[5])
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"  max error / max |delta| = {err:.2e}", flush=True)
    assert err <= 1e-5


def test_pm_skips_non_finite_points_and_repeats_bit_for_bit():
    """Points at NaN / inf take no part in the deposit: every finite point's force equals,
    bit for bit, its force in a session without them (no edges, so every mass is 1, and the
    fixed-point scale, set by the total mass, is the same for 3000 and 3003 points); theirs is
    NaN. The deposit is an exact integer sum, so a second session repeats the forces exactly."""
    rng = np.random.default_rng(7)
    n = 3000
    xy = np.concatenate([rng.normal(0, 40.0, (n // 2, 2)),
                         rng.uniform(-150, 150, (n - n // 2, 2))])
    bad = np.array([[np.nan, 1.0], [np.inf, 0.0], [2.0, -np.inf]])
    at = np.array([5, 1700, 2999])
    mixed = np.insert(xy, at - np.arange(3), bad, axis=0)
    finite = np.isfinite(mixed).all(axis=1)
    assert finite.sum() == n and (~finite).sum() == 3
    e = np.zeros(0, np.uint32)
    print("pm non-finite case n=3003 side=1024", flush=True)
    with pm_session(n + 3, e, e, mixed, 1024) as sess:
        f = sess.forces(2)
    with pm_session(n + 3, e, e, mixed, 1024) as sess:
        f2 = sess.forces(2)
    with pm_session(n, e, e, mixed[finite], 1024) as sess:
        ref = sess.forces(2)
    assert np.array_equal(f, f2, equal_nan=True)
    assert np.array_equal(f[finite], ref)
    assert np.isnan(f[~finite]).all()
