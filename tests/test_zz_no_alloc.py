"""gr_pm_charge_noalloc (no device allocation; every buffer in the library's own array)
against gr_pm_charge through the C ABI: the same charge deltas and frame, bit for bit, on
two calls in a row, the second smaller, and again after it; sizes past the array's limits
are refused before any device work."""
import os

import numpy as np
import pytest

import graph_render_amd as G
from graph_render_amd import pm

pytestmark = pytest.mark.gpu


def points(n, spread, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.normal(0, spread / 4, (n // 2, 2)),
                           rng.uniform(-spread, spread, (n - n // 2, 2))])


def both(xy, side=0, **law):
    a, fa = pm.charge_pass(xy, side, **law)
    b, fb = pm.charge_pass(xy, side, allocate=False, **law)
    return a, fa, b, fb


def test_noalloc_equals_the_allocating_pass_on_calls_in_a_row():
    big = points(60000, 400.0, 1)
    small = points(3000, 80.0, 2)
    cases = [(big, 512, {}), (small, 0, dict(charge=-30.0, alpha=0.5, distance_max=90.0)),
             (big, 512, {})]
    first = None
    for k, (xy, side, law) in enumerate(cases):
        print(f"noalloc call {k}: n={xy.shape[0]} side={side} law={law}", flush=True)
        a, fa, b, fb = both(xy, side, **law)
        assert fa is not None and fa == fb, (fa, fb)
        assert np.array_equal(a, b), np.abs(a - b).max()
        if k == 0:
            first = b
    # the third call repeats the first: nothing of the smaller call in between leaks in
    assert np.array_equal(cases[2][0], cases[0][0])
    assert np.array_equal(b, first)


def test_noalloc_reproduces_the_reference_fixture_frame_and_deltas():
    f = np.load(os.path.join(os.path.dirname(__file__), "golden", "pm_d3_mesh1k.npz"))
    for state in ("start", "settled"):
        xy = f[f"{state}_xy"]
        print(f"noalloc fixture {state}", flush=True)
        a, fa, b, fb = both(xy)
        assert fa == fb and np.array_equal(a, b)


def test_noalloc_largest_sizes_and_refusals():
    n = pm.MAX_NODES_NOALLOC
    xy = points(n, 3000.0, 3)
    print(f"noalloc largest: n={n} side={pm.MAX_SIDE_NOALLOC}", flush=True)
    a, fa, b, fb = both(xy, pm.MAX_SIDE_NOALLOC)
    assert fa == fb and np.array_equal(a, b)
    with pytest.raises(G.CapacityError):
        pm.charge_pass(points(n + 1, 10.0, 4), 0, allocate=False)
    with pytest.raises(G.CapacityError):
        pm.charge_pass(points(100, 10.0, 5), 2 * pm.MAX_SIDE_NOALLOC, allocate=False)
    with pytest.raises(G.ParamError):
        pm.charge_pass(points(100, 10.0, 6), 300, allocate=False)
