"""The oracle's layout.force.particle_mesh passes (oracle/pm_d3_oracle.c: the reference's
radix-2 transform on libm's twiddles, the slot-ordered CIC deposit, the kernel spectrum, the
hashed-grid collide) against the reference's own fixtures/gpu/mesh-1k-{start,settled}.gmfx
(copied to tests/golden/pm_d3_mesh1k.npz by tests/golden/make_pm_d3_fixture.py): the kernel
spectrum's digest, the charge and collide deltas from a fresh mesh, bit for bit."""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def mesh():
    return np.load(os.path.join(HERE, "golden", "pm_d3_mesh1k.npz"))


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_charge_pass_is_the_references_bit_for_bit(oracle, mesh, state):
    xy, fr = mesh[f"{state}_xy"], mesh[f"{state}_frame"]
    side = oracle.pm_side_for(len(xy))
    assert side == int(fr[6]) == 128
    got, frame, spec = oracle.pmd3_charge(xy, side)
    assert (frame["step"], frame["h"], frame["ox"], frame["oy"], frame["cells"],
            frame["reach"]) == (int(fr[0]), fr[1], fr[2], fr[3], int(fr[4]), int(fr[5]))
    assert same_bits(got, mesh[f"{state}_delta_charge"])
    digest = hashlib.sha256(spec[0].astype("<f8").tobytes() +
                            spec[1].astype("<f8").tobytes()).hexdigest()
    assert digest == str(mesh[f"{state}_spectrum_sha256"])


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_hashed_grid_collide_is_the_references_bit_for_bit(oracle, mesh, state):
    xy = mesh[f"{state}_xy"]
    got, order = oracle.pmd3_collide(xy, 16.0, seed=0, tick=0)
    assert same_bits(got, mesh[f"{state}_delta_collide"])
    assert sorted(order.tolist()) == list(range(len(xy)))


def test_negative_controls_miss(oracle, mesh):
    # another deposit order rounds some cell sums differently (the settled state's cells
    # hold enough nodes for it; the spiral's hardly any); a radius an ulp off moves the
    # collide pushes; a cutoff off moves the spectrum
    xy = mesh["settled_xy"]
    n = len(xy)
    got, _, _ = oracle.pmd3_charge(xy, 128, order=np.arange(n)[::-1])
    assert not same_bits(got, mesh["settled_delta_charge"])
    assert np.abs(got - mesh["settled_delta_charge"]).max() < 1e-9  # rounding only
    xy = mesh["start_xy"]
    got, _ = oracle.pmd3_collide(xy, np.nextafter(16.0, 17.0))
    assert not same_bits(got, mesh["start_delta_collide"])
    got, _, spec = oracle.pmd3_charge(xy, 128, distance_max=519.0)
    digest = hashlib.sha256(spec[0].tobytes() + spec[1].tobytes()).hexdigest()
    assert digest != str(mesh["start_spectrum_sha256"])


def test_two_charges_feel_the_law(oracle):
    # two nodes on a grid line: the transform pair convolves the unit deposits with G, so
    # each node feels the other's sampled law (a wrong sign, twiddle or transpose fails)
    xy = np.array([[0.0, 0.0], [130.0, 0.0]])
    got, frame, _ = oracle.pmd3_charge(xy, 128, charge=1.0, alpha=1.0, distance_min=1.0,
                                       distance_max=520.0)
    # G(r) = -r / |r|^2 with r = (node - source): at a positive charge node 0 moves toward
    # node 1 by 1/130 (a negative charge, the layout's, repels), to the mesh's accuracy
    assert got[0, 0] == pytest.approx(1.0 / 130.0, rel=1e-3)
    assert abs(got[0, 1]) < 1e-9
    assert got[1, 0] == pytest.approx(-1.0 / 130.0, rel=1e-3)


def test_the_mesh_layout_reduces_to_the_tree_layout_without_charge_or_collision(oracle):
    # with no many-body strength and no collision reach both layouts are the springs and
    # the centring alone, so their positions agree exactly
    s, t, w = oracle.gate_model(400)
    kw = dict(ticks=40, charge=0.0, collide_radius=0.0)
    bx, by = oracle.d3sim(400, s, t, w, **kw)
    mx, my = oracle.d3sim(400, s, t, w, mesh_side=128, **kw)
    assert np.array_equal(bx, mx) and np.array_equal(by, my)


def test_the_mesh_layout_settles(oracle):
    s, t, w = oracle.gate_model(1000)
    x, y = oracle.d3sim(1000, s, t, w, mesh_side=oracle.pm_side_for(1000))
    assert np.isfinite(x).all() and np.isfinite(y).all()
    bx, by = oracle.d3sim(1000, s, t, w)
    # not the tree layout's bytes, but the same scale of picture
    span = lambda a, b: max(np.ptp(a), np.ptp(b))
    assert 0.5 < span(x, y) / span(bx, by) < 2.0
