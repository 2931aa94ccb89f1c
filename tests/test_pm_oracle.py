"""Pins the particle-mesh oracle (oracle/pm_oracle.c), the FA2-law PM arm's checker.

The reference's PM fixtures cover d3's law only; this arm is pinned by the mathematics: the
FFT against numpy's, the frame's rung rules, the mesh field against the exact pair sum for
well-separated clusters, and the vanishing self-force of a lone node."""
import numpy as np
import pytest


@pytest.mark.parametrize("side", [8, 64])
def test_fft2_is_numpys(oracle, side):
    rng = np.random.default_rng(side)
    z = rng.normal(size=(side, side)) + 1j * rng.normal(size=(side, side))
    assert np.allclose(oracle.fft2(z, -1), np.fft.fft2(z), rtol=1e-12, atol=1e-9)
    assert np.allclose(oracle.fft2(z, +1), np.fft.ifft2(z) * side * side, rtol=1e-12, atol=1e-9)


def test_rung_ladder_values(oracle):
    assert oracle.lib().gro_pm_h(0) == 1.0 and oracle.lib().gro_pm_h(4) == 2.0
    assert oracle.lib().gro_pm_h(-8) == 0.25
    for s in range(-40, 40):
        assert abs(oracle.lib().gro_pm_h(s) - 2.0 ** (s / 4)) <= 1e-15 * 2.0 ** (s / 4)


def test_frame_fits_and_holds_every_node(oracle):
    rng = np.random.default_rng(1)
    xy = rng.uniform(-100, 100, (500, 2))
    _, fr = oracle.pm_forces(xy, np.ones(500), 2.0, 256)
    assert fr is not None and fr["cells"] + fr["reach"] <= 256
    u = (xy - [fr["ox"], fr["oy"]]) / fr["h"]
    assert (u >= 0).all() and (u <= fr["cells"] - 1).all()
    # the smallest rung that fits: one rung finer does not
    h2 = oracle.lib().gro_pm_h(fr["step"] - 1)
    span = (xy.max(0) - xy.min(0)).max()
    c2 = np.floor(span / h2) + 3
    assert c2 > 256 or 2 * c2 - 1 > 256


def test_a_lone_node_feels_no_force(oracle):
    f, fr = oracle.pm_forces(np.array([[0.3, 0.7], [50.0, 50.0]]), np.array([1.0, 0.0]), 2.0, 64)
    assert np.abs(f[0]).max() < 1e-12


def exact(xy, m, k):
    d = xy[:, None, :] - xy[None, :, :]
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    return k * m[:, None] * ((m[None, :, None] * d) / r2[..., None]).sum(1)


def test_mesh_field_follows_the_pair_sum_at_range(oracle):
    """Two clusters far apart: the cross-cluster force dominates and the mesh resolves it."""
    rng = np.random.default_rng(2)
    a = rng.normal(0, 1.0, (200, 2))
    b = rng.normal(0, 1.0, (200, 2)) + [60.0, 0.0]
    xy = np.concatenate([a, b])
    m = rng.integers(1, 5, 400).astype(float)
    f, _ = oracle.pm_forces(xy, m, 2.0, 512)
    # compare the cluster-mean force (intra-cluster terms cancel in the mean)
    ex = exact(xy, m, 2.0)
    for sl in (slice(0, 200), slice(200, 400)):
        fm, em = f[sl].sum(0), ex[sl].sum(0)
        assert np.linalg.norm(fm - em) <= 2e-2 * np.linalg.norm(em)


def test_mesh_is_linear_in_mass(oracle):
    rng = np.random.default_rng(3)
    xy = rng.uniform(0, 10, (300, 2))
    m = rng.uniform(1, 3, 300)
    f1, _ = oracle.pm_forces(xy, m, 1.0, 128)
    f2, _ = oracle.pm_forces(xy, m, 2.0, 128)
    assert np.allclose(f2, 2 * f1, rtol=1e-12, atol=1e-12)


def _d3_fixture():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "pm_d3_mesh1k.npz")
    return np.load(path)


@pytest.mark.parametrize("state", ["start", "settled"])
def test_d3_law_charge_pass_reproduces_the_reference_fixture(oracle, state):
    """fixtures/gpu/mesh-1k-*.gmfx: the reference's own frame and charge-pass deltas (f64)
    for n = 1000 on a 128 mesh, at layout.force's defaults and alpha 1."""
    f = _d3_fixture()
    xy, want, fr = f[f"{state}_xy"], f[f"{state}_delta_charge"], f[f"{state}_frame"]
    got, frame = oracle.pm_charge(xy, int(fr[6]), -90.0, 1.0, 1.0, 520.0)
    assert frame["step"] == int(fr[0]) and frame["h"] == fr[1]
    assert (frame["ox"], frame["oy"]) == (fr[2], fr[3])
    assert frame["cells"] == int(fr[4]) and frame["reach"] == int(fr[5])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
