"""The golden spiral's cosine and sine (libm 0.2.16, fdlibm via musl: oracle/libm_trig.c),
the gate model and the link and collide passes of the d3 simulation, against the reference's
own fixture fixtures/gpu/mesh-1k-{start,settled}.gmfx (copied to
tests/golden/pm_d3_mesh1k.npz by tests/golden/make_pm_d3_fixture.py): the spiral of 1000
nodes, the 128-point twiddle table of libm's cos/sin, the collapsed gate-model edges with
their strengths, and the velocity the link and collide passes add from rest at alpha 1."""
import math
import os
from fractions import Fraction

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


def test_the_spiral_is_the_references_bit_for_bit(oracle, mesh):
    x, y = oracle.golden_spiral(1000)
    assert same_bits(np.stack([x, y], 1), mesh["start_xy"])
    # the simulation's own start is that spiral
    e = np.zeros(0, np.uint32)
    sx, sy = oracle.d3sim(1000, e, e, np.zeros(0), ticks=0)
    assert same_bits(np.stack([sx, sy], 1), mesh["start_xy"])


def test_the_twiddles_are_libms_cos_and_sin(oracle, mesh):
    # fft.rs Plan::new: root k = (cos, sin)(-tau k / side); the forward table holds them all
    side = mesh["twiddle"].shape[0]
    root = {(oracle.libm_cos(-2 * math.pi * k / side), oracle.libm_sin(-2 * math.pi * k / side))
            for k in range(side // 2)}
    table = set(zip(mesh["twiddle"][:, 0].tolist(), mesh["twiddle"][:, 1].tolist()))
    assert table == root


def test_a_libm_other_than_fdlibms_misses_the_spiral(mesh):
    # negative control: the platform's own cos/sin (not fdlibm's) misses some of the spiral
    i = np.arange(1000, dtype=np.float64)
    x = np.array([math.cos(a) for a in i * 2.399963229728653]) * (12.0 * np.sqrt(i + 1.0))
    assert not same_bits(x, mesh["start_xy"][:, 0])


def _pi(bits=400):
    def atan_inv(x, b):
        one, total, term, k, sign = 1 << b, 0, (1 << b) // x, 1, 1
        while term:
            total += sign * (term // k)
            term //= x * x
            k += 2
            sign = -sign
        return total
    return Fraction(16 * atan_inv(5, bits) - 4 * atan_inv(239, bits), 1 << bits)


def test_the_reduction_constants_follow_from_pi():
    # rem_pio2: 2/pi rounded, pi/2 = p1 + p2 + p3 + tail with each piece 33 bits wide
    import struct
    pi = _pi()
    src = open(os.path.join(HERE, "..", "oracle", "libm_trig.c")).read()

    def const(name):
        line = next(ln for ln in src.splitlines() if f"{name} =" in ln)
        return float(line.split("=")[1].split(";")[0])

    def top33(f):
        e = 0
        while f >= 2:
            f, e = f / 2, e + 1
        while f < 1:
            f, e = f * 2, e - 1
        return Fraction(int(f * (1 << 32)), 1 << 32) * Fraction(2) ** e

    assert const("INVPIO2") == float(2 / pi)
    rest = pi / 2
    for k in (1, 2, 3):
        piece = top33(rest)
        assert Fraction(const(f"PIO2_{k}")) == piece, k
        assert const(f"PIO2_{k}T") == float(rest - piece), k
        rest -= piece
    assert struct.pack("<d", const("TOINT")) == struct.pack("<d", 1.5 * 2.0 ** 52)


def test_libm_agrees_with_the_exact_cosine_to_one_ulp(oracle):
    # fdlibm's cos/sin are within one ulp of the exact values; Taylor series over Fractions
    # with the argument reduced exactly give those values
    pi = _pi()
    rng = np.random.default_rng(5)
    for a in list(rng.uniform(-50, 50, 40)) + [i * 2.399963229728653 for i in (3, 77, 5000)]:
        r = Fraction(a) - round(Fraction(a) / (pi / 2)) * (pi / 2)
        q = round(Fraction(a) / (pi / 2)) % 4
        c, s, term = Fraction(0), Fraction(0), Fraction(1)
        for k in range(40):
            if k % 2 == 0:
                c += term * (-1) ** (k // 2)
            else:
                s += term * (-1) ** (k // 2)
            term = term * r / (k + 1)
        cos_exact = [c, -s, -c, s][q]
        sin_exact = [s, c, -s, -c][q]
        for got, want in ((oracle.libm_cos(a), cos_exact), (oracle.libm_sin(a), sin_exact)):
            assert abs(Fraction(got) - want) <= Fraction(math.ulp(float(want))), a


def _ipio2_tables():
    src = open(os.path.join(HERE, "..", "oracle", "ipio2.h")).read()
    body = src.split("IPIO2[66] = {")[1].split("};")[0]
    ipio2 = [int(v, 16) for v in body.replace("\n", " ").split(",") if v.strip()]
    body = src.split("PIO2_24[8] = {")[1].split("};")[0]
    pio2 = [float.fromhex(v.strip()) for v in body.split(",") if v.strip()]
    return ipio2, pio2


def test_the_large_reduction_tables_follow_from_pi():
    # 2/pi in 24-bit chunks and pi/2 in 24-bit pieces, against pi computed exactly, and their
    # leading entries against the values fdlibm publishes for __kernel_rem_pio2
    ipio2, pio2 = _ipio2_tables()
    assert len(ipio2) == 66 and len(pio2) == 8
    pi = _pi(1800)
    for k, v in enumerate(ipio2):
        assert v == int((2 / pi) * 2 ** (24 * (k + 1))) % 2 ** 24, k
    rest = pi / 2
    for k, v in enumerate(pio2):
        scale = Fraction(2) ** (23 + 24 * k)
        assert Fraction(v) == Fraction(int(rest * scale), 1) / scale, k
        rest -= Fraction(v)
    assert ipio2[:4] == [0xA2F983, 0x6E4E44, 0x1529FC, 0x2757D1]
    assert pio2[:4] == [1.57079625129699707031e+00, 7.54978941586159635335e-08,
                        5.39030252995776476554e-15, 3.28200341580791294123e-22]


def _exact_cos_sin(a, pi):
    """cos and sin of the double a as Fractions: a reduced exactly, then Taylor series."""
    q = round(Fraction(a) / (pi / 2))
    r = Fraction(a) - q * (pi / 2)
    c, s, term = Fraction(0), Fraction(0), Fraction(1)
    for k in range(40):
        if k % 2 == 0:
            c += term * (-1) ** (k // 2)
        else:
            s += term * (-1) ** (k // 2)
        term = term * r / (k + 1)
    return [c, -s, -c, s][q % 4], [s, c, -s, -c][q % 4]


def test_large_arguments_are_within_one_ulp_of_exact(oracle):
    # past 2^20 pi/2 (the golden spiral from node 686 302 on) the reduction is rem_pio2_large's;
    # fdlibm's cos/sin stay within one ulp of the exact values there too
    pi = _pi(600)
    rng = np.random.default_rng(11)
    nodes = [686_302, 686_303, 999_999, 4_000_000, 2 ** 31 + 5] + list(
        rng.integers(686_302, 1 << 32, 20))
    args = [float(i) * 2.399963229728653 for i in nodes]
    args += [2.0 ** 20 * math.pi / 2 * 1.0000001, -3.0e9, 1.0e15, 2.0 ** 60 + 2.0 ** 10]
    for a in args:
        cos_exact, sin_exact = _exact_cos_sin(a, pi)
        for got, want in ((oracle.libm_cos(a), cos_exact), (oracle.libm_sin(a), sin_exact)):
            assert math.isfinite(got)
            assert abs(Fraction(got) - want) <= Fraction(math.ulp(float(want))), a


def test_the_spiral_is_finite_past_the_medium_range(oracle):
    x, y = oracle.golden_spiral(686_400)
    assert np.isfinite(x).all() and np.isfinite(y).all()
    i = np.arange(686_290, 686_400, dtype=np.float64)
    r = 12.0 * np.sqrt(i + 1.0)
    # the radius is the reference's 12 sqrt(i + 1) to rounding, on either side of the switch
    assert np.allclose(np.hypot(x[686_290:], y[686_290:]), r, rtol=1e-14, atol=0)


def test_the_gate_model_collapses_to_the_references_edges(oracle, mesh):
    s, t, w = oracle.gate_model(1000)
    seen, lo, hi, st = set(), [], [], []
    for e in range(len(s)):
        a, b = int(s[e]), int(t[e])
        if a == b or (min(a, b), max(a, b)) in seen:
            continue
        seen.add((min(a, b), max(a, b)))
        lo.append(min(a, b))
        hi.append(max(a, b))
        st.append(w[e])
    assert np.array_equal(lo, mesh["edge_lo"]) and np.array_equal(hi, mesh["edge_hi"])
    assert same_bits(st, mesh["edge_strength"])


def test_the_link_pass_from_rest_is_the_references(oracle, mesh):
    # one tick with only the springs acting (no charge, no collision reach, no centring,
    # no velocity decay, alpha held at 1): the velocity is the link pass from rest
    s, t, w = oracle.gate_model(1000)
    x, y, vx, vy = oracle.d3sim(1000, s, t, w, ticks=1, velocities=True, charge=0.0,
                                collide_radius=0.0, center_strength=0.0, velocity_decay=1.0,
                                alpha_decay=0.0, initial_alpha=1.0)
    assert same_bits(np.stack([vx, vy], 1), mesh["start_delta_link"])


@pytest.mark.parametrize("state", ["start", "settled"])
def test_the_collide_pass_agrees_to_rounding(oracle, mesh, state):
    # the fixture's collide column comes from the mesh layout's hashed grid, which sums a
    # node's pushes in another order than the quadtree: each push is at most reach / 2 = 16
    # per axis, so a row of k pushes agrees to within 2 k eps 16 (reordered rounding)
    xy = mesh[f"{state}_xy"]
    got = oracle.d3_collide(xy, 16.0, seed=0, tick=0)
    want = mesh[f"{state}_delta_collide"]
    d2 = ((xy[:, None, :] - xy[None, :, :]) ** 2).sum(-1)
    k = (d2 < 32.0 ** 2).sum(1, keepdims=True) - 1
    assert (np.abs(got - want) <= 2 * k * np.finfo(float).eps * 16.0).all()
    assert ((got == 0) == (want == 0)).all()  # the same nodes collide
