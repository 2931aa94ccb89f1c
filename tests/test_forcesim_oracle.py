"""Pins of the oracle's layout.forceatlas2.forcesim (oracle/forcesim_oracle.c).

Every value below is one the reference holds as bits, measured against SciGraphs / numpy
2.3.3 (crates/graph-core/src/layout/force/forcesim/tests.rs, forcesim/seed.rs,
forcesim/reduce.rs, rng/pcg64/tests.rs), or numpy's own default_rng, importable here."""
import numpy as np
import pytest

SEED = 1_767_573_729


def f32_bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32).tolist()


def test_pcg64_is_numpys_default_rng(oracle):
    # rng/pcg64/tests.rs:10-48
    draws = [0x3FE96B8FA5E53C4A, 0x3FE5525D058739CE, 0x3FE3A71AE513D623,
             0x3FC1C7A8FC6458F4, 0x3FEB5006A7EEBDE4, 0x3FE9C989DD74267F]
    zero = [0x3FE461FD79FB3850, 0x3FD1442F7E20B674, 0x3FA4FA7B529D9BD0,
            0x3F90EC9ED84D0BC0, 0x3FEA064F4F059BCA, 0x3FED354B2F34C803]
    raw = [0xCB5C7D2F29E25674, 0xAA92E82C39CE7606, 0x9D38D7289EB11871, 0x238F51F8C8B1EC2D]
    assert oracle.pcg64_doubles(SEED, 6).view(np.uint64).tolist() == draws
    assert oracle.pcg64_doubles(0, 6).view(np.uint64).tolist() == zero
    assert oracle.pcg64_raw(SEED, 4).tolist() == raw
    assert oracle.pcg64_state(SEED) == (60223290111031304053904000740221457340,
                                        232180640892196755669630849829386553819)
    # negative control: a neighbouring seed moves every draw
    assert not np.array_equal(oracle.pcg64_doubles(SEED - 1, 6).view(np.uint64), draws)


@pytest.mark.parametrize("seed", [0, 1, 7, 12345, SEED, 2**32 - 1, 2**32, 2**40 + 3,
                                  2**63 + 99])
def test_pcg64_follows_numpy_at_any_seed(oracle, seed):
    rng = np.random.default_rng(seed)
    st = rng.bit_generator.state["state"]
    assert oracle.pcg64_state(seed) == (st["state"], st["inc"])
    assert np.array_equal(oracle.pcg64_doubles(seed, 300), rng.random(300))


def test_the_first_start_row_is_the_measured_one(oracle):
    # forcesim/seed.rs:35-44: random_positions(77, 5.0, seed)[0] as f32 bits
    d = oracle.pcg64_doubles(SEED, 3)
    row = ((d - 0.5) * 5.0).astype(np.float32)
    assert f32_bits(row) == [0x3FBC6739, 0x3F54DE89, 0x3F121C34]


MEASURED = [0x41EBD930, 0x410CED93, 0x40BFB904, 0x418EFD17, 0x41E85794, 0x40D3C252,
            0x40C6F1B9, 0x416ED4ED, 0x424AABFB, 0x41A2A046, 0x4112F274, 0x41D39C1C,
            0x40F6930C, 0x40DDC3EB, 0x41869969]


def test_pairwise_f32_is_numpys_answer_on_the_measured_fifteen(oracle):
    # reduce.rs:163-203, with its negative control: the remainder summed on its own is one
    # ulp higher
    v = np.array(MEASURED, np.uint32).view(np.float32)
    got = oracle.pairwise_f32(v)
    assert f32_bits([got]) == [0x43808671]
    lanes = v[:8]
    tree = np.float32((lanes[0] + lanes[1]) + (lanes[2] + lanes[3])) + \
        np.float32((lanes[4] + lanes[5]) + (lanes[6] + lanes[7]))
    grouped = np.float32(0.0)
    for x in v[8:]:
        grouped = np.float32(grouped + x)
    assert f32_bits([np.float32(tree + grouped)])[0] == f32_bits([got])[0] + 1
    assert got == np.add.reduce(v)  # numpy's own pairwise sum of a contiguous f32 array


@pytest.mark.parametrize("n", [1, 7, 8, 9, 77, 128, 129, 130, 200, 1000, 4097])
def test_pairwise_sums_match_numpy(oracle, n):
    rng = np.random.default_rng(n)
    v32 = (rng.random(n) * 1e4).astype(np.float32)
    v64 = rng.random(n) * 1e4
    assert oracle.pairwise_f32(v32) == np.add.reduce(v32)
    assert oracle.pairwise_f64(v64) == np.add.reduce(v64)


def test_dot3_is_the_f32_association(oracle):
    # reduce.rs:229-243
    row = np.array([0x3C9169F9, 0x3AE0ADC0, 0xBC0D7A4B], np.uint32).view(np.float32)
    assert f32_bits([oracle.dot3_f32(row, row)]) == [0x39CDD4CA]
    assert np.einsum("ij,ij->i", row[None], row[None])[0] == oracle.dot3_f32(row, row)


def exact_cbrt_is(y: float, x: int) -> bool:
    """y is the double nearest the cube root of x: x lies between the cubes of the midpoints
    to y's neighbours (exact rational arithmetic)."""
    from fractions import Fraction as F
    lo, hi = float(np.nextafter(y, 0.0)), float(np.nextafter(y, np.inf))
    return ((F(lo) + F(y)) / 2) ** 3 <= x <= ((F(y) + F(hi)) / 2) ** 3


def test_the_cube_root_is_correctly_rounded(oracle):
    # k = 5 / cbrt(n) takes libm's correctly rounded cube root; the C library's is not
    # (cbrt(27) = 3.0000000000000004 in glibc), so the oracle corrects it by exact cubes
    rng = np.random.default_rng(9)
    ns = list(range(1, 3000)) + [27, 64, 125, 1000, 2**30, 2**32 - 1, 2**32] + \
        [int(v) for v in rng.integers(1, 2**32, 500)]
    bad = [n for n in ns if not exact_cbrt_is(oracle.cbrt_cr(n), n)]
    assert not bad, bad[:10]
    assert oracle.cbrt_cr(27) == 3.0 and oracle.cbrt_cr(15) == 2.4662120743304703


def test_the_constants_follow_the_reference_arithmetic(oracle):
    # state.rs:117-132: k from the correctly rounded cube root; the mean mass is numpy's
    # pairwise f32 mean; the powers are products (k * k, mm * mm * mm) as the port writes them
    for n, mass in [(4, [3, 3, 3, 3]), (15, [2] * 8 + [3] * 6 + [4]), (77, list(range(1, 78)))]:
        m = np.array(mass, np.float32)
        k = 5.0 / max(oracle.cbrt_cr(max(n, 1)), 1.0)
        mm = float(m.mean())
        rep = 2.0 * (k * k / max(mm * mm * mm, 1e-9))
        grv = 1.0 * 0.1 * (0.38 * 2.0 * (5.0 * 5.0) / (max(k, 1e-9) * max(mm * mm, 1e-9)))
        assert oracle.forcesim_constants(m) == (k, rep, grv)


DIAMOND = [(0, 1), (0, 2), (1, 3), (2, 3)]
TREE = [(14, 7), (11, 7), (7, 8), (14, 0), (0, 1), (4, 0), (1, 3), (1, 2), (4, 5), (6, 4),
        (8, 9), (8, 10), (3, 9), (11, 12), (13, 11)]
GATE04_RAW = [(1, 0), (1, 0), (2, 1), (2, 0), (3, 0), (4, 0), (4, 0), (5, 1), (5, 1)]


def run(oracle, n, pairs, **kw):
    s = np.array([a for a, _ in pairs], np.uint32)
    t = np.array([b for _, b in pairs], np.uint32)
    xyz, st, _ = oracle.forcesim(n, s, t, **kw)
    assert st == 0
    return xyz


def test_the_diamond_layout_is_the_measured_one(oracle):
    # forcesim/tests.rs:62-79
    xyz = run(oracle, 4, DIAMOND)
    assert f32_bits(xyz[:, 0]) == [0x4082FC14, 0xBD9BF065, 0x3D9BF04D, 0xC082FC14]
    assert f32_bits(xyz[:, 1]) == [0xBFD4B204, 0x40A00000, 0xC0A00000, 0x3FD4B204]
    assert f32_bits(xyz[:, 2]) == [0x405BC46D, 0x4020A903, 0xC020A903, 0xC05BC46E]


def test_the_tree_layout_x_column_is_the_measured_one(oracle):
    # forcesim/tests.rs:81-123
    xyz = run(oracle, 15, TREE)
    assert f32_bits(xyz[:, 0]) == [0xBF31B6C9, 0xC00AF28C, 0xC03CE39C, 0xC0025F77, 0xBEDCAE74,
                                   0xBB4704E5, 0xBF2B75BB, 0x3F897283, 0xBE18BCE2, 0xBFAF023C,
                                   0xBDC676AF, 0x402E25C4, 0x40688EA7, 0x40371CAC, 0x3E8FF4C8]


def test_the_gate_layout_x_column_is_the_measured_one(oracle):
    # forcesim/tests.rs:125-157
    xyz = run(oracle, 6, GATE04_RAW)
    assert f32_bits(xyz[:, 0]) == [0xBBC1C7FB, 0x3E18749E, 0xBF4E0B13, 0xBF00F81B, 0x3F33F82B,
                                   0x3EECE2D6]


def test_parallel_edges_and_self_loops_do_not_change_the_layout(oracle):
    plain = run(oracle, 3, [(0, 1), (1, 2), (2, 0)])
    noisy = run(oracle, 3, [(0, 1), (1, 0), (1, 2), (2, 2), (2, 0), (1, 2)])
    assert np.array_equal(plain.view(np.uint32), noisy.view(np.uint32))


def test_a_neighbouring_seed_moves_every_coordinate(oracle):
    a = run(oracle, 4, DIAMOND)
    b = run(oracle, 4, DIAMOND, seed=SEED - 1)
    assert (a != b).all()


def test_zero_iterations_still_run_one_step(oracle):
    zero = run(oracle, 4, DIAMOND, iterations=0)
    one = run(oracle, 4, DIAMOND, iterations=1)
    assert np.array_equal(zero, one)
    assert not np.array_equal(zero, run(oracle, 4, DIAMOND))


def test_too_small_to_step(oracle):
    assert run(oracle, 0, []).shape == (0, 3)
    one = run(oracle, 1, [])
    assert one.shape == (1, 3) and (one == 0.0).all()


def test_the_output_scale_resizes_without_rearranging(oracle):
    wide = run(oracle, 4, DIAMOND, scale=50.0)
    narrow = run(oracle, 4, DIAMOND)
    assert np.allclose(wide[:, 0] / narrow[:, 0], 10.0, rtol=1e-4)
