"""CPU pins for the reverse-mode checks: the torch restatement of dense FA2 (tests/torch_fa2.py,
the FP64 reference gradient) against the oracle's forward, its autograd gradient against
central differences of itself, and Kruskal stress-1 against the reference's own unit tests
(crates/graph-cli/src/bench/tests/plan.rs:69-92)."""
import numpy as np
import pytest
import torch

from tests.fa2_autograd import stress1, stress1_np, stress_sources
from tests import torch_fa2


def graph(n, seed):
    rng = np.random.default_rng(seed)
    m = 3 * n
    s = rng.integers(0, n, m).astype(np.uint32)
    t = rng.integers(0, n, m).astype(np.uint32)
    return s, t


def test_torch_restatement_follows_the_oracle(oracle):
    n, iters = 60, 15
    s, t = graph(n, 1)
    rng = np.random.default_rng(2)
    w_raw = rng.uniform(0.5, 1.5, s.size)
    start = rng.uniform(-50, 50, (n, 2))
    ref = oracle.Fa2(n, s, t, start=start, weights=w_raw, tree=False)
    for _ in range(iters):
        ref.iterate()
    lo, hi, raw = torch_fa2.simple_edges(s, t)
    p = torch_fa2.run(torch.tensor(start), lo, hi, torch.tensor(w_raw[raw]),
                      torch.tensor(2.0, dtype=torch.float64),
                      torch.tensor(1.0, dtype=torch.float64), iters)
    want = ref.positions()
    extent = np.ptp(want, axis=0).max()
    assert np.abs(p.numpy() - want).max() / extent < 1e-11


def test_torch_gradient_matches_its_central_differences():
    n, iters = 40, 12
    s, t = graph(n, 3)
    lo, hi, _ = torch_fa2.simple_edges(s, t)
    rng = np.random.default_rng(4)
    start = torch.tensor(rng.uniform(-30, 30, (n, 2)))
    c = torch.tensor(rng.normal(size=(n, 2)))
    w = torch.tensor(rng.uniform(0.5, 1.5, lo.size), requires_grad=True)
    k = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    (torch_fa2.run(start, lo, hi, w, k, g, iters) * c).sum().backward()

    def loss(w_, k_, g_):
        with torch.no_grad():
            return float((torch_fa2.run(start, lo, hi, w_, k_, g_, iters) * c).sum())

    h = 1e-6
    for e in (0, 5, 11):
        wp, wm = w.detach().clone(), w.detach().clone()
        wp[e] += h
        wm[e] -= h
        fd = (loss(wp, k.detach(), g.detach()) - loss(wm, k.detach(), g.detach())) / (2 * h)
        assert abs(fd - float(w.grad[e])) <= 1e-5 * max(abs(fd), 1e-3)
    for which, par in (("k", k), ("g", g)):
        hp = [k.detach().clone(), g.detach().clone()]
        hm = [k.detach().clone(), g.detach().clone()]
        j = 0 if which == "k" else 1
        hp[j] = hp[j] + h
        hm[j] = hm[j] - h
        fd = (loss(w.detach(), *hp) - loss(w.detach(), *hm)) / (2 * h)
        assert abs(fd - float(par.grad)) <= 1e-5 * max(abs(fd), 1e-3)


@pytest.mark.parametrize("xs, ys, src, dst, want", [
    ([0.0, 2.0, 4.0], [0.0] * 3, [0, 1], [1, 2], 0.0),             # isometric path
    ([0.0, 1.0, 0.0], [0.0] * 3, [0, 1], [1, 2], np.sqrt(8 / 12)),  # folded back
    ([5.0] * 3, [5.0] * 3, [0, 1], [1, 2], 1.0),                    # collapsed
    ([0.0, 1.0], [0.0, 0.0], [], [], 1.0),                          # edgeless
    ([0.0, 1.0, 100.0, 101.0], [0.0] * 4, [0, 2], [1, 3], 0.0),     # two components
])
def test_stress_matches_the_reference_cases(xs, ys, src, dst, want):
    from tests.fa2_autograd import bfs_hops
    n = len(xs)
    pos = np.stack([xs, ys], 1).astype(np.float64)
    srcs = stress_sources(n)
    hops = bfs_hops(n, np.array(src, np.int64), np.array(dst, np.int64), srcs)
    assert stress1_np(pos, srcs, hops) == want
    assert float(stress1(torch.tensor(pos), srcs, hops)) == pytest.approx(want, abs=1e-15)


def test_stress_sources_are_evenly_spaced():
    assert stress_sources(10000).tolist()[:3] == [0, 39, 78]
    assert stress_sources(100).tolist() == list(range(100))
