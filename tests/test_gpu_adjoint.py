"""Reverse mode through the C ABI (gr_fa2_session_record / gr_fa2_session_adjoint).

Checks, each against something the adjoint does not share code with:
- dense FP64: d loss / d (raw edge weights, scaling ratio, gravity, start positions) equal
  torch.autograd through the torch restatement (tests/torch_fa2.py, itself pinned to the
  oracle's forward on the CPU) to 1e-8 relative;
- tree at theta 0 sums exactly the dense pairs, so its adjoint equals the dense one;
- tree at theta 0.8 (cell centres moving with their bodies, the part no dense check sees)
  against central differences of the device's own FP64 forward, h small enough that no
  accept/open decision flips;
- the checkpoint interval changes nothing; the session is left as the recording left it;
- FP32 forward + adjoint close to FP64 over a short horizon; the torch.autograd.Function
  wrapper returns what the ABI returns.
"""
import numpy as np
import pytest
import torch

import graph_render_amd as G
from tests import fa2_autograd as AG
from tests import torch_fa2
from tests.refgraphs import bh_graph

pytestmark = pytest.mark.gpu


def rand_graph(n, seed, deg=3):
    rng = np.random.default_rng(seed)
    m = deg * n
    return (rng.integers(0, n, m).astype(np.uint32), rng.integers(0, n, m).astype(np.uint32))


def device(topo, w_raw, k, g, iters, every, cvec, start, precision="fp64",
           repulsion="dense", theta=None):
    p = G.Fa2Params(max_iter=iters, scaling_ratio=k, gravity=g)
    x = G.Fa2Ext(precision=precision, repulsion=repulsion, theta=theta, start=start,
                 edge_weights=w_raw)
    with G.Session(topo, p, x) as s:
        done = s.record(iters, every)
        assert done == iters
        pos = s.positions()
        grads = s.adjoint(cvec)
        after = s.positions()
    assert np.array_equal(after, pos)
    return pos, grads


def relerr(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def small():
    n, iters = 150, 25
    s, t = rand_graph(n, 5)
    rng = np.random.default_rng(6)
    w_raw = rng.uniform(0.5, 1.5, s.size)
    start = rng.uniform(-60, 60, (n, 2))
    cvec = rng.normal(size=(n, 2))
    return n, iters, G.Topology(n, s, t), s, t, w_raw, start, cvec


def test_dense_adjoint_equals_torch_autograd(small):
    n, iters, topo, s, t, w_raw, start, cvec = small
    print(f"dense adjoint n={n} iters={iters}", flush=True)
    pos, gd = device(topo, w_raw, 2.0, 1.0, iters, 7, cvec, start)
    lo, hi, raw = torch_fa2.simple_edges(s, t)
    w = torch.tensor(w_raw[raw], requires_grad=True)
    k = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    p0 = torch.tensor(start, requires_grad=True)
    pt = torch_fa2.run(p0, lo, hi, w, k, g, iters)
    (pt * torch.tensor(cvec)).sum().backward()
    assert np.abs(pt.detach().numpy() - pos).max() / np.ptp(pos, 0).max() < 1e-10
    want_w = np.zeros(s.size)
    want_w[raw] = w.grad.numpy()
    assert relerr(gd["weights"], want_w) < 1e-8
    assert abs(gd["scaling_ratio"] - float(k.grad)) <= 1e-8 * abs(float(k.grad))
    assert abs(gd["gravity"] - float(g.grad)) <= 1e-8 * abs(float(g.grad))
    assert relerr(gd["start"], p0.grad.numpy()) < 1e-8
    # repeated and self-loop raw edges carry no gradient
    dropped = np.setdiff1d(np.arange(s.size), raw)
    assert np.all(gd["weights"][dropped] == 0.0)


def test_tree_theta0_adjoint_equals_dense(small):
    n, iters, topo, s, t, w_raw, start, cvec = small
    _, gd = device(topo, w_raw, 2.0, 1.0, iters, 7, cvec, start)
    _, gt = device(topo, w_raw, 2.0, 1.0, iters, 7, cvec, start, repulsion="tree", theta=0.0)
    for key in ("weights", "start"):
        assert relerr(gt[key], gd[key]) < 1e-9, key
    for key in ("scaling_ratio", "gravity"):
        assert abs(gt[key] - gd[key]) <= 1e-9 * abs(gd[key]), key


def test_checkpoint_interval_does_not_change_the_gradient(small):
    n, iters, topo, s, t, w_raw, start, cvec = small
    ref = device(topo, w_raw, 2.0, 1.0, iters, 1, cvec, start)[1]
    # an interval past the run is one segment: the tape is sized by the longest segment
    # recorded, not by the interval asked for (1 << 30 would not fit any device)
    for every in (4, 25, 1000, 1 << 30):
        got = device(topo, w_raw, 2.0, 1.0, iters, every, cvec, start)[1]
        for key in ("weights", "start"):
            assert np.array_equal(got[key], ref[key]), (every, key)
        assert got["scaling_ratio"] == ref["scaling_ratio"]
        assert got["gravity"] == ref["gravity"]


def test_tree_adjoint_matches_central_differences():
    n, iters = 600, 12
    s, t = bh_graph(n)
    topo = G.Topology(n, s, t)
    rng = np.random.default_rng(8)
    w_raw = rng.uniform(0.5, 1.5, s.size)
    start = rng.uniform(-80, 80, (n, 2))
    cvec = rng.normal(size=(n, 2))
    print(f"tree adjoint vs FD n={n} iters={iters}", flush=True)
    _, ga = device(topo, w_raw, 2.0, 1.0, iters, 5, cvec, start, repulsion="tree", theta=0.8)

    def loss(w, k=2.0, g=1.0):
        p = G.Fa2Params(max_iter=iters, scaling_ratio=k, gravity=g)
        x = G.Fa2Ext(precision="fp64", repulsion="tree", theta=0.8, start=start,
                     edge_weights=w)
        with G.Session(topo, p, x) as sess:
            sess.step(iters)
            return float((sess.positions() * cvec).sum())

    h = 1e-6
    scale = np.abs(ga["weights"]).max()
    for e in rng.choice(np.nonzero(ga["weights"])[0], 6, replace=False):
        wp, wm = w_raw.copy(), w_raw.copy()
        wp[e] += h
        wm[e] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - ga["weights"][e]) <= 1e-4 * max(abs(fd), 1e-2 * scale), (e, fd)
    for key, kw in (("scaling_ratio", "k"), ("gravity", "g")):
        base = {"k": 2.0, "g": 1.0}
        up, dn = dict(base), dict(base)
        up[kw] += h
        dn[kw] -= h
        fd = (loss(w_raw, **up) - loss(w_raw, **dn)) / (2 * h)
        assert abs(fd - ga[key]) <= 1e-4 * abs(fd), (key, fd, ga[key])


def test_session_continues_after_the_adjoint(small):
    n, iters, topo, s, t, w_raw, start, cvec = small
    p = G.Fa2Params(max_iter=iters + 5)
    x = G.Fa2Ext(precision="fp64", repulsion="tree", start=start, edge_weights=w_raw)
    with G.Session(topo, p, x) as a:
        a.record(iters, 6)
        a.adjoint(cvec)
        a.step(5)
        pa, sa = a.positions(), a.scalars()
    with G.Session(topo, p, x) as b:
        b.step(iters + 5)
        pb, sb = b.positions(), b.scalars()
    assert np.array_equal(pa, pb)
    assert sa == sb


def test_fp32_adjoint_is_close_to_fp64(small):
    n, _, topo, s, t, w_raw, start, cvec = small
    iters = 8
    _, g64 = device(topo, w_raw, 2.0, 1.0, iters, 4, cvec, start, repulsion="tree")
    _, g32 = device(topo, w_raw, 2.0, 1.0, iters, 4, cvec, start, precision="fp32",
                    repulsion="tree")
    print("fp32 vs fp64 adjoint:", relerr(g32["weights"], g64["weights"]),
          relerr(g32["start"], g64["start"]), flush=True)
    assert relerr(g32["weights"], g64["weights"]) < 1e-3
    assert relerr(g32["start"], g64["start"]) < 1e-3
    assert abs(g32["gravity"] - g64["gravity"]) <= 1e-3 * abs(g64["gravity"])


def test_autograd_function_with_stress_loss():
    n = 400
    s, t = bh_graph(n)
    topo = G.Topology(n, s, t)
    rng = np.random.default_rng(9)
    start = rng.uniform(-50, 50, (n, 2))
    spec = AG.UnrolledSpec(topology=topo, iters=15, checkpoint_every=4, precision="fp64",
                           repulsion="tree", start=start)
    srcs = AG.stress_sources(n, 64)
    hops = AG.bfs_hops(n, s, t, srcs)
    fn = AG.make_function()
    w = torch.ones(s.size, dtype=torch.float64, requires_grad=True)
    k = torch.tensor(2.0, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    loss = AG.stress1(fn.apply(w, k, g, spec), srcs, hops)
    loss.backward()
    # the same gradient straight from the ABI, with d loss / d positions from torch
    sess, pos = AG.run_forward(spec, np.ones(s.size), 2.0, 1.0, record=True)
    pt = torch.tensor(pos, requires_grad=True)
    AG.stress1(pt, srcs, hops).backward()
    direct = sess.adjoint(pt.grad.numpy())
    sess.close()
    # the tree's cell accumulators add in arrival order: equal to rounding, not bit for bit
    np.testing.assert_allclose(w.grad.numpy(), direct["weights"], rtol=1e-12,
                               atol=1e-12 * np.abs(direct["weights"]).max())
    assert float(k.grad) == pytest.approx(direct["scaling_ratio"], rel=1e-12)
    assert float(g.grad) == pytest.approx(direct["gravity"], rel=1e-12)
    assert float(loss) == pytest.approx(AG.stress1_np(pos, srcs, hops), rel=1e-12)


def test_record_refuses_what_it_cannot_reverse():
    s, t = bh_graph(50)
    with G.Session(G.Topology(50, s, t), G.Fa2Params(),
                   G.Fa2Ext(repulsion="pm", pm_grid=64)) as pm:
        with pytest.raises(G.StageError):
            pm.record(3, 2)
    topos = [G.Topology(50, s, t), G.Topology(50, s, t)]
    with G.Session.batch(topos) as bs:
        with pytest.raises(G.StageError):
            bs.record(3, 2)
    with G.Session(G.Topology(50, s, t)) as one:
        with pytest.raises(G.StageError):
            one.adjoint(np.zeros((50, 2)))  # nothing recorded


def test_tree_adjoint_is_bit_reproducible():
    """Cell and many-body-leaf accumulators add exactly (integer limbs), so the order the
    waves' atomics land in moves no bit: two reverse sweeps give the same gradients, with
    coincident points putting several bodies in one leaf."""
    n, iters = 3000, 8
    s, t = bh_graph(n)
    topo = G.Topology(n, s, t)
    rng = np.random.default_rng(12)
    start = rng.uniform(-80, 80, (n, 2))
    start[1000:1040] = start[:40]  # 40 leaves of two bodies
    w_raw = rng.uniform(0.5, 1.5, s.size)
    cvec = rng.normal(size=(n, 2))
    runs = []
    for _ in range(2):
        _, g = device(topo, w_raw, 2.0, 1.0, iters, 3, cvec, start, repulsion="tree",
                      theta=0.8)
        runs.append(g)
    for key in ("weights", "start"):
        assert np.array_equal(runs[0][key], runs[1][key]), key
    assert runs[0]["scaling_ratio"] == runs[1]["scaling_ratio"]
    assert runs[0]["gravity"] == runs[1]["gravity"]
    assert np.isfinite(runs[0]["start"]).all()
