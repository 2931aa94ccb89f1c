"""Batched layouts: many graphs in one session. Each graph's result must be bit-identical
(FP64) to the same graph run alone: the batch shares launches, not arithmetic."""
import numpy as np
import pytest

import graph_render_amd as G
from tests.refgraphs import bh_graph

pytestmark = pytest.mark.gpu


def graphs(seed, count, lo=2, hi=3000):
    rng = np.random.default_rng(seed)
    out = []
    for b in range(count):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if b % 3 == 0:
            s, t = bh_graph(n)
        else:
            m = int(n * rng.uniform(1, 5))
            s = rng.integers(0, n, m).astype(np.uint32)
            t = rng.integers(0, n, m).astype(np.uint32)
        out.append(G.Topology(n, s, t))
    return out


@pytest.mark.parametrize("repulsion", ["tree", "dense"])
def test_batch_equals_graphs_run_alone(repulsion):
    topos = graphs(7, 24, hi=2000 if repulsion == "tree" else 400)
    topos.insert(5, G.Topology(1, [], []))      # a lone node
    topos.insert(9, G.Topology(0, [], []))      # an empty graph
    p = G.Fa2Params(max_iter=40)
    x = G.Fa2Ext(repulsion=repulsion)
    with G.Session.batch(topos, p, x) as bs:
        done = bs.step(12)
        pb = bs.positions()
        scal = [bs.graph_scalars(b) for b in range(len(topos))]
    assert done == 12
    off = 0
    for b, t in enumerate(topos):
        with G.Session(t, p, G.Fa2Ext(repulsion=repulsion)) as s1:
            d1 = s1.step(12)
            p1 = s1.positions()
            sc = s1.scalars()
        assert np.array_equal(pb[off:off + t.n], p1), b
        assert scal[b]["iters_done"] == d1 and scal[b]["speed"] == sc["speed"], b
        off += t.n


def test_batch_early_exit_is_per_graph():
    # a 2-node graph settles long before a 600-node one; the big one keeps moving
    topos = [G.Topology(2, [0], [1]), G.Topology(600, *bh_graph(600))]
    with G.Session.batch(topos, G.Fa2Params(max_iter=300)) as bs:
        bs.step(300)
        a, b = bs.graph_scalars(0), bs.graph_scalars(1)
    with G.Session(topos[0], G.Fa2Params(max_iter=300)) as s0:
        d0 = s0.step(300)
    assert a["iters_done"] == d0
    assert b["iters_done"] == 300


def test_batch_forces_match_single():
    topos = graphs(11, 10, hi=5000)
    with G.Session.batch(topos, G.Fa2Params(), G.Fa2Ext(precision="fp32")) as bs:
        fb = bs.forces(7)
    off = 0
    for t in topos:
        with G.Session(t, G.Fa2Params(), G.Fa2Ext(precision="fp32")) as s1:
            f1 = s1.forces(7)
        assert np.array_equal(fb[off:off + t.n], f1)
        off += t.n


def test_large_batch_equals_graphs_run_alone():
    """A batch of about two million nodes: its centre-of-mass tiles (4096 cells each, about
    nine span cells per tile) leave several thousand span cells, more than the one-workgroup
    pass holds in LDS, so the span pass runs level by level over the GPU. Graphs picked from
    across the batch are still bit-identical (FP64) to themselves run alone."""
    rng = np.random.default_rng(11)
    sizes = rng.integers(20000, 30000, 80)
    topos = [G.Topology(int(n), *bh_graph(int(n))) for n in sizes]
    p = G.Fa2Params(max_iter=40)
    print(f"large batch: {len(topos)} graphs, {int(sizes.sum())} nodes", flush=True)
    with G.Session.batch(topos, p) as bs:
        assert bs.step(6) == 6
        pb = bs.positions()
        tr = bs.tree(bounds=False)
    # span cells: a cell whose subtree [k, skip) runs past the end of its 4096-cell tile; the
    # one-workgroup pass holds 3072 of them
    c = tr["ncells"]
    k = np.arange(c)
    tile_end = np.minimum((k // 4096 + 1) * 4096, c)
    spans = int((tr["skip"] > tile_end).sum())
    print(f"large batch: {c} cells, {spans} span cells", flush=True)
    assert spans > 3072
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for b in (0, 27, 53, 79):
        with G.Session(topos[b], p) as s1:
            assert s1.step(6) == 6
            p1 = s1.positions()
        assert np.array_equal(pb[offs[b]:offs[b + 1]], p1), b
