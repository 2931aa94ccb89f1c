"""Launch schedules and the autotuner: a schedule changes how kernels are launched, never a
result bit. Every candidate the autotuner can pick gives the default schedule's forces and
trajectory bit for bit (FP64 and FP32, one graph and a batch); the autotuner installs one of
them, leaves the state where it was and hands it to later sessions of the same size class."""
import itertools

import numpy as np
import pytest

import graph_render_amd as G
from tests.refgraphs import bh_graph

pytestmark = pytest.mark.gpu

CANDIDATES = [dict(walk_block=wb, walk_xcd=x, sort_items=si, finish_split=fs)
              for wb, x, si, fs in itertools.product((64, 128, 256), (0, 1), (4, 8, 16),
                                                     (1, 2, 4, 8))]


def run(topo, ext, sched=None, steps=3):
    with G.Session(topo, G.Fa2Params(max_iter=100), ext) as s:
        if sched:
            s.set_schedule(**sched)
            assert s.schedule() == sched
        f = s.forces(7)
        s.step(steps)
        return f, s.positions()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_every_candidate_schedule_gives_the_same_bits(precision):
    n = 20000
    s, t = bh_graph(n)
    topo = G.Topology(n, s, t)
    ext = G.Fa2Ext(theta=1.2, precision=precision)
    f0, p0 = run(topo, ext)
    rng = np.random.default_rng(1)
    picks = [CANDIDATES[i] for i in rng.choice(len(CANDIDATES), 12, replace=False)]
    picks += [dict(walk_block=64, walk_xcd=0, sort_items=4, finish_split=8),
              dict(walk_block=256, walk_xcd=1, sort_items=16, finish_split=1)]
    for sc in picks:
        print(f"schedule {sc} {precision}", flush=True)
        f, p = run(topo, ext, sc)
        assert np.array_equal(f, f0) and np.array_equal(p, p0), sc


def test_batch_schedules_give_the_same_bits():
    topos = [G.Topology(n, *bh_graph(n)) for n in (300, 5000, 77, 12000)]
    ext = G.Fa2Ext(theta=1.0, precision="fp32")
    out = []
    for sc in (None, dict(walk_block=64, walk_xcd=0, sort_items=4, finish_split=8),
               dict(walk_block=128, walk_xcd=1, sort_items=16, finish_split=2)):
        with G.Session.batch(topos, G.Fa2Params(max_iter=100), ext) as b:
            if sc:
                b.set_schedule(**sc)
            f = b.forces(7)
            b.step(4)
            out.append((f, b.positions()))
    for f, p in out[1:]:
        assert np.array_equal(f, out[0][0]) and np.array_equal(p, out[0][1])


def test_autotune_installs_a_candidate_keeps_the_state_and_caches_it():
    n = 30000
    s, t = bh_graph(n)
    topo = G.Topology(n, s, t)
    ext = G.Fa2Ext(theta=1.2, precision="fp32")
    with G.Session(topo, G.Fa2Params(max_iter=100), ext) as a:
        a.step(2)
        p = a.positions()
        f = a.forces(7)
        tuned = a.autotune(3)
        print(f"autotuned schedule: {tuned}", flush=True)
        assert tuned in CANDIDATES and a.schedule() == tuned
        assert np.array_equal(a.positions(), p) and np.array_equal(a.forces(7), f)
        a.step(3)
        pa = a.positions()
    with G.Session(topo, G.Fa2Params(max_iter=100), ext) as b:
        assert b.schedule() == tuned  # the size class's cached choice
        b.step(5)
        assert np.array_equal(b.positions(), pa)


def test_set_schedule_refuses_values_outside_the_sets():
    n = 500
    s, t = bh_graph(n)
    with G.Session(G.Topology(n, s, t)) as sess:
        for bad in (dict(walk_block=96), dict(sort_items=32), dict(finish_split=3),
                    dict(walk_xcd=2)):
            with pytest.raises(G.ParamError):
                sess.set_schedule(**bad)
        sess.step(2)
