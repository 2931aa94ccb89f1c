"""Distinct handles run concurrently (the drop-in boundary's threading rule: reentrant per
handle, one HIP stream per session, no hidden global state besides the device context and
the per-thread last error). ctypes releases the interpreter lock inside each foreign call, so
the threads' sessions and stage runs overlap on the device.

Bar: bit for bit. Each run gives the same bytes run alone and run beside another."""
import threading

import numpy as np
import pytest

import graph_render_amd as G
from graph_render_amd import _abi as A
from graph_render_amd import barnes_hut as BH
from graph_render_amd import forcesim as FS
from tests.refgraphs import bh_graph

pytestmark = pytest.mark.gpu


def fp64_run(n, steps, theta):
    s, t = bh_graph(n)
    with G.Session(G.Topology(n, s, t), G.Fa2Params(), G.Fa2Ext(theta=theta)) as sess:
        sess.step(steps)
        return sess.positions()


def stage_run(n):
    s, t = bh_graph(n)
    g = G.find("layout.forceatlas2.barnes_hut").run(G.Topology(n, s, t), G.Fa2Params(max_iter=15))
    return np.stack([g.x, g.y], 1)


def d3_run(n, ticks, mesh):
    s, t = bh_graph(n)
    with BH.ForceSession(G.Topology(n, s, t), None) as S:
        if mesh:
            S.with_particle_mesh()
        S.step(ticks)
        x, y, vx, vy = S.state()
    return np.stack([x, y, vx, vy], 1)


def forcesim_run(graphs):
    rng = np.random.default_rng(graphs)
    topos, params = [], []
    for b in range(graphs):
        n = int(rng.integers(2, 300))
        s, t = bh_graph(n)
        topos.append(G.Topology(n, s, t))
        params.append(FS.Fa2ForceSimParams(iterations=int(rng.integers(1, 60)), seed=b))
    return np.concatenate([xyz.ravel() for xyz, _ in FS.batch(topos, params)])


def run_threads(jobs):
    out, errs = [None] * len(jobs), []

    def work(k, fn):
        try:
            out[k] = fn()
        except Exception as e:  # noqa: BLE001 - reported below
            errs.append(repr(e))

    ts = [threading.Thread(target=work, args=(k, fn)) for k, fn in enumerate(jobs)]
    for th in ts:
        th.start()
    for th in ts:
        th.join(timeout=120)
    assert not any(th.is_alive() for th in ts), "a thread did not finish"
    assert not errs, errs
    return out


def test_sessions_on_two_threads_give_their_bits_alone():
    jobs = [lambda: fp64_run(20000, 12, 0.8), lambda: fp64_run(30000, 9, 1.2),
            lambda: stage_run(15000), lambda: fp64_run(5000, 30, 0.8)]
    alone = [fn() for fn in jobs]
    print("concurrent sessions: 4 threads", flush=True)
    together = run_threads(jobs)
    for a, b in zip(alone, together):
        assert a.shape == b.shape and np.array_equal(a.view(np.uint64) if a.dtype == np.float64
                                                     else a, b.view(np.uint64)
                                                     if b.dtype == np.float64 else b)


def test_every_layout_family_on_its_own_thread():
    # the d3 simulations run a side stream with events beside their session stream; forcesim
    # one workgroup per graph on its own stream; each is a thread of its own here
    jobs = [lambda: d3_run(20000, 6, False), lambda: d3_run(30000, 5, True),
            lambda: forcesim_run(64), lambda: fp64_run(25000, 10, 1.0),
            lambda: d3_run(8000, 12, True)]
    alone = [fn() for fn in jobs]
    print("concurrent layouts: 5 threads", flush=True)
    together = run_threads(jobs)
    for a, b in zip(alone, together):
        a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
        assert a.dtype == b.dtype and a.shape == b.shape
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_the_last_error_is_per_thread():
    # a refusal on one thread leaves another thread's last error as its own call left it
    ready, go = threading.Event(), threading.Event()
    seen = {}

    def refuse():
        import struct
        s, t = bh_graph(50)
        topo = G.Topology(50, s, t)
        # max_iter 0 lies outside the published range: refused, never clamped
        try:
            G.run_params("layout.forceatlas2.barnes_hut", topo, struct.pack("<5d", 0, 1, 2, 1, 0))
            seen["refuse_code"] = A.GR_OK
        except G.ParamError as e:
            seen["refuse_code"] = e.code
        ready.set()
        go.wait(30)
        seen["refuse_last"] = A.last_error()

    def succeed():
        ready.wait(30)
        seen["ok"] = fp64_run(2000, 3, 0.8) is not None
        seen["ok_last"] = A.last_error()
        go.set()

    run_threads([refuse, succeed])
    assert seen["refuse_code"] == A.GR_ERR_PARAM_OUT_OF_RANGE
    assert seen["refuse_last"][0] == A.GR_ERR_PARAM_OUT_OF_RANGE
    assert seen["ok"] and seen["ok_last"][0] == A.GR_OK
