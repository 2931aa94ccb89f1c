#!/usr/bin/env python3
"""FA2-BH iterations/s on one MI355X: 1M-node / ~5M-edge stochastic block model, 2D, FP32
forces with FP64 swing/traction reductions, theta 1.2 (the large Barnes-Hut workload).

Prints one JSON line. `value` is FP32 iterations/s with the graph and state already
resident in HBM; the same line carries FP64 iterations/s, the FP32-vs-FP64 force error
(per-node rms and whole-field, as the reference's BH tests measure it), the walk kernel's
roofline position (HIP events on the session stream) and the CPU oracle's rate on a
bounded sample of the same workload (one core).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FA2-BH iters/sec @1M nodes/5M edges, FP32 vs FP64, force err vs FP64"
HBM_PEAK_GBS = 8000.0


def errors(approx, exact):
    d2 = ((approx - exact) ** 2).sum(1)
    e2 = (exact ** 2).sum(1)
    ok = e2 > 0
    return float(np.sqrt((d2[ok] / e2[ok]).mean())), float(np.sqrt(d2.sum() / e2.sum()))


def walk_bytes(n: int, ncells: int, precision: str) -> int:
    """Algorithmic bytes of one walk launch, the per-unit figures of the FA2-BH byte model
    (2D, FP32 forces): per body 20 B (its sorted position, mass and point index read, its
    force written) and per cell 48 B (one read of its record); doubled float sizes in FP64."""
    if precision == "fp64":
        return n * 36 + ncells * 72
    return n * 20 + ncells * 48


def iteration_bytes(n: int, ncells: int, m_simple: int) -> int:
    """Algorithmic bytes of one whole FA2-BH iteration (2D, FP32 forces): keys 20n, sort
    24n, build 8n + 32C, centres of mass 12n + 28C, walk 20n + 48C, attraction
    4(n + 1) + 16 * 2m + 24n, centroid 8n, swing/traction 20n, apply 28n."""
    return 168 * n + 108 * ncells + 32 * m_simple + 4


def walk_traffic(n: int):
    """HBM bytes per k_walk<float,2> launch on this workload, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (profiles/walk_1M_pmc.md), or None for other sizes."""
    path = os.path.join(ROOT, "profiles", "walk_1M_pmc.md")
    if n != 1_000_000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_walk_fp32_1M"):
            return int(line.split("=")[1])
    return None


def collide_traffic(n: int):
    """HBM bytes per k_sim_collide launch at the 100k-node gate model, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/d3sim_100k_kernel_stats.md), or None for
    other sizes."""
    path = os.path.join(ROOT, "profiles", "d3sim_100k_kernel_stats.md")
    if n != 100_000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_collide_100k"):
            return int(line.split("=")[1])
    return None


def profile_figure(name: str, n: int, profile: str, at_n: int):
    """A `name = value` figure from a committed rocprofv3 summary under profiles/ (HBM bytes
    per launch from FETCH_SIZE/WRITE_SIZE), or None for other sizes."""
    path = os.path.join(ROOT, "profiles", profile)
    if n != at_n or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(name + " ="):
            return int(line.split("=")[1])
    return None


def cpu_info() -> dict:
    """The host CPU as the baseline states it: model, logical cores the process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "host_cores": len(os.sched_getaffinity(0)),
            "host_cpus_total": os.cpu_count()}


def pin_one_core() -> int:
    """Pins the calling thread (the one that runs the oracle) to the first core it may use:
    the reference path is single-threaded, so its baseline runs on one core."""
    core = sorted(os.sched_getaffinity(0))[0]
    os.sched_setaffinity(0, {core})
    return core


def cpu_baseline(w, params, budget_s: float):
    """The oracle (the reference's FA2-BH restated in C, FP64, one thread) on the same graph
    and start, pinned to one core, iterated until about budget_s seconds are spent; then the
    speed check on the reference's own timed graph, seeded_model(0, 1e6, 8) at theta 0.8
    with its Mulberry32 start (published: 123 031.69 ms for 100 iterations, 1.23 s each)."""
    from oracle import oracle as O
    info = cpu_info()
    core = pin_one_core()
    t0 = time.perf_counter()
    st = O.Fa2(w["n"], w["source"], w["target"], start=w["start"], weights=w["weights"],
               tree=True, theta2=w["theta"] ** 2, scaling_ratio=params.scaling_ratio,
               gravity=params.gravity, jitter_tolerance=params.jitter_tolerance)
    setup = time.perf_counter() - t0
    its, t0 = 0, time.perf_counter()
    while True:
        st.iterate()
        its += 1
        el = time.perf_counter() - t0
        if el >= budget_s or its >= 50:
            break
    del st
    # speed check against the published single-thread figure
    gs, gt = O.gate_edges(1_000_000)
    ref = O.Fa2(1_000_000, gs, gt, tree=True)
    ref.iterate()  # warm
    t1 = time.perf_counter()
    for _ in range(3):
        ref.iterate()
    s_per_it = (time.perf_counter() - t1) / 3
    del ref
    return {"value": its / el, "unit": "iters/s", "cores": 1, "kind": "port",
            "sample": f"{its} FA2-BH FP64 iterations of the same 1M-node graph and start "
                      f"(oracle/fa2_oracle.c, gcc -O2, one thread pinned to core {core}; "
                      f"{setup:.1f}s setup untimed)",
            **info,
            "speed_check": {"graph": "seeded_model(0, 1000000, 8), 1 549 929 raw edges, "
                                     "theta 0.8, Mulberry32 start, iterations 2-4",
                            "edges": int(gs.size), "s_per_iter": round(s_per_it, 4),
                            "published_s_per_iter": 1.2303,
                            "ratio_to_published": round(s_per_it / 1.2303, 3)}}


def batch_cpu_baseline(topos, starts, offs, params, budget_s: float):
    """The oracle on every 32nd graph of the batch (FP64, one thread, theta 1.0), eight
    iterations each after an untimed set-up; the batch's CPU time per iteration is the
    sample's time scaled by total nodes over sampled nodes."""
    from oracle import oracle as O
    info = cpu_info()
    core = pin_one_core()
    sample = list(range(0, len(topos), 32))
    spent, nodes, its = 0.0, 0, 8
    for b in sample:
        t = topos[b]
        st = O.Fa2(t.n, t.source, t.target, start=starts[offs[b]:offs[b] + t.n], tree=True,
                   theta2=1.0, scaling_ratio=params.scaling_ratio, gravity=params.gravity,
                   jitter_tolerance=params.jitter_tolerance)
        t0 = time.perf_counter()
        for _ in range(its):
            st.iterate()
        spent += (time.perf_counter() - t0) / its
        nodes += t.n
        if spent * its > budget_s:
            break
    total = sum(t.n for t in topos)
    per_batch_iter = spent * total / nodes
    return {"value": 1.0 / per_batch_iter, "unit": "batch iters/s", "cores": 1, "kind": "port",
            "sample": f"{its} FA2-BH FP64 iterations on each of {len(sample)} graphs (every "
                      f"32nd, {nodes} nodes), scaled by {total}/{nodes} nodes "
                      f"(oracle/fa2_oracle.c, gcc -O2, one thread pinned to core {core})",
            **info}


def bench_batch(args):
    """4096 graphs (log-uniform 50..20000 nodes, half preferential attachment, half
    Erdos-Renyi) laid out together in one session, FP32, theta 1.0: batch iterations/s
    (one iteration advances every graph), graph-iterations/s and node-iterations/s."""
    import graph_render_amd as G
    from graph_render_amd import synth
    t0 = time.perf_counter()
    sizes, srcs, dsts, starts = synth.batch_workload(args.graphs)
    topos = [G.Topology(n, s, t) for n, s, t in zip(sizes, srcs, dsts)]
    offs = np.concatenate([[0], np.cumsum(sizes)])
    gen_s = time.perf_counter() - t0
    params = G.Fa2Params(max_iter=args.warmup + args.steps)
    t0 = time.perf_counter()
    bs = G.Session.batch(topos, params, G.Fa2Ext(theta=1.0, precision="fp32", start=starts))
    create_s = time.perf_counter() - t0
    bs.step(args.warmup)
    t0 = time.perf_counter()
    done = bs.step(args.steps)
    el = time.perf_counter() - t0
    live = sum(1 for b in range(len(topos)) if not bs.graph_scalars(b)["done"])
    walk_ms = bs.time("walk", 5)
    tree_ms = bs.time("tree", 3)
    it_ms = bs.time("iteration", 3)
    bs.close()
    its = done / el
    n = int(offs[-1])
    cpu = None if args.no_cpu else batch_cpu_baseline(topos, starts, offs, params,
                                                      args.cpu_seconds)
    line = {
        "metric": "FA2-BH batched layout, batch iterations/s (every graph advances one "
                  "iteration)",
        "value": round(its, 3), "unit": "batch iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 / its, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(topos)} graphs, log-uniform 50..20000 nodes, even "
                               "graphs preferential attachment m 1..6, odd Erdos-Renyi mean "
                               "degree 4..12, seeds 0..4095, starts U[-100,100]^2, 2D FP32, "
                               "theta 1.0",
                   "graphs": len(topos), "nodes": n,
                   "raw_edges": int(sum(len(s) for s in srcs))},
        "graph_iters_per_s": round(its * len(topos), 1),
        "node_iters_per_s": round(its * n, 1),
        "graphs_still_moving": live, "walk_ms": round(walk_ms, 4),
        "tree_build_ms": round(tree_ms, 4), "iteration_ms": round(it_ms, 4),
        "generate_s": round(gen_s, 2), "session_create_s": round(create_s, 2),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def max_rel(approx, exact) -> float:
    """The largest per-node relative error |a - e| / |e| (nodes with e != 0)."""
    d = np.sqrt(((approx - exact) ** 2).sum(1))
    e = np.sqrt((exact ** 2).sum(1))
    ok = e > 0
    return float((d[ok] / e[ok]).max())


def conditioned(approx, exact, parts):
    """The per-node error against the node's force scale: |a - e| / (|attraction| + |repulsion|
    + |gravity|) of the FP64 parts at the same positions. Where the three nearly cancel, |e| is
    a small fraction of that scale and |a - e| / |e| grows by the same factor whatever the
    arithmetic; this is the error with that factor taken out. Also the count of nodes whose
    error relative to |e| passes 1e-5, and the cancellation |e| / scale at the worst of them."""
    d = np.sqrt(((approx - exact) ** 2).sum(1))
    e = np.sqrt((exact ** 2).sum(1))
    scale = sum(np.sqrt((q ** 2).sum(1)) for q in parts)
    ok = scale > 0
    r = np.zeros_like(d)
    r[e > 0] = d[e > 0] / e[e > 0]
    worst = int(np.argmax(r))
    return {"max_vs_force_scale": float((d[ok] / scale[ok]).max()),
            "nodes_over_1e-5": int((r > 1e-5).sum()),
            "cancellation_at_max": float(e[worst] / scale[worst]) if scale[worst] > 0 else None}


def rel_errors(approx, exact):
    """Per-node relative error |a - e| / |e| (max and mean over nodes with e != 0) and the
    whole-field ratio sqrt(sum |a - e|^2 / sum |e|^2)."""
    d = np.sqrt(((approx - exact) ** 2).sum(1))
    e = np.sqrt((exact ** 2).sum(1))
    ok = e > 0
    r = d[ok] / e[ok]
    return {"max": float(r.max()), "mean": float(r.mean()),
            "whole_field": float(np.sqrt((d ** 2).sum() / (e ** 2).sum()))}


def bench_precision(args):
    """3D FA2-BH (octree) precision study on a 250K-node random geometric graph (unit cube,
    mean degree 8, seed 2): FP64, FP32 and FP32-Kahan forces at theta 0 and 0.8 against the
    exact O(n^2) FP64 pair sum, and the position drift after `steps` iterations against the
    exact FP64 layout from the same start (drift in units of the exact layout's extent)."""
    import graph_render_amd as G
    from graph_render_amd import synth
    n = args.nodes or 250_000
    s, t, _ = synth.random_geometric_3d(n, 8.0, 2)
    topo = G.Topology(n, s, t)
    iters = args.steps
    params = G.Fa2Params(max_iter=iters, dim=3, seed=2)
    with G.Session(topo, params, G.Fa2Ext(repulsion="dense", precision="fp64")) as ex:
        start = ex.positions()
        f_exact = ex.forces(7)
        t0 = time.perf_counter()
        ex.step(iters)
        exact_its = iters / (time.perf_counter() - t0)
        p_exact = ex.positions()
    extent = float((p_exact.max(0) - p_exact.min(0)).max())
    rows = []
    for theta in (0.0, 0.8):
        f64 = None
        for prec in ("fp64", "fp32", "fp32_kahan"):
            with G.Session(topo, params, G.Fa2Ext(theta=theta, precision=prec,
                                                  start=start)) as ss:
                f = ss.forces(7)
                t0 = time.perf_counter()
                d = ss.step(iters)
                its = d / (time.perf_counter() - t0)
                p = ss.positions()
            if prec == "fp64":
                f64 = f
            drift = np.sqrt(((p - p_exact) ** 2).sum(1)) / extent
            rows.append({"theta": theta, "precision": prec, "iters_per_s": round(its, 3),
                         "force_err_vs_exact_fp64": rel_errors(f, f_exact),
                         "force_err_vs_fp64_same_theta": rel_errors(f, f64),
                         "drift_after_iters": {"max": float(drift.max()),
                                               "mean": float(drift.mean())}})
    head = next(r for r in rows if r["theta"] == 0.8 and r["precision"] == "fp32")
    line = {
        "metric": "FA2-BH 3D precision study: force error vs exact FP64 and position drift",
        "value": head["iters_per_s"], "unit": "iters/s", "n_gpus": 1, "steps": iters,
        "warmup": 0, "ms_per_step": round(1e3 / head["iters_per_s"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"FA2-BH 3D octree, random geometric graph in the unit cube, "
                               f"{n} nodes, mean degree 8, seed 2, Mulberry32 start seed 2",
                   "nodes": n, "edges": int(s.size), "edges_sha256": synth.digest(s, t)},
        "exact_fp64_dense_iters_per_s": round(exact_its, 3),
        "exact_layout_extent": extent,
        "rows": rows,
    }
    print(json.dumps(line), flush=True)


def bench_adjoint(args):
    """Reverse mode through unrolled FA2-BH: 10K-node SBM (16 equal blocks, p_in 0.01,
    p_out 0.0005, seed 3), iterations recorded with a checkpoint every 20, loss = Kruskal
    stress-1 against BFS hops from 256 evenly spaced sources. Times forward and reverse over
    `steps` iterations in FP32 and FP64. Then, at horizons of 10 .. `steps` iterations (and
    with dense repulsion, whose map has no accept/open switches, at 10 .. 50), compares
    d loss / d weight on 64 random edges, d scaling ratio and d gravity with central
    differences of the FP64 forward at h 1e-3 and 1e-6: the two step sizes agreeing with
    each other is what says the finite difference measures a derivative at all."""
    import torch
    import graph_render_amd as G
    from tests import fa2_autograd as AG
    from graph_render_amd import synth
    n = args.nodes or 10_000
    s, t = synth.sbm_equal(n, 16, 0.01, 0.0005, seed=3)
    topo = G.Topology(n, s, t)
    every = 20
    srcs = AG.stress_sources(n, 256)
    hops = AG.bfs_hops(n, s, t, srcs)
    w0 = np.ones(s.size)

    def reverse(precision, iters, repulsion="tree"):
        spec = AG.UnrolledSpec(topology=topo, iters=iters, checkpoint_every=every,
                               precision=precision, repulsion=repulsion)
        t0 = time.perf_counter()
        sess, pos = AG.run_forward(spec, w0, 2.0, 1.0, record=True)
        t_fwd = time.perf_counter() - t0
        pt = torch.tensor(pos, requires_grad=True)
        loss = AG.stress1(pt, srcs, hops)
        loss.backward()
        t0 = time.perf_counter()
        g = sess.adjoint(pt.grad.numpy())
        t_bwd = time.perf_counter() - t0
        sess.close()
        return float(loss.detach()), g, t_fwd, t_bwd, spec.applied[-1]

    L32, g32, f32, b32, a32 = reverse("fp32", args.steps)
    L64, g64, f64, b64, a64 = reverse("fp64", args.steps)

    # raw edges the simple graph keeps (first occurrence of each unordered pair)
    lo, hi = np.minimum(s, t).astype(np.int64), np.maximum(s, t).astype(np.int64)
    _, first = np.unique(lo * n + hi, return_index=True)
    first = first[lo[first] != hi[first]]
    rng = np.random.default_rng(3)
    pick = np.sort(rng.choice(first, 64, replace=False))

    def vec_err(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

    checks = [("tree", h) for h in (10, 25, 50) if h < args.steps] + [("tree", args.steps)]
    checks += [("dense", h) for h in (10, 25, 50) if h <= args.steps]
    rows = []
    t_fd = time.perf_counter()
    for rep, H in checks:
        def loss64(w, k=2.0, g=1.0):
            spec = AG.UnrolledSpec(topology=topo, iters=H, precision="fp64", repulsion=rep)
            _, pos = AG.run_forward(spec, w, k, g)
            return AG.stress1_np(pos, srcs, hops)

        fds = {}
        for h in (1e-3, 1e-6):
            fw = []
            for e in pick:
                wp, wm = w0.copy(), w0.copy()
                wp[e] += h
                wm[e] -= h
                fw.append((loss64(wp) - loss64(wm)) / (2 * h))
            fk = (loss64(w0, k=2.0 + h) - loss64(w0, k=2.0 - h)) / (2 * h)
            fg = (loss64(w0, g=1.0 + h) - loss64(w0, g=1.0 - h)) / (2 * h)
            fds[h] = (np.array(fw), fk, fg)
        if rep == "tree" and H == args.steps:
            ga64, ga32 = g64, g32
        else:
            ga64 = reverse("fp64", H, rep)[1]
            ga32 = reverse("fp32", H, rep)[1]
        row = {"repulsion": rep, "iterations": H,
               "grad_norm_fp64_adjoint": float(np.linalg.norm(ga64["weights"][pick])),
               "fd_h1e-3_vs_fd_h1e-6": vec_err(fds[1e-3][0], fds[1e-6][0])}
        for name, ga in (("fp64", ga64), ("fp32", ga32)):
            for h in (1e-3, 1e-6):
                fw, fk, fg = fds[h]
                row[f"adjoint_{name}_vs_fd_h{h:.0e}"] = {
                    "weights_vector_rel_err": vec_err(ga["weights"][pick], fw),
                    "weights_max_abs_err_over_max_fd":
                        float(np.abs(ga["weights"][pick] - fw).max() / np.abs(fw).max()),
                    "scaling_ratio_rel_err": float(abs(ga["scaling_ratio"] - fk) / abs(fk)),
                    "gravity_rel_err": float(abs(ga["gravity"] - fg) / abs(fg))}
        rows.append(row)
        print(f"{rep} horizon {H}: fd(1e-3) vs fd(1e-6) {row['fd_h1e-3_vs_fd_h1e-6']:.2e}; "
              f"fp64 adjoint vs fd(1e-3) "
              f"{row['adjoint_fp64_vs_fd_h1e-03']['weights_vector_rel_err']:.2e}, vs fd(1e-6) "
              f"{row['adjoint_fp64_vs_fd_h1e-06']['weights_vector_rel_err']:.2e}",
              file=sys.stderr, flush=True)
    fd_s = time.perf_counter() - t_fd

    its = args.steps / (f32 + b32)
    line = {
        "metric": "FA2-BH adjoint: unrolled iterations/s, forward + reverse (checkpoint 20)",
        "value": round(its, 3), "unit": "iters/s", "n_gpus": 1, "steps": args.steps,
        "warmup": 0, "ms_per_step": round(1e3 / its, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"SBM {n} nodes, 16 equal blocks, p_in 0.01, p_out 0.0005, "
                               f"seed 3; FA2-BH theta 0.8, {args.steps} iterations, checkpoint "
                               "every 20; loss Kruskal stress-1 vs BFS hops from 256 sources",
                   "nodes": n, "raw_edges": int(s.size), "edges_sha256": synth.digest(s, t)},
        "loss_fp32": L32, "loss_fp64": L64, "iterations_applied": [a32, a64],
        "fp32": {"forward_s": round(f32, 4), "reverse_s": round(b32, 4)},
        "fp64": {"forward_s": round(f64, 4), "reverse_s": round(b64, 4)},
        "fd": {"edges": 64, "step_sizes": [1e-3, 1e-6], "seconds": round(fd_s, 2)},
        "checks": rows,
    }
    print(json.dumps(line), flush=True)


def bench_autotune(args):
    """The autotuner on each benchmark workload: iterations/s with the default schedule and
    with the one gr_fa2_session_autotune picks (two sessions from the same start, 5 untimed
    then `steps` timed iterations each), and the schedules themselves: the table
    gr_fa2_schedule_select is written from."""
    import graph_render_amd as G
    from graph_render_amd import synth

    def single(name, precision, theta):
        w = synth.workload(name)
        topo = G.Topology(w["n"], w["source"], w["target"])
        p = G.Fa2Params(max_iter=10000, dim=w["dim"], seed=w["seed"])

        def make():
            return G.Session(topo, p, G.Fa2Ext(theta=theta, precision=precision,
                                                start=w["start"], edge_weights=w["weights"]))
        return w["n"], int(w["source"].size), make

    def batch():
        sizes, srcs, dsts, starts = synth.batch_workload(4096)
        topos = [G.Topology(n, s_, t_) for n, s_, t_ in zip(sizes, srcs, dsts)]

        def make():
            return G.Session.batch(topos, G.Fa2Params(max_iter=10000),
                                   G.Fa2Ext(theta=1.0, precision="fp32", start=starts))
        return int(sum(sizes)), int(sum(s_.size for s_ in srcs)), make

    cases = [("sbm10k", "fp32", 0.8), ("sbm10k", "fp64", 0.8), ("rgg250k", "fp32", 0.8),
             ("sbm1m", "fp32", 1.2), ("batch4096", "fp32", 1.0)]
    rows = []
    for name, precision, theta in cases:
        n, m, make = batch() if name == "batch4096" else single(name, precision, theta)
        rates = {}
        for mode in ("default", "autotuned"):
            sess = make()
            default = sess.schedule()
            sess.step(5)
            tuned = sess.autotune(5) if mode == "autotuned" else None
            t0 = time.perf_counter()
            d = sess.step(args.steps)
            rates[mode] = d / (time.perf_counter() - t0)
            sess.close()
            if tuned is not None:
                chosen = tuned
        rows.append({"workload": name, "precision": precision, "nodes": n, "raw_edges": m,
                     "default_schedule": default, "autotuned_schedule": chosen,
                     "iters_per_s_default": round(rates["default"], 2),
                     "iters_per_s_autotuned": round(rates["autotuned"], 2),
                     "speedup": round(rates["autotuned"] / rates["default"], 3)})
        print(json.dumps(rows[-1]), file=sys.stderr, flush=True)
    head = next(r for r in rows if r["workload"] == "sbm1m")
    print(json.dumps({
        "metric": "FA2-BH iterations/s with autotuned launch schedules",
        "value": head["iters_per_s_autotuned"], "unit": "iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 5, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "the benchmark workloads, default vs autotuned schedule"},
        "rows": rows}), flush=True)


def golden_spiral(n: int) -> np.ndarray:
    """barnes_hut/seed.rs: node i at 12 sqrt(i + 1) along the golden-angle spiral (numpy's
    cos/sin; the device and the oracle read the same array)."""
    i = np.arange(n, dtype=np.float64)
    r = 12.0 * np.sqrt(i + 1.0)
    a = i * 2.399963229728653
    return np.stack([np.cos(a) * r, np.sin(a) * r], 1)


def bench_d3(args):
    """d3's many-body pass over the quadtree (layout.force's Barnes-Hut charge pass): 1M
    nodes at the simulation's golden-spiral seed positions, theta 0.9, chargeStrength -90,
    distanceMin 1, chargeDistanceMax 520, alpha 1. Passes/s with the positions resident on the
    device (tree build, bodies, node ids, walk), the walk's roofline position, bit-equality of
    the device deltas with the oracle's at the full size, and the oracle's pass timed on one
    core."""
    from graph_render_amd import barnes_hut
    n = args.nodes or 1_000_000
    xy = golden_spiral(n)
    kw = dict(theta=0.9, charge=-90.0, alpha=1.0, distance_min=1.0, distance_max=520.0,
              seed=0, tick=0)
    got, ms_pass, ms_walk, ncells = barnes_hut.time_pass(xy, reps=max(args.steps, 1), **kw)
    print(f"d3 BH n={n}: pass {ms_pass:.3f} ms, walk {ms_walk:.3f} ms, {ncells} cells",
          file=sys.stderr, flush=True)
    # algorithmic bytes of one walk launch: per node its position (16 B), point index (4 B)
    # and delta (16 B); per cell one read of its 40-byte body
    wb = 36 * n + 40 * ncells
    achieved = wb / (ms_walk * 1e-3) / 1e9
    line = {
        "metric": "d3 Barnes-Hut charge passes/s @1M nodes (tree build + bodies + walk), FP64",
        "value": round(1e3 / ms_pass, 3), "unit": "passes/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(ms_pass, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"d3 many-body pass, {n} nodes at the golden-spiral seed "
                               "positions, theta 0.9, charge -90, distanceMin 1, "
                               "chargeDistanceMax 520, alpha 1", "nodes": n, "cells": ncells},
        "walk_ms": round(ms_walk, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": None, "kernel": "k_d3_walk",
                     "algorithmic_bytes": wb},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        info = cpu_info()
        core = pin_one_core()
        t0 = time.perf_counter()
        want = O.d3_charge_bh(xy, **kw)
        el = time.perf_counter() - t0
        line["bit_exact_vs_oracle"] = bool(np.array_equal(got.view(np.uint64),
                                                          want.view(np.uint64)))
        line["cpu_baseline"] = {"value": round(1.0 / el, 4), "unit": "passes/s", "cores": 1,
                                "kind": "port",
                                "sample": f"one whole d3 pass over the same {n} positions "
                                          "(oracle/fa2_oracle.c: insertion quadtree, bodies, "
                                          f"walk; gcc -O2, one thread pinned to core {core})",
                                **info}
    print(json.dumps(line), flush=True)


def bench_ba2000(args):
    """The reference's CPU-runnable configuration: preferential attachment, 2000 nodes, m 5,
    seed 0, starts U[-100,100]^2, theta 1.2, FP64 with the reference's node-order global sums,
    100 iterations (the whole layout). Layouts are timed on fresh sessions (graph upload and
    the iteration's graph capture included), then 100 more iterations of one resident session;
    the device's 100-iteration positions compared bit for bit with the oracle's, whose run of
    the same 100 iterations on one core is the CPU rate."""
    from graph_render_amd import synth
    import graph_render_amd as G
    w = synth.workload("ba2000")
    n, iters = w["n"], w["iterations"]
    topo = G.Topology(n, w["source"], w["target"])
    params = G.Fa2Params(dim=2, seed=w["seed"], max_iter=iters)
    ext = G.Fa2Ext(theta=w["theta"], precision="fp64", start=w["start"])
    layouts, final = [], None
    for r in range(max(args.warmup, 1) + max(args.steps // 10, 5)):
        t0 = time.perf_counter()
        with G.Session(topo, params, ext) as s:
            done = s.step(iters)
            final = s.positions()
        el = time.perf_counter() - t0
        assert done == iters, (done, iters)
        if r >= max(args.warmup, 1):
            layouts.append(el)
    layout_s = float(np.median(layouts))
    with G.Session(topo, params, ext) as s:
        s.step(iters)
        t0 = time.perf_counter()
        more = s.step(iters)
        steady = time.perf_counter() - t0
        walk_ms = s.time("walk", 20)
        ncells = s.tree(bounds=False)["ncells"]
    wb = walk_bytes(n, ncells, "fp64")
    achieved = wb / (walk_ms * 1e-3) / 1e9
    print(f"ba2000: layout {layout_s * 1e3:.2f} ms, walk {walk_ms * 1e3:.1f} us, {ncells} cells",
          file=sys.stderr, flush=True)
    line = {
        "metric": "FA2-BH layouts/s @BA(2000, m 5), 100 iterations, FP64 node-order sums",
        "value": round(1.0 / layout_s, 3), "unit": "layouts/s", "n_gpus": 1,
        "steps": len(layouts), "warmup": max(args.warmup, 1),
        "ms_per_step": round(layout_s * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "preferential attachment, 2000 nodes, m 5, seed 0, starts "
                               "U[-100,100]^2 (seed 0), theta 1.2, FP64, 100 iterations",
                   "nodes": n, "raw_edges": int(len(w["source"]))},
        "iters_per_s_in_layout": round(iters / layout_s, 1),
        "steady_iters_per_s": round(more / steady, 1),
        "walk_ms": round(walk_ms, 5), "tree_cells": ncells,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": None, "kernel": "k_walk<double,2>", "algorithmic_bytes": wb},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        info = cpu_info()
        core = pin_one_core()
        t0 = time.perf_counter()
        ref = O.Fa2(n, w["source"], w["target"], dim=2, seed=w["seed"], start=w["start"],
                    weights=None, tree=True, theta2=w["theta"] ** 2)
        ran = ref.run(iters)
        el = time.perf_counter() - t0
        line["bit_exact_vs_oracle"] = bool(ran == iters and np.array_equal(
            final.view(np.uint64), ref.positions().view(np.uint64)))
        line["cpu_baseline"] = {"value": round(1.0 / el, 4), "unit": "layouts/s", "cores": 1,
                                "kind": "port",
                                "sample": "the whole 100-iteration layout of the same graph and "
                                          "start (oracle/fa2_oracle.c, gcc -O2, one thread "
                                          f"pinned to core {core}, graph setup included)",
                                **info}
    print(json.dumps(line), flush=True)


def bench_d3sim(args):
    """layout.force.barnes_hut, the frozen d3 force simulation, on the device: the
    reference's gate model at 100k nodes (FORCE_CEILING, registry/force.rs:31), 112 ticks of
    link, many-body over the quadtree, center, collide over the quadtree of the projected
    positions and integration, FP64 in the reference's orders. Layouts/s with the graph
    uploaded (the tick loop timed with HIP events); the charge walk's roofline position; the
    state after a sample of ticks compared bit for bit with the oracle's, whose run of the same
    ticks on one core gives the CPU rate (scaled to 112 ticks)."""
    from graph_render_amd import barnes_hut as BH
    from graph_render_amd import synth
    import graph_render_amd as G
    n = args.nodes or 100_000
    s, t, w = synth.gate_model(n)
    top = G.Topology(n, s, t)
    BH.BarnesHut.time_run(top, w, ticks=8)  # warm-up: code objects, allocator
    runs = []
    for _ in range(max(args.steps, 1)):
        geo, ms = BH.BarnesHut.time_run(top, w, ticks=BH.TICKS)
        runs.append(ms)
        print(f"d3sim n={n}: 112 ticks {ms:.2f} ms", file=sys.stderr, flush=True)
    ms = float(np.median(runs))
    # the collide walk, the longest kernel on the tick's critical path, timed again on the
    # final tick's tree; algorithmic bytes per launch: per node its tree-order position (16 B),
    # point index (4 B) and delta (16 B); per cell its skip/start/end (12 B) and square (32 B)
    with BH.ForceSession(top, w) as S:
        S.step(BH.TICKS)
        ms_charge, ms_collide, c_charge, c_collide = S.time_walks(10)
    wb = 36 * n + 44 * c_collide
    achieved = wb / (ms_collide * 1e-3) / 1e9
    print(f"d3sim walks: many-body {ms_charge:.3f} ms ({c_charge} cells), collide "
          f"{ms_collide:.3f} ms ({c_collide} cells)", file=sys.stderr, flush=True)
    # the reference's own stage time on this graph (gate model, 100k nodes, 154 978 edges):
    # 13 735.36 ms, one thread, one run (docs/measurements/perf-p2-pm.md:36)
    published_ms = 13735.36 if (n == 100_000 and len(s) == 154_978) else None
    line = {
        "metric": "layout.force.barnes_hut layouts/s @100k nodes (112 ticks), FP64",
        "value": round(1e3 / ms, 4), "unit": "layouts/s", "n_gpus": 1,
        "steps": len(runs), "warmup": 1, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(published_ms / ms, 2) if published_ms else None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"the gate model (synthetic.rs seeded_model, Mulberry32 "
                               f"0x051042) at {n} nodes, {len(s)} raw edges, golden-spiral "
                               "start, ForceParams defaults, 112 ticks", "nodes": n,
                   "raw_edges": int(len(s)), "ticks": BH.TICKS},
        "ms_per_tick": round(ms / BH.TICKS, 4),
        "walk_ms": {"many_body": round(ms_charge, 4), "collide": round(ms_collide, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": collide_traffic(n), "kernel": "k_sim_collide",
                     "algorithmic_bytes": wb},
        "finite": bool(np.isfinite(geo.x).all() and np.isfinite(geo.y).all()),
    }
    if published_ms:
        line["baseline"] = {"value_ms": published_ms, "source": "the reference's stage time, "
                            "one thread on a shared x86_64 host (perf-p2-pm.md:36)"}
    if not args.no_cpu:
        from oracle import oracle as O
        k = 12  # about 6 s of oracle work at 100k nodes
        with BH.ForceSession(top, w) as S:
            S.step(k)
            dev = S.state()
        info = cpu_info()
        core = pin_one_core()
        t0 = time.perf_counter()
        want = O.d3sim(n, s, t, w, ticks=k, velocities=True)
        el = time.perf_counter() - t0
        line["bit_exact_vs_oracle"] = all(
            bool(np.array_equal(a.view(np.uint64), b.view(np.uint64))) for a, b in zip(dev, want))
        line["cpu_baseline"] = {
            "value": round(1.0 / (el * BH.TICKS / k), 5), "unit": "layouts/s", "cores": 1,
            "kind": "port",
            "sample": f"the first {k} ticks of the same layout through oracle/fa2_oracle.c "
                      f"(gro_d3sim_run, {el:.1f}s, one thread pinned to core {core}), scaled "
                      "to 112 ticks; the device state after those ticks compared bit for bit",
            **info}
    print(json.dumps(line), flush=True)


def bench_pmsim(args):
    """layout.force.particle_mesh on the device: the gate model at 1M nodes (1 549 929 raw
    edges, the graph of the reference's own stage timing), 112 ticks of link, many-body on
    the side_for(n) mesh (frame, slot-ordered CIC deposit, radix-2 FP64 transforms, kernel
    spectrum, CIC read), center, collide on the hashed grid and integration, FP64 in the
    reference's orders. Layouts/s (the tick loop timed with HIP events) against the
    reference's published 68 483.65 ms per 1M layout; the state after a sample of ticks
    compared bit for bit with the oracle's, whose run of those ticks gives the CPU rate."""
    from graph_render_amd import barnes_hut as BH
    from graph_render_amd import synth
    import graph_render_amd as G
    n = args.nodes or 1_000_000
    s, t, w = synth.gate_model(n)
    top = G.Topology(n, s, t)
    BH.ParticleMesh.time_run(top, w, ticks=8)  # warm-up: code objects, allocator
    runs = []
    for _ in range(max(args.steps, 1)):
        geo, ms = BH.ParticleMesh.time_run(top, w, ticks=BH.TICKS)
        runs.append(ms)
        print(f"pmsim n={n}: 112 ticks {ms:.2f} ms", file=sys.stderr, flush=True)
    ms = float(np.median(runs))
    side = BH.pm_side_for(n)
    # the tick's longest kernel, the node-order centring sums (two dependent chains of n FP64
    # adds, on a side stream beside link and charge; the tick waits for them), and the main
    # stream's longest, the hashed grid's pushes, timed again after the layout's last tick.
    # Algorithmic bytes: the sums read x and y (16 B per node); the pushes read a node's
    # projected position (16 B), its grid slot (4 B) and its three rows' bucket bounds (24 B)
    # and write its delta (16 B): 60 B per node
    with BH.ForceSession(top, w) as S:
        S.with_particle_mesh()
        S.step(BH.TICKS)
        ms_center, ms_gather, _, buckets = S.time_walks(10)
    cb, gb = 16 * n, 60 * n
    ach_c = cb / (ms_center * 1e-3) / 1e9
    ach_g = gb / (ms_gather * 1e-3) / 1e9
    print(f"pmsim kernels: centring sums {ms_center:.3f} ms, hashed-grid pushes "
          f"{ms_gather:.3f} ms ({buckets} buckets)", file=sys.stderr, flush=True)
    # the reference's stage time on this graph: 68 483.65 ms, median of 3, one thread
    # (docs/measurements/perf-p2-pm.md:39)
    published_ms = 68483.65 if (n == 1_000_000 and len(s) == 1_549_929) else None
    line = {
        "metric": "layout.force.particle_mesh layouts/s @1M nodes (112 ticks), FP64",
        "value": round(1e3 / ms, 4), "unit": "layouts/s", "n_gpus": 1,
        "steps": len(runs), "warmup": 1, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak",
        "vs_baseline": round(published_ms / ms, 2) if published_ms else None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"the gate model (synthetic.rs seeded_model, Mulberry32 "
                               f"0x051042) at {n} nodes, {len(s)} raw edges, golden-spiral "
                               f"start, ForceParams defaults, {side}x{side} mesh, 112 ticks",
                   "nodes": n, "raw_edges": int(len(s)), "ticks": BH.TICKS, "mesh_side": side},
        "ms_per_tick": round(ms / BH.TICKS, 4),
        "kernel_ms": {"centring_sums": round(ms_center, 4), "grid_pushes": round(ms_gather, 4)},
        "roofline": {"bound": "hbm", "achieved": round(ach_c, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach_c / HBM_PEAK_GBS, 6),
                     "traffic": profile_figure("traffic_bytes_center_sum_1M", n,
                                               "pmsim_1M_kernel_stats.md", 1_000_000),
                     "kernel": "k_sim_center_sum", "algorithmic_bytes": cb,
                     "note": "two chains of n dependent FP64 adds in node order (the "
                             "reference's rounding): bound by the add latency, not by bytes"},
        "roofline_grid_pushes": {"bound": "hbm", "achieved": round(ach_g, 2),
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(ach_g / HBM_PEAK_GBS, 5),
                                 "traffic": profile_figure("traffic_bytes_grid_pushes_1M", n,
                                                           "pmsim_1M_kernel_stats.md",
                                                           1_000_000),
                                 "kernel": "k_pmd3_gather", "algorithmic_bytes": gb},
        "finite": bool(np.isfinite(geo.x).all() and np.isfinite(geo.y).all()),
    }
    if published_ms:
        line["baseline"] = {"value_ms": published_ms, "source": "the reference's stage time, "
                            "median of 3, one thread on a shared x86_64 host "
                            "(perf-p2-pm.md:39)"}
    if not args.no_cpu:
        from oracle import oracle as O
        k = 4  # about 2 s of oracle work at 1M nodes
        with BH.ForceSession(top, w) as S:
            S.with_particle_mesh()
            S.step(k)
            dev = S.state()
        info = cpu_info()
        core = pin_one_core()
        t0 = time.perf_counter()
        want = O.d3sim(n, s, t, w, ticks=k, velocities=True, mesh_side=side)
        el = time.perf_counter() - t0
        line["bit_exact_vs_oracle"] = all(
            bool(np.array_equal(a.view(np.uint64), b.view(np.uint64))) for a, b in zip(dev, want))
        line["cpu_baseline"] = {
            "value": round(1.0 / (el * BH.TICKS / k), 6), "unit": "layouts/s", "cores": 1,
            "kind": "port",
            "sample": f"the first {k} ticks of the same layout through oracle/pm_d3_oracle.c "
                      f"(gro_d3sim_run on the mesh, {el:.1f}s with setup, one thread pinned to "
                      f"core {core}), scaled to 112 ticks; the device state after those ticks "
                      "compared bit for bit",
            **info}
    print(json.dumps(line), flush=True)


def bench_forcesim(args):
    """layout.forceatlas2.forcesim over a batch of small graphs, one launch (a workgroup per
    graph): graphs with n uniform in [10, 400] (the reference's DIRECT_MAX) and random edges
    at mean degree 4, the stage's defaults (50 iterations) with a seed per graph. Layouts/s
    with the inputs resident; every 2nd graph is checked bit for bit against the oracle and
    timed there on one core, the CPU rate scaled by the batch's pair work (sum of n^2)."""
    import graph_render_amd as G
    from graph_render_amd import forcesim as FS
    ng = args.graphs if args.graphs != 4096 else 1024
    rng = np.random.default_rng(5)
    topos, params = [], []
    for b in range(ng):
        n = int(rng.integers(10, 401))
        m = 2 * n
        topos.append(G.Topology(n, rng.integers(0, n, m).astype(np.uint32),
                                rng.integers(0, n, m).astype(np.uint32)))
        params.append(FS.Fa2ForceSimParams(seed=1_767_573_729 + b))
    res, ms = FS.time_batch(topos, params, reps=max(args.steps, 1))
    nodes = np.array([t.n for t in topos], np.float64)
    pairs = float((nodes ** 2).sum()) * 50
    # FP64 work per pair and iteration: the coefficient (2 mul, 1 div), three products and
    # three sums, one pairwise add
    flops = 10.0 * pairs
    achieved = flops / (ms * 1e-3) / 1e12
    print(f"forcesim batch {ng} graphs: {ms:.3f} ms", file=sys.stderr, flush=True)
    line = {
        "metric": "layout.forceatlas2.forcesim layouts/s, batch of small graphs (50 iterations)",
        "value": round(ng / (ms * 1e-3), 2), "unit": "layouts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{ng} random graphs, n uniform 10..400, mean degree 4, "
                               "forcesim defaults (50 iterations, seed per graph)",
                   "graphs": ng, "nodes": int(nodes.sum()), "pairs_x_iterations": pairs},
        "roofline": {"bound": "fp64", "achieved": round(achieved, 3), "peak": 78.6,
                     "unit": "TFLOP/s", "frac": round(achieved / 78.6, 5), "traffic": None,
                     "kernel": "k_fs_layout", "algorithmic_flops": flops},
        "nonfinite_graphs": int(sum(st != 0 for _, st in res)),
    }
    if not args.no_cpu:
        from oracle import oracle as O
        info = cpu_info()
        core = pin_one_core()
        sample = list(range(0, ng, 2))  # about 5 s of CPU work at the default 1024 graphs
        same = True
        t0 = time.perf_counter()
        for b in sample:
            t = topos[b]
            want, _, _ = O.forcesim(t.n, t.source, t.target, seed=params[b].seed)
            same &= bool(np.array_equal(want.view(np.uint32), res[b][0].view(np.uint32)))
        el = time.perf_counter() - t0
        sw = float((nodes[sample] ** 2).sum())
        line["bit_exact_sample"] = same
        line["cpu_baseline"] = {
            "value": round(len(sample) / el * (sw / len(sample)) / ((nodes ** 2).sum() / ng), 3),
            "unit": "layouts/s", "cores": 1, "kind": "port",
            "sample": f"every 2nd graph ({len(sample)} graphs, {el:.1f}s) through "
                      "oracle/forcesim_oracle.c on one core (core "
                      f"{core}), scaled to the batch's mean pair work", **info}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nodes", type=int, default=None)
    ap.add_argument("--fp64-steps", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the FP32 headline (for profiler runs)")
    ap.add_argument("--workload", choices=["sbm1m", "batch", "precision", "adjoint",
                                           "autotune", "d3", "forcesim", "d3sim", "pmsim",
                                           "ba2000"],
                    default="sbm1m")
    ap.add_argument("--graphs", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "precision":
        bench_precision(args)
        return
    if args.workload == "adjoint":
        bench_adjoint(args)
        return
    if args.workload == "autotune":
        bench_autotune(args)
        return
    if args.workload == "ba2000":
        bench_ba2000(args)
        return
    if args.workload == "d3":
        bench_d3(args)
        return
    if args.workload == "d3sim":
        bench_d3sim(args)
        return
    if args.workload == "pmsim":
        bench_pmsim(args)
        return
    if args.workload == "forcesim":
        bench_forcesim(args)
        return

    import graph_render_amd as G
    from graph_render_amd import synth

    w = synth.workload("sbm1m")
    if args.nodes:  # a smaller graph of the same family, for trials
        s_, t_, w_ = synth.sbm_zipf(args.nodes, 64, 1.1, 10.0, seed=1)
        w.update(n=args.nodes, source=s_, target=t_, weights=w_,
                 start=synth.uniform_start(args.nodes, 2, seed=1))
    n, s, t = w["n"], w["source"], w["target"]
    topo = G.Topology(n, s, t)
    params = G.Fa2Params(max_iter=500, scaling_ratio=2.0, gravity=1.0, jitter_tolerance=1.0)

    def make(precision, **ext):
        return G.Session(topo, params, G.Fa2Ext(theta=w["theta"], precision=precision,
                                                 start=w["start"], edge_weights=w["weights"],
                                                 **ext))

    # FP32 (the headline): W untimed iterations, then exactly K timed ones
    s32 = make("fp32")
    m_simple = len(s32.graph()[0])
    s32.step(args.warmup)
    t0 = time.perf_counter()
    done = s32.step(args.steps)
    el = time.perf_counter() - t0
    assert done == args.steps, (done, args.steps)
    fp32_its = done / el
    # the walk (the dominant kernel) and the tree build, timed with HIP events on the
    # session's stream, and the walk's interactions from its counting variant
    walk_ms = s32.time("walk", 20)
    tree_ms = s32.time("tree", 10)
    ws = s32.walk_stats()
    ncells = s32.tree(bounds=False)["ncells"]
    wb = walk_bytes(n, ncells, "fp32")
    achieved = wb / (walk_ms * 1e-3) / 1e9
    ib = iteration_bytes(n, ncells, m_simple)
    pos32 = s32.positions()
    f32 = s32.forces(7)
    s32.close()
    if args.headline_only:
        print(json.dumps({"metric": METRIC, "value": round(fp32_its, 3), "unit": "iters/s",
                          "steps": args.steps, "walk_ms": round(walk_ms, 4),
                          "tree_build_ms": round(tree_ms, 4), "tree_cells": int(ncells),
                          "walk": ws}), flush=True)
        return

    # the whole 500-iteration run from the start, in blocks of 100
    full = make("fp32")
    blocks = []
    for _ in range(5):
        t0 = time.perf_counter()
        d = full.step(100)
        blocks.append(d / (time.perf_counter() - t0))
    full.close()

    # FP32 with compensated accumulation, same start and steps
    sk = make("fp32_kahan")
    sk.step(args.warmup)
    t0 = time.perf_counter()
    dk = sk.step(args.steps)
    kahan_its = dk / (time.perf_counter() - t0)
    sk.set_positions(pos32)
    fk = sk.forces(7)
    sk.close()

    # FP64: with block-reduced global sums (fp64_iters_per_s) and with the reference's own
    # node-order sums (bit-exact trajectories), each over fp64_steps iterations after two
    # untimed
    def fp64_rate(**ext):
        s64 = make("fp64", **ext)
        s64.step(2)
        t0 = time.perf_counter()
        d64 = s64.step(args.fp64_steps)
        r = d64 / (time.perf_counter() - t0)
        return s64, r
    s64, fp64_its = fp64_rate()
    s64.set_positions(pos32)
    f64 = s64.forces(7)
    f64_parts = [s64.forces(1), s64.forces(2), s64.forces(4)]  # attraction, repulsion, gravity
    s64.close()
    s64b, fp64_block_its = fp64_rate(block_sums=True)
    s64b.close()
    node_err, field_err = errors(f32, f64)
    knode_err, kfield_err = errors(fk, f64)

    # particle-mesh variant on a 2048^2 mesh, same graph and start, FP32
    pm = make("fp32", repulsion="pm", pm_grid=2048)
    pm.step(2)
    t0 = time.perf_counter()
    dpm = pm.step(args.steps)
    pm_its = dpm / (time.perf_counter() - t0)
    pm_ms = pm.time("pm", 5)
    pm.set_positions(pos32)
    fpm = pm.forces(2)
    pm.close()
    s64r = make("fp64")
    s64r.set_positions(pos32)
    f64r = s64r.forces(2)
    s64r.close()
    pm_node, pm_field = errors(fpm, f64r)
    # the exact pair sum (device FP64 dense, O(n^2)) at the same positions: how far the
    # Barnes-Hut approximation at theta 1.2 and the 2048^2 mesh each are from it
    ex = make("fp64", repulsion="dense")
    ex.set_positions(pos32)
    t0 = time.perf_counter()
    fex = ex.forces(2)
    exact_s = time.perf_counter() - t0
    ex.close()
    pm_ex_node, pm_ex_field = errors(fpm, fex)
    bh_ex_node, bh_ex_field = errors(f64r, fex)
    # the FA2 law reaches the whole frame, so a side-P transform resolves the layout on
    # P/2 cells per axis (the rest is the zero padding of the free-space convolution): the
    # 4096^2 transform resolves it on the 2048 cells the 2048^2 label suggests
    pm4 = make("fp32", repulsion="pm", pm_grid=4096)
    pm4.step(2)
    t0 = time.perf_counter()
    dpm4 = pm4.step(args.steps)
    pm4_its = dpm4 / (time.perf_counter() - t0)
    pm4_ms = pm4.time("pm", 5)
    pm4.set_positions(pos32)
    fpm4 = pm4.forces(2)
    pm4.close()
    pm4_ex_node, pm4_ex_field = errors(fpm4, fex)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(w, params, args.cpu_seconds)

    line = {
        "metric": METRIC,
        "value": round(fp32_its, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / fp32_its, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "FA2-BH, 2D, SBM 1M nodes / 64 Zipf(1.1) blocks / mean degree 10, "
                        "lognormal(0,0.5) weights, seed 1, theta 1.2, FP32 forces + FP64 "
                        "reductions",
            "nodes": n, "raw_edges": int(s.size), "simple_edges": int(m_simple),
            "edges_sha256": synth.digest(s, t),
        },
        "full_run_500": {"iters_per_s_per_100": [round(b, 1) for b in blocks],
                         "iters_per_s": round(500 / sum(100 / b for b in blocks), 3)},
        "fp64_iters_per_s": round(fp64_block_its, 3),
        "fp64_node_order_sums_iters_per_s": round(fp64_its, 3),
        "fp64_steps": args.fp64_steps,
        "force_err_fp32_vs_fp64": {"per_node_rms": node_err, "whole_field": field_err,
                                   "max_per_node": max_rel(f32, f64),
                                   **conditioned(f32, f64, f64_parts)},
        "fp32_kahan_iters_per_s": round(kahan_its, 3),
        "force_err_fp32_kahan_vs_fp64": {"per_node_rms": knode_err,
                                         "whole_field": kfield_err,
                                         "max_per_node": max_rel(fk, f64),
                                         **conditioned(fk, f64, f64_parts)},
        "pm_2048_iters_per_s": round(pm_its, 3),
        "pm_2048_repulsion_ms": round(pm_ms, 4),
        "pm_4096_iters_per_s": round(pm4_its, 3),
        "pm_4096_repulsion_ms": round(pm4_ms, 4),
        "pm_4096_vs_exact_fp64_repulsion_err": {"per_node_rms": pm4_ex_node,
                                                "whole_field": pm4_ex_field},
        "pm_vs_bh_fp64_repulsion_err": {"per_node_rms": pm_node, "whole_field": pm_field},
        "pm_vs_exact_fp64_repulsion_err": {"per_node_rms": pm_ex_node,
                                           "whole_field": pm_ex_field},
        "bh_theta_1_2_vs_exact_fp64_repulsion_err": {"per_node_rms": bh_ex_node,
                                                     "whole_field": bh_ex_field},
        "exact_fp64_dense_repulsion_s": round(exact_s, 3),
        "walk_ms": round(walk_ms, 4),
        "tree_build_ms": round(tree_ms, 4),
        "tree_cells": int(ncells),
        "walk": {"interactions": ws["accepted"] + ws["pairs"],
                 "interactions_per_s": round((ws["accepted"] + ws["pairs"]) /
                                             (walk_ms * 1e-3), 1),
                 "cells_accepted": ws["accepted"], "leaf_pairs": ws["pairs"],
                 "scan_steps": ws["steps"],
                 "steps_per_wave": round(ws["steps"] / max(1, -(-n // 64)), 1),
                 "acting_lane_fraction": round(ws["acting_lanes"] / max(1, 64 * ws["steps"]),
                                               4)},
        "iteration_roofline": {"algorithmic_bytes": int(ib),
                               "achieved": round(ib * fp32_its / 1e9, 2),
                               "unit": "GB/s",
                               "frac": round(ib * fp32_its / 1e9 / HBM_PEAK_GBS, 5)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": walk_traffic(n), "kernel": "k_walk<float,2>",
                     "algorithmic_bytes": int(wb)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
