"""torch.autograd wrapper around the device adjoint, and the stress loss it is checked with.
Test and measurement harness: the product is the C ABI's record/adjoint pair; this module
only drives it.

`Fa2Unrolled.apply(weights, scaling_ratio, gravity, spec)` runs `spec.iters` FA2 iterations
on the device (gr_fa2_session_record) and returns the final positions; backward() is the
hand-written reverse sweep (gr_fa2_session_adjoint), so a torch loss of the positions gives
gradients for the raw edge weights, the scaling ratio and gravity.

`stress1` is Kruskal stress-1 at the optimal uniform scale against BFS hop distances from
evenly spaced sources, as graph-cli measures layout quality (crates/graph-cli/src/
bench.rs:255-294): sum (d - a e)^2 / sum d^2 with a = sum d e / sum e^2, square-rooted,
unreachable pairs and the source itself skipped, 1 for a collapsed layout or a graph with
no reachable pair.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from graph_render_amd.fa2 import Fa2Ext, Fa2Params, Session, Topology


@dataclass
class UnrolledSpec:
    topology: Topology
    iters: int = 200
    checkpoint_every: int = 20
    precision: str = "fp32"
    repulsion: str = "tree"
    theta: Optional[float] = None
    start: Optional[np.ndarray] = None
    seed: int = 0
    jitter_tolerance: float = 1.0
    applied: list = field(default_factory=list)  # iterations each forward applied


def run_forward(spec: UnrolledSpec, weights: np.ndarray, scaling_ratio: float,
                gravity: float, precision: Optional[str] = None, record: bool = False):
    """One forward run: a fresh session with these parameters, `spec.iters` iterations.
    Returns (session, final positions); the session is open (for adjoint) if record."""
    p = Fa2Params(max_iter=spec.iters, scaling_ratio=float(scaling_ratio),
                  gravity=float(gravity), seed=spec.seed,
                  jitter_tolerance=spec.jitter_tolerance)
    x = Fa2Ext(theta=spec.theta, precision=precision or spec.precision,
               repulsion=spec.repulsion, start=spec.start,
               edge_weights=np.asarray(weights, np.float64))
    s = Session(spec.topology, p, x)
    if record:
        spec.applied.append(s.record(spec.iters, spec.checkpoint_every))
    else:
        spec.applied.append(s.step(spec.iters))
    pos = s.positions()
    if not record:
        s.close()
        s = None
    return s, pos


def _torch():
    import torch
    return torch


def make_function():
    torch = _torch()

    class Fa2Unrolled(torch.autograd.Function):
        @staticmethod
        def forward(ctx, weights, scaling_ratio, gravity, spec):
            sess, pos = run_forward(spec, weights.detach().cpu().double().numpy(),
                                    float(scaling_ratio), float(gravity), record=True)
            ctx.sess = sess
            return torch.from_numpy(pos).to(weights.device)

        @staticmethod
        def backward(ctx, gpos):
            g = ctx.sess.adjoint(gpos.detach().cpu().double().numpy())
            ctx.sess.close()
            dev = gpos.device
            return (torch.from_numpy(g["weights"]).to(dev),
                    torch.tensor(g["scaling_ratio"], dtype=torch.float64, device=dev),
                    torch.tensor(g["gravity"], dtype=torch.float64, device=dev), None)

    return Fa2Unrolled


def bfs_hops(n: int, source, target, sources) -> np.ndarray:
    """Hop distances (int32, -1 where unreachable) from each of `sources` over the
    undirected graph."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import shortest_path
    s = np.asarray(source, np.int64)
    t = np.asarray(target, np.int64)
    a = coo_matrix((np.ones(s.size), (s, t)), shape=(n, n)).tocsr()
    d = shortest_path(a, method="D", directed=False, unweighted=True, indices=sources)
    out = np.where(np.isfinite(d), d, -1).astype(np.int32)
    return out


def stress_sources(n: int, count: int = 256) -> np.ndarray:
    """Evenly spaced BFS sources, the step_by(n / count) of graph-cli's stress."""
    return np.arange(0, n, max(n // count, 1), dtype=np.int64)


def stress1(pos, sources: np.ndarray, hops: np.ndarray):
    """Kruskal stress-1 (torch, differentiable in pos): hops[s, j] is the hop distance from
    sources[s] to j, -1 when unreachable."""
    torch = _torch()
    h = torch.as_tensor(hops, device=pos.device)
    mask = h > 0
    dist = h.to(pos.dtype)
    diff = pos[None, :, :] - pos[torch.as_tensor(sources, device=pos.device)][:, None, :]
    sq = (diff * diff).sum(-1)
    e = torch.sqrt(torch.where(mask, sq, torch.ones_like(sq)))
    e = torch.where(mask, e, torch.zeros_like(e))
    de = (dist * e)[mask].sum()
    ee = (e * e)[mask].sum()
    dd = (dist * dist)[mask].sum()
    if float(ee.detach()) == 0.0 or float(dd.detach()) == 0.0:  # collapsed or no pair
        return (pos * 0.0).sum() + 1.0
    return torch.sqrt(torch.clamp((dd - de / ee * de) / dd, min=0.0))


def stress1_np(pos: np.ndarray, sources: np.ndarray, hops: np.ndarray) -> float:
    """stress1 in numpy float64 (for finite differences)."""
    mask = hops > 0
    diff = pos[None, :, :] - pos[sources][:, None, :]
    e = np.sqrt((diff * diff).sum(-1))
    dist = hops.astype(np.float64)
    de = (dist * e)[mask].sum()
    ee = (e * e)[mask].sum()
    dd = (dist * dist)[mask].sum()
    if ee == 0.0 or dd == 0.0:  # collapsed layout or no reachable pair
        return 1.0
    return float(np.sqrt(max((dd - de / ee * de) / dd, 0.0)))
