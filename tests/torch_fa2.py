"""A plain torch (float64, CPU) restatement of dense FA2 iterations, differentiable by
torch.autograd: the FP64 reference gradient the device adjoint is checked against.

Follows the same iteration as the oracle's dense arm (oracle/fa2_oracle.c:gro_fa2_iterate,
after layout/forceatlas2/state.rs:130-148 and state/schedule.rs:16-77): weighted
attraction, dense repulsion, centroid gravity, cumulative swing/traction, networkx's
estimate_factor and the per-node step. Only the summation order differs from the oracle
(vectorised sums), so forward positions agree to rounding; test_adjoint_reference pins that.
"""
from __future__ import annotations

import math

import numpy as np
import torch


def simple_edges(source, target):
    """simple_graph (layout/force/mod.rs:101-120): self-loops and repeated unordered pairs
    dropped, the first raw edge kept. Returns lo, hi and each kept edge's raw index."""
    seen, lo, hi, raw = set(), [], [], []
    for e, (a, b) in enumerate(zip(np.asarray(source).tolist(), np.asarray(target).tolist())):
        if a == b:
            continue
        key = (min(a, b), max(a, b))
        if key in seen:
            continue
        seen.add(key)
        lo.append(key[0])
        hi.append(key[1])
        raw.append(e)
    return np.array(lo, np.int64), np.array(hi, np.int64), np.array(raw, np.int64)


def run(start, lo, hi, w, k, g, iters, jitter_tolerance=1.0):
    """Positions after `iters` dense FA2 iterations. start (n, dim), w (simple edges), k and
    g are tensors (any may require grad); masses are simple degree + 1."""
    n = start.shape[0]
    lo_t = torch.as_tensor(lo)
    hi_t = torch.as_tensor(hi)
    deg = torch.zeros(n, dtype=torch.float64)
    deg.index_add_(0, lo_t, torch.ones(len(lo), dtype=torch.float64))
    deg.index_add_(0, hi_t, torch.ones(len(hi), dtype=torch.float64))
    mass = deg + 1.0
    eye = torch.eye(n, dtype=torch.float64)
    p = start
    S = torch.tensor(1.0, dtype=torch.float64)
    T = torch.tensor(1.0, dtype=torch.float64)
    speed = torch.tensor(1.0, dtype=torch.float64)
    eff = 1.0
    nn = float(n)
    jt = float(jitter_tolerance)
    opt = 0.05 * math.sqrt(nn)
    minj = math.sqrt(opt)
    for _ in range(iters):
        d = p[hi_t] - p[lo_t]
        att = d * w[:, None]
        u = torch.zeros_like(p).index_add(0, lo_t, att).index_add(0, hi_t, -att)
        r = p[:, None, :] - p[None, :, :]
        d2 = (r * r).sum(-1) + eye
        f = mass[:, None] * mass[None, :] / d2 * (1.0 - eye) * k
        u = u + (f[..., None] * r).sum(1)
        c = p - p.mean(0)
        nc = torch.sqrt((c * c).sum(1, keepdim=True))
        u = u - g * mass[:, None] * c / nc
        sw = (mass * torch.sqrt(((p - u) ** 2).sum(1))).sum()
        tr = (0.5 * mass * torch.sqrt(((p + u) ** 2).sum(1))).sum()
        S = S + sw
        T = T + tr
        other = torch.clamp(opt * T / (nn * nn), max=10.0)
        jitter = jt * torch.clamp(other, min=minj)
        if float((S / T).detach()) > 2.0:
            if eff > 0.05:
                eff *= 0.5
            jitter = torch.clamp(jitter, min=jt)
        target = jitter * eff * T / S
        if float(S.detach()) > float((jitter * T).detach()):
            if eff > 0.05:
                eff *= 0.7
        elif float(speed.detach()) < 1000.0:
            eff *= 1.3
        speed = speed + torch.minimum(target - speed, 0.5 * speed)
        nu = torch.sqrt((u * u).sum(1))
        factor = speed / (1.0 + torch.sqrt(speed * mass * nu))
        p = p + u * factor[:, None]
    return p
