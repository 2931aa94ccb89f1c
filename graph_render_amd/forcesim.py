"""layout.forceatlas2.forcesim on the device (gr_forcesim_run / gr_forcesim_batch): SciGraphs'
own ForceSim, FA2 with direct repulsion in f32 state and f64 coefficient rows, as the
reference ports it (crates/graph-core/src/layout/force/forcesim.rs)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _abi as A
from .fa2 import Geometry, Topology, _check


@dataclass
class Fa2ForceSimParams:
    """forcesim.rs:104-153 and its Default."""

    iterations: int = 50
    scale: float = 5.0
    scaling_ratio: float = 2.0
    gravity: float = 1.0
    jitter_tolerance: float = 1.0
    seed: int = 1_767_573_729

    def _c(self) -> A.GrForceSimParams:
        return A.GrForceSimParams(int(self.iterations), float(self.scale),
                                  float(self.scaling_ratio), float(self.gravity),
                                  float(self.jitter_tolerance), int(self.seed))


class ForceAtlas2ForceSim:
    """Stage `layout.forceatlas2.forcesim` (forcesim.rs:155-163): x, y, z columns."""

    ID = "layout.forceatlas2.forcesim"

    @classmethod
    def run(cls, topology: Topology, params: Optional[Fa2ForceSimParams] = None) -> Geometry:
        n = int(topology.n)
        cols = [np.zeros(max(n, 1), np.float32) for _ in range(3)]
        p = (params or Fa2ForceSimParams())._c()
        src = topology.source if topology.m else np.zeros(1, np.uint32)
        dst = topology.target if topology.m else np.zeros(1, np.uint32)
        _check(A.lib().gr_forcesim_run(n, src.ctypes.data, dst.ctypes.data, topology.m,
                                       C.addressof(p), *[c.ctypes.data for c in cols]))
        return Geometry(3, cols[0][:n], cols[1][:n], cols[2][:n])


def _batch_args(topologies: Sequence[Topology], params):
    g = len(topologies)
    nodes = np.array([t.n for t in topologies] or [0], np.uint32)
    edges = np.array([t.m for t in topologies] or [0], np.uint64)
    src = np.concatenate([t.source for t in topologies] + [np.zeros(1, np.uint32)])
    dst = np.concatenate([t.target for t in topologies] + [np.zeros(1, np.uint32)])
    if params is None:
        params = [Fa2ForceSimParams()] * g
    elif isinstance(params, Fa2ForceSimParams):
        params = [params] * g
    arr = (A.GrForceSimParams * max(g, 1))(*[p._c() for p in params])
    total = int(nodes[:g].sum()) if g else 0
    cols = [np.zeros(max(total, 1), np.float32) for _ in range(3)]
    status = np.zeros(max(g, 1), np.int32)
    return g, nodes, edges, src, dst, arr, cols, status, total


def _split(topologies, cols, status, total):
    out, off = [], 0
    for b, t in enumerate(topologies):
        xyz = np.stack([c[off:off + t.n] for c in cols], 1)
        out.append((xyz, int(status[b])))
        off += t.n
    return out


def batch(topologies: Sequence[Topology], params=None) -> List[Tuple[np.ndarray, int]]:
    """Every graph laid out in one launch: [(xyz (n, 3) f32, status)] (status 2 = NonFinite)."""
    g, nodes, edges, src, dst, arr, cols, status, total = _batch_args(topologies, params)
    _check(A.lib().gr_forcesim_batch(g, nodes.ctypes.data, edges.ctypes.data, src.ctypes.data,
                                     dst.ctypes.data, C.addressof(arr),
                                     *[c.ctypes.data for c in cols], status.ctypes.data))
    return _split(topologies, cols, status, total)


def time_batch(topologies: Sequence[Topology], params=None, reps: int = 3):
    """batch(), then the whole batch timed from its resident inputs: (results, ms)."""
    g, nodes, edges, src, dst, arr, cols, status, total = _batch_args(topologies, params)
    ms = C.c_double(0.0)
    _check(A.lib().gr_forcesim_batch_time(g, nodes.ctypes.data, edges.ctypes.data,
                                          src.ctypes.data, dst.ctypes.data, C.addressof(arr),
                                          *[c.ctypes.data for c in cols], status.ctypes.data,
                                          int(reps), C.byref(ms)))
    return _split(topologies, cols, status, total), ms.value
