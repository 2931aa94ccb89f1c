"""d3's many-body force over the quadtree (layout.force's Barnes-Hut arm), on the device
through gr_d3_charge_bh: one charge pass, the velocity each node gains, and the cells'
insertion node ids its jiggle is keyed on (gr_quadtree_node_ids)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from .fa2 import _check


def charge_pass(xy, theta: float = 0.9, charge: float = -90.0, alpha: float = 1.0,
                distance_min: float = 1.0, distance_max: float = 520.0, seed: int = 0,
                tick: int = 0) -> np.ndarray:
    """barnes_hut/charge.rs:41-53: deltas (n, 2), NaN for a non-finite node. The defaults are
    layout.force's (params.rs: theta 0.9, chargeStrength -90, distanceMin 1,
    chargeDistanceMax 520) at alpha 1; seed and tick key the coincidence jiggle."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    _check(A.lib().gr_d3_charge_bh(n, xy.ctypes.data if n else None, float(theta),
                                   float(charge), float(alpha), float(distance_min),
                                   float(distance_max), int(seed) & 0xFFFFFFFF,
                                   int(tick) & 0xFFFFFFFF, out.ctypes.data))
    return out[: 2 * n].reshape(n, 2)


def time_pass(xy, reps: int = 5, theta: float = 0.9, charge: float = -90.0,
              alpha: float = 1.0, distance_min: float = 1.0, distance_max: float = 520.0,
              seed: int = 0, tick: int = 0):
    """charge_pass, then the pass timed on the device-resident positions: (deltas, whole pass
    ms, walk ms, cells); medians of three timings of `reps` runs."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    ms = np.zeros(3)
    _check(A.lib().gr_d3_charge_bh_time(n, xy.ctypes.data if n else None, float(theta),
                                        float(charge), float(alpha), float(distance_min),
                                        float(distance_max), int(seed) & 0xFFFFFFFF,
                                        int(tick) & 0xFFFFFFFF, int(reps), out.ctypes.data,
                                        ms.ctypes.data))
    return out[: 2 * n].reshape(n, 2), float(ms[0]), float(ms[1]), int(ms[2])


def node_ids(xy) -> np.ndarray:
    """Quadtree::node_id of every cell in preorder (quadtree.rs:64-88): the shape slot the
    reference's insertion pushed it as."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    if n == 0:
        return np.zeros(0, np.uint32)
    cells = C.c_uint32(0)
    ptr = xy.ctypes.data
    _check(A.lib().gr_quadtree_node_ids(n, ptr, 0, C.byref(cells), None))
    out = np.zeros(max(cells.value, 1), np.uint32)
    _check(A.lib().gr_quadtree_node_ids(n, ptr, cells.value, C.byref(cells), out.ctypes.data))
    return out[: cells.value]
