"""The particle mesh under the reference's own d3 many-body law (layout.force's PM arm),
on the device through gr_pm_charge: one charge pass, the velocity each node gains."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi as A
from .fa2 import _check


def side_for(n: int) -> int:
    """mesh.rs:31-34: ceil(sqrt n) rounded up to a power of two, held to 128..1024."""
    return int(A.lib().gr_pm_side_for(int(n)))


MAX_NODES_NOALLOC = 262144  # GR_NOALLOC_MAX_NODES
MAX_SIDE_NOALLOC = 1024     # GR_NOALLOC_MAX_SIDE


def charge_pass(xy, side: int = 0, charge: float = -90.0, alpha: float = 1.0,
                distance_min: float = 1.0, distance_max: float = 520.0,
                allocate: bool = True):
    """charge.rs:29-47: (deltas (n, 2), frame dict or None). The defaults are layout.force's
    (params.rs: chargeStrength -90, distanceMin 1, chargeDistanceMax 520) at alpha 1.
    allocate=False runs gr_pm_charge_noalloc: no device allocation, n <= 262144 and a mesh
    side <= 1024."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    fr = np.zeros(7)
    fn = A.lib().gr_pm_charge if allocate else A.lib().gr_pm_charge_noalloc
    _check(fn(n, xy.ctypes.data if n else None, int(side), float(charge), float(alpha),
              float(distance_min), float(distance_max), out.ctypes.data, fr.ctypes.data))
    frame = None
    if fr[6]:
        frame = dict(step=int(fr[0]), h=fr[1], ox=fr[2], oy=fr[3], cells=int(fr[4]),
                     reach=int(fr[5]))
    return out[: 2 * n].reshape(n, 2), frame
