"""layout.force.barnes_hut on the device: the whole frozen d3 force simulation
(BarnesHut::run, gr_d3_layout_run), its many-body pass alone (gr_d3_charge_bh: the velocity
each node gains) and the cells' insertion node ids that pass's jiggle is keyed on
(gr_quadtree_node_ids)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields
from typing import Optional

import numpy as np

from . import _abi as A
from .fa2 import Geometry, ParamError, Topology, _check, _raise

# params.rs: 0.94^k < 0.001 first holds at k = 112
TICKS = 112


@dataclass
class ForceParams:
    """ForceParams (layout/force/params.rs:26-77), the frozen force set."""

    charge_strength: float = -90.0
    theta: float = 0.9
    distance_min: float = 1.0
    distance_max: float = 520.0
    link_distance: float = 60.0
    link_strength_scale: float = 0.15
    collide_radius: float = 16.0
    center_strength: float = 1.0
    velocity_decay: float = 1.0 - 0.42
    alpha_decay: float = 0.06
    alpha_min: float = 0.001
    initial_alpha: float = 1.0

    def _c(self) -> A.GrD3Params:
        out = A.GrD3Params()
        for f in fields(self):
            setattr(out, f.name, float(getattr(self, f.name)))
        return out


def _strength(topology: Topology, strength):
    if strength is None:
        return None
    s = np.ascontiguousarray(strength, np.float64).reshape(-1)
    if s.size != topology.m:
        raise ParamError(A.GR_ERR_PARAM, "strength", "one strength per raw edge")
    return s


class ForceSession:
    """ForceSession (layout/force/session.rs:118-270) resident in HBM: step(ticks) runs the
    ticks on the device; step(112), 112 x step(1) and 7 x step(16) give the same bytes."""

    def __init__(self, topology: Topology, strength=None,
                 params: Optional[ForceParams] = None, particle_mesh: bool = False):
        self.n = int(topology.n)
        s = _strength(topology, strength)
        cp = (params or ForceParams())._c()
        m = int(topology.m)
        src = topology.source if m else np.zeros(1, np.uint32)
        dst = topology.target if m else np.zeros(1, np.uint32)
        self._h = A.lib().gr_d3_session_create(self.n, src.ctypes.data, dst.ctypes.data,
                                               s.ctypes.data if s is not None and m else None,
                                               m, C.byref(cp))
        if not self._h:
            _raise()
        if particle_mesh:
            self.with_particle_mesh()

    def with_particle_mesh(self):
        """session.rs with_particle_mesh: later ticks run the many-body pass on the mesh and
        collision on the hashed grid."""
        _check(A.lib().gr_d3_session_with_particle_mesh(self._h))
        return self

    def close(self):
        if self._h:
            A.lib().gr_d3_session_release(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, ticks: int):
        """StepReport: (alpha, settled, ticks_run)."""
        a = C.c_double(0.0)
        st = C.c_int32(0)
        _check(A.lib().gr_d3_session_step(self._h, int(ticks), C.byref(a), C.byref(st)))
        return float(a.value), bool(st.value), int(ticks)

    def state(self):
        """(x, y, vx, vy) f64 columns in node order."""
        cols = [np.zeros(max(self.n, 1)) for _ in range(4)]
        _check(A.lib().gr_d3_session_state(self._h, *[c.ctypes.data for c in cols]))
        return tuple(c[: self.n] for c in cols)

    def xs(self):
        return self.state()[0]

    def ys(self):
        return self.state()[1]

    def time_walks(self, reps: int = 10):
        """The last tick's many-body and collide walks timed again on their own trees:
        (many-body ms, collide ms, many-body cells, collide cells). On a mesh session: (the
        node-order centring sums' ms, the hashed grid's pushes' ms, mesh cells, grid
        buckets)."""
        ms = np.zeros(2)
        cells = np.zeros(2, np.uint32)
        _check(A.lib().gr_d3_session_time(self._h, int(reps), ms.ctypes.data, cells.ctypes.data))
        return float(ms[0]), float(ms[1]), int(cells[0]), int(cells[1])

    @property
    def alpha(self) -> float:
        return float(A.lib().gr_d3_session_alpha(self._h))

    def reheat(self, alpha: float):
        _check(A.lib().gr_d3_session_reheat(self._h, float(alpha)))

    def set_alpha_target(self, target: float):
        _check(A.lib().gr_d3_session_set_alpha_target(self._h, float(target)))


class BarnesHut:
    """layout/force/barnes_hut.rs:139-200: the simple graph with its strengths, the golden
    spiral start (seed.rs), TICKS ticks of link, many-body, center, collide and integrate
    (sim.rs), x and y narrowed to f32 (NonFinite{node.x} refused)."""

    ID = "layout.force.barnes_hut"

    RUN = "gr_d3_layout_run"

    @classmethod
    def _call(cls, topology: Topology, strength, params, ticks, ms):
        n, m = int(topology.n), int(topology.m)
        s = _strength(topology, strength)
        cp = (params or ForceParams())._c()
        xs = np.zeros(max(n, 1), np.float32)
        ys = np.zeros(max(n, 1), np.float32)
        src = topology.source if m else np.zeros(1, np.uint32)
        dst = topology.target if m else np.zeros(1, np.uint32)
        _check(getattr(A.lib(), cls.RUN)(n, src.ctypes.data, dst.ctypes.data,
                                         s.ctypes.data if s is not None and m else None, m,
                                         C.byref(cp), int(ticks), xs.ctypes.data,
                                         ys.ctypes.data, ms))
        return Geometry(2, xs[:n], ys[:n], None)

    @classmethod
    def run(cls, topology: Topology, strength=None, params: Optional[ForceParams] = None,
            ticks: int = TICKS) -> Geometry:
        """strength: one per raw edge (None = 1.0 each); a simple edge keeps its first raw
        edge's."""
        return cls._call(topology, strength, params, ticks, None)

    @classmethod
    def time_run(cls, topology: Topology, strength=None, params: Optional[ForceParams] = None,
                 ticks: int = TICKS):
        """run, with the tick loop's device time in ms (HIP events around the ticks)."""
        ms = C.c_double(0.0)
        g = cls._call(topology, strength, params, ticks, C.byref(ms))
        return g, ms.value


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


def collide_pass(xy, radius: float = 16.0, seed: int = 0, tick: int = 0) -> np.ndarray:
    """collide.rs:37-127 over these positions: deltas (n, 2) in node order, every node's push
    from the others closer than 2 * radius (params.rs: collideRadius 16)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    _check(A.lib().gr_d3_collide(n, xy.ctypes.data if n else None, float(radius),
                                 int(seed) & 0xFFFFFFFF, int(tick) & 0xFFFFFFFF,
                                 out.ctypes.data))
    return out[: 2 * n].reshape(n, 2)


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


class ParticleMesh(BarnesHut):
    """layout/force/particle_mesh.rs:96-131: BarnesHut::run's session with the many-body pass
    on a side_for(n) mesh (d3's law) and collision on the hashed grid."""

    ID = "layout.force.particle_mesh"
    RUN = "gr_pm_layout_run"


def pm_charge_pass(xy, charge: float = -90.0, alpha: float = 1.0, distance_min: float = 1.0,
                   distance_max: float = 520.0):
    """particle_mesh/charge.rs from rest on a fresh mesh over these positions: (deltas (n, 2)
    node order, frame dict or None)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    fr = np.zeros(7)
    _check(A.lib().gr_pm_d3_charge(n, xy.ctypes.data if n else None, float(charge),
                                   float(alpha), float(distance_min), float(distance_max),
                                   out.ctypes.data, fr.ctypes.data))
    frame = None if fr[6] == 0 else dict(step=int(fr[0]), h=float(fr[1]), ox=float(fr[2]),
                                         oy=float(fr[3]), cells=int(fr[4]), reach=int(fr[5]),
                                         side=int(fr[6]))
    return out[: 2 * n].reshape(n, 2), frame


def pm_collide_pass(xy, radius: float = 16.0, seed: int = 0, tick: int = 0):
    """particle_mesh/collide.rs over these positions (the hashed grid): (deltas (n, 2) node
    order, the grid order)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    order = np.zeros(max(n, 1), np.uint32)
    _check(A.lib().gr_pm_d3_collide(n, xy.ctypes.data if n else None, float(radius),
                                    int(seed) & 0xFFFFFFFF, int(tick) & 0xFFFFFFFF,
                                    out.ctypes.data, order.ctypes.data))
    return out[: 2 * n].reshape(n, 2), order[:n]


def pm_side_for(n: int) -> int:
    """mesh.rs:31-34: ceil(sqrt n) rounded up to a power of two, held to 128..1024."""
    return int(A.lib().gr_pm_side_for(int(n)))

