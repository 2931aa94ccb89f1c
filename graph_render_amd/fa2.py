"""Host-side mirror of graph-core's ForceAtlas2 interface over the HIP library.

Names, argument meaning and error behaviour follow the reference (paths relative to
crates/graph-core/src/):

    Fa2Params            layout/forceatlas2/state.rs:20-54 (+ Default)
    Stage / StageError   stage.rs:23-52
    ForceAtlas2          layout/forceatlas2.rs:32-41       id "layout.forceatlas2"
    ForceAtlas2BarnesHut layout/forceatlas2.rs:49-61       id "layout.forceatlas2.barnes_hut"
    run_3d               layout/forceatlas2.rs:63-68       id "layout.forceatlas2.3d"
    LAYOUTS / run_params registry/layouts.rs, registry/params.rs:144-202

Everything runs through libgraph_render_fa2.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass, field
from typing import Callable, Dict, Optional

import numpy as np

from . import _abi as A

# ---------------------------------------------------------------- errors (stage.rs:35-52)


class StageError(Exception):
    """A refused or failed stage run: (code, name, rule) as the library reports it."""

    def __init__(self, code: int, name: str = "", rule: str = ""):
        self.code, self.name, self.rule = code, name, rule
        super().__init__(f"[{code}] {name}: {rule}")


class CapacityError(StageError):
    pass


class NonFiniteError(StageError):
    pass


class ParamError(StageError):
    pass


class DeviceError(StageError):
    pass


_ERROR_TYPES = {
    A.GR_ERR_CAPACITY: CapacityError,
    A.GR_ERR_NON_FINITE: NonFiniteError,
    A.GR_ERR_PARAM: ParamError,
    A.GR_ERR_PARAM_OUT_OF_RANGE: ParamError,
    A.GR_ERR_PARAMS_MALFORMED: ParamError,
    A.GR_ERR_DEVICE: DeviceError,
}


def _raise(code: Optional[int] = None):
    c, name, rule = A.last_error()
    code = c if code is None else code
    raise _ERROR_TYPES.get(code, StageError)(code, name, rule)


def _check(code: int):
    if code != A.GR_OK:
        _raise(code)


# ---------------------------------------------------------------- parameters


@dataclass
class Fa2Params:
    """Fa2Params (state.rs:20-54); defaults are Fa2Params::default()."""

    max_iter: int = 100
    jitter_tolerance: float = 1.0
    scaling_ratio: float = 2.0
    gravity: float = 1.0
    seed: int = 0
    dim: int = 2

    # schema order (registry/tunable.rs:229-240); `dim` is not published
    WIRE_FIELDS = ("max_iter", "jitter_tolerance", "scaling_ratio", "gravity", "seed")

    def to_wire(self) -> bytes:
        return struct.pack("<5d", *(float(getattr(self, f)) for f in self.WIRE_FIELDS))

    @classmethod
    def from_wire(cls, data: bytes) -> "Fa2Params":
        out = A.GrFa2Params()
        buf = C.create_string_buffer(bytes(data), max(len(data), 1))
        _check(A.lib().gr_fa2_params_decode(buf, len(data), C.byref(out)))
        return cls(out.max_iter, out.jitter_tolerance, out.scaling_ratio, out.gravity,
                   out.seed, out.dim)

    def _c(self) -> A.GrFa2Params:
        return A.GrFa2Params(int(self.max_iter), float(self.jitter_tolerance),
                             float(self.scaling_ratio), float(self.gravity), int(self.seed),
                             int(self.dim))


PRECISIONS = {"fp64": A.GR_PREC_FP64, "fp32": A.GR_PREC_FP32,
              "fp32_kahan": A.GR_PREC_FP32_KAHAN}
REPULSIONS = {None: A.GR_REP_DEFAULT, "tree": A.GR_REP_TREE, "dense": A.GR_REP_DENSE,
              "pm": A.GR_REP_PM}


@dataclass
class Fa2Ext:
    """What the reference does not have. The defaults are the reference's semantics."""

    theta: Optional[float] = None          # None: the reference's 0.8 (barnes_hut.rs:24)
    precision: str = "fp64"
    repulsion: Optional[str] = None        # None: what the layout id runs
    edge_weights: Optional[np.ndarray] = None  # one per raw edge
    start: Optional[np.ndarray] = None     # (n, dim) start positions
    pm_grid: int = 0
    export_tree: bool = False
    direct_launch: bool = False            # True: no HIP-graph replay of iterations
    block_sums: bool = False               # FP64: block-reduced global sums (faster, not
                                           # the reference's rounding order)
    _keep: list = field(default_factory=list, repr=False)

    def _c(self, n: int, m: int, dim: int) -> A.GrFa2Ext:
        x = A.GrFa2Ext()
        x.struct_size = C.sizeof(A.GrFa2Ext)
        x.flags = ((A.GR_EXT_EXPORT_TREE if self.export_tree else 0) |
                   (A.GR_EXT_DIRECT_LAUNCH if self.direct_launch else 0) |
                   (A.GR_EXT_BLOCK_SUMS if self.block_sums else 0))
        x.theta = -1.0 if self.theta is None else float(self.theta)
        x.precision = PRECISIONS[self.precision]
        x.repulsion = REPULSIONS[self.repulsion]
        self._keep = []
        if self.edge_weights is not None:
            w = np.ascontiguousarray(self.edge_weights, np.float64).reshape(-1)
            if w.size != m:
                raise ParamError(A.GR_ERR_PARAM, "edge_weights", "one weight per raw edge")
            self._keep.append(w)
            x.edge_weights = w.ctypes.data
        if self.start is not None:
            s = np.ascontiguousarray(self.start, np.float64).reshape(-1)
            if s.size != n * dim:
                raise ParamError(A.GR_ERR_PARAM, "start", "n*dim start coordinates")
            self._keep.append(s)
            x.start = s.ctypes.data
        x.pm_grid = int(self.pm_grid)
        return x


# ---------------------------------------------------------------- topology / geometry


@dataclass
class Topology:
    """The part of graph-core's Topology the FA2 path reads: the node count and the raw
    edge columns in edge-index order (index.rs, edges().source/target)."""

    n: int
    source: np.ndarray
    target: np.ndarray

    def __post_init__(self):
        self.source = np.ascontiguousarray(self.source, np.uint32).reshape(-1)
        self.target = np.ascontiguousarray(self.target, np.uint32).reshape(-1)
        if self.source.size != self.target.size:
            raise ParamError(A.GR_ERR_PARAM, "edges", "equal source and target columns")

    @property
    def m(self) -> int:
        return int(self.source.size)


@dataclass
class Geometry:
    """Geometry::points (layout/mod.rs:116-122): f32 columns, z only at dim 3."""

    dim: int
    x: np.ndarray
    y: np.ndarray
    z: Optional[np.ndarray] = None


# ---------------------------------------------------------------- stages


def _run_id(layout_id: str, topology: Topology, wire: bytes, ext: Optional[Fa2Ext],
            dim: int) -> Geometry:
    n = int(topology.n)
    xs = np.zeros(max(n, 1), np.float32)
    ys = np.zeros(max(n, 1), np.float32)
    zs = np.zeros(max(n, 1), np.float32) if dim == 3 else None
    xc = ext._c(n, topology.m, dim) if ext is not None else None
    buf = C.create_string_buffer(bytes(wire), max(len(wire), 1))
    src = topology.source if topology.m else np.zeros(1, np.uint32)
    dst = topology.target if topology.m else np.zeros(1, np.uint32)
    code = A.lib().gr_layout_run(layout_id.encode(), n, src.ctypes.data, dst.ctypes.data,
                                 topology.m, buf, len(wire),
                                 C.addressof(xc) if xc is not None else None,
                                 xs.ctypes.data, ys.ctypes.data,
                                 zs.ctypes.data if zs is not None else None)
    _check(code)
    return Geometry(dim, xs[:n], ys[:n], zs[:n] if zs is not None else None)


class Stage:
    """stage.rs:23-31: run(topology, params) -> Geometry, a pure function of its inputs."""

    ID: str = ""
    DIM: int = 2

    @classmethod
    def run(cls, topology: Topology, params: Optional[Fa2Params] = None,
            ext: Optional[Fa2Ext] = None) -> Geometry:
        p = params or Fa2Params()
        return _run_id(cls.ID, topology, p.to_wire(), ext, cls.DIM)


class ForceAtlas2(Stage):
    ID = "layout.forceatlas2"


class ForceAtlas2BarnesHut(Stage):
    ID = "layout.forceatlas2.barnes_hut"


class ForceAtlas2_3d(Stage):
    ID = "layout.forceatlas2.3d"
    DIM = 3


def run_3d(topology: Topology, params: Optional[Fa2Params] = None,
           ext: Optional[Fa2Ext] = None) -> Geometry:
    """layout/forceatlas2.rs:63-68: the dense arm with dim forced to 3."""
    return ForceAtlas2_3d.run(topology, params, ext)


@dataclass
class Capability:
    """One registry entry (registry/capability.rs): id, runner, parameter schema, ledger."""

    id: str
    run: Callable[..., Geometry]
    params: tuple
    scale_ceiling: int
    oracle: str


_SCHEMA = (  # registry/tunable.rs:229-240: name, kind, min, max, default, step
    ("max_iter", "Int", 1.0, 10000.0, 100.0, 1.0),
    ("jitter_tolerance", "Float", 0.0, 1e6, 1.0, 0.1),
    ("scaling_ratio", "Float", 0.0, 1e6, 2.0, 0.1),
    ("gravity", "Float", 0.0, 1e6, 1.0, 0.1),
    ("seed", "Int", 0.0, 4294967295.0, 0.0, 1.0),
)

LAYOUTS: Dict[str, Capability] = {
    ForceAtlas2.ID: Capability(ForceAtlas2.ID, ForceAtlas2.run, _SCHEMA, 14_000,
                               "networkx 3.6 forceatlas2_layout"),
    ForceAtlas2BarnesHut.ID: Capability(ForceAtlas2BarnesHut.ID, ForceAtlas2BarnesHut.run,
                                        _SCHEMA, 250_000, "layout.forceatlas2"),
    ForceAtlas2_3d.ID: Capability(ForceAtlas2_3d.ID, run_3d, _SCHEMA, 14_000,
                                  "networkx 3.6 forceatlas2_layout at dim 3"),
}


def _forcesim_run(topology: Topology, params=None, ext=None) -> Geometry:
    from .forcesim import ForceAtlas2ForceSim
    if ext is not None:
        raise StageError(A.GR_ERR_UNSUPPORTED, "ext", "no extension applies to this layout")
    return ForceAtlas2ForceSim.run(topology, params)


# registry/forceatlas2_forcesim.rs:67-72: no published parameters, ceiling 2000
LAYOUTS["layout.forceatlas2.forcesim"] = Capability(
    "layout.forceatlas2.forcesim", _forcesim_run, (), 2_000,
    "SciGraphs' own ForceSim (model FA2, direct repulsion)")


def _d3_run(topology: Topology, params=None, ext=None) -> Geometry:
    # the registry form: no published parameters; the strength column as ext.edge_weights
    if params is not None:
        raise StageError(A.GR_ERR_UNSUPPORTED, "params", "layout.force.barnes_hut publishes none")
    return _run_id("layout.force.barnes_hut", topology, b"", ext, 2)


# registry/layouts.rs:92-97: no published parameters; ceiling FORCE_CEILING 100 000
# (registry/force.rs:31, :92-126)
LAYOUTS["layout.force.barnes_hut"] = Capability(
    "layout.force.barnes_hut", _d3_run, (), 100_000,
    "d3-force@3.0.0 src/{manyBody,link,center,collide}.js at TICKS=112")


def _pm_run(topology: Topology, params=None, ext=None) -> Geometry:
    if params is not None:
        raise StageError(A.GR_ERR_UNSUPPORTED, "params",
                         "layout.force.particle_mesh publishes none")
    return _run_id("layout.force.particle_mesh", topology, b"", ext, 2)


# registry/force.rs:291-298: no published parameters; ceiling PM_CEILING 800 000 (:137)
LAYOUTS["layout.force.particle_mesh"] = Capability(
    "layout.force.particle_mesh", _pm_run, (), 800_000,
    "graph-cli stress --oracle d3 --layout layout.force.particle_mesh")


def find(layout_id: str) -> Capability:
    if layout_id not in LAYOUTS:
        raise StageError(A.GR_ERR_UNKNOWN_LAYOUT, "layout_id", "a registered layout id")
    return LAYOUTS[layout_id]


def schedule_select(n: int, m: int, dim: int = 2, precision: str = "fp64",
                    ngraphs: int = 1) -> dict:
    """The measured table's launch schedule for these sizes (a pure function)."""
    out = A.GrFa2Schedule()
    _check(A.lib().gr_fa2_schedule_select(int(n), int(m), int(dim), PRECISIONS[precision],
                                          int(ngraphs), C.byref(out)))
    return {f: int(getattr(out, f)) for f, _ in A.GrFa2Schedule._fields_}


def run_params(layout_id: str, topology: Topology, wire: bytes,
               ext: Optional[Fa2Ext] = None) -> Geometry:
    """registry/params.rs:144-160: the layout at the values a wire buffer carries (empty =
    defaults), refused (never clamped) when malformed or out of range."""
    find(layout_id)
    dim = 3 if layout_id in (ForceAtlas2_3d.ID, "layout.forceatlas2.forcesim") else 2
    return _run_id(layout_id, topology, wire, ext, dim)


# ---------------------------------------------------------------- resident session


class Session:
    """A Fa2State kept resident in HBM (state.rs:60-80) and stepped on the device."""

    def __init__(self, topology: Topology, params: Optional[Fa2Params] = None,
                 ext: Optional[Fa2Ext] = None):
        self.params = params or Fa2Params()
        self.ext = ext or Fa2Ext()
        self.n, self.dim = int(topology.n), int(self.params.dim)
        self.m_raw = int(topology.m)
        xc = self.ext._c(self.n, topology.m, self.dim)
        cp = self.params._c()
        src = topology.source if topology.m else np.zeros(1, np.uint32)
        dst = topology.target if topology.m else np.zeros(1, np.uint32)
        self._h = A.lib().gr_fa2_session_create(self.n, src.ctypes.data, dst.ctypes.data,
                                                topology.m, C.byref(cp), C.addressof(xc))
        if not self._h:
            _raise()

    def close(self):
        if getattr(self, "_h", None):
            A.lib().gr_fa2_session_release(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @classmethod
    def batch(cls, topologies, params: Optional[Fa2Params] = None,
              ext: Optional[Fa2Ext] = None) -> "Session":
        """Many independent graphs in one session, laid out by the same launches. Positions,
        weights and starts are the graphs' concatenated in order."""
        self = cls.__new__(cls)
        self.params = params or Fa2Params()
        self.ext = ext or Fa2Ext()
        self.dim = int(self.params.dim)
        counts = np.array([t.n for t in topologies], np.uint32)
        edges = np.array([t.m for t in topologies], np.uint64)
        self.n = int(counts.sum())
        self.node_counts = counts
        src = np.concatenate([t.source for t in topologies] + [np.zeros(1, np.uint32)])
        dst = np.concatenate([t.target for t in topologies] + [np.zeros(1, np.uint32)])
        m = int(edges.sum())
        self.m_raw = m
        xc = self.ext._c(self.n, m, self.dim)
        cp = self.params._c()
        self._h = A.lib().gr_fa2_batch_create(len(topologies), counts.ctypes.data,
                                              edges.ctypes.data, src.ctypes.data,
                                              dst.ctypes.data, C.byref(cp), C.addressof(xc))
        if not self._h:
            _raise()
        return self

    def graph_scalars(self, graph: int) -> dict:
        out = np.zeros(7)
        _check(A.lib().gr_fa2_session_graph_scalars(self._h, int(graph), out.ctypes.data))
        keys = ("swing", "traction", "speed", "speed_efficiency", "iter_no", "iters_done",
                "done")
        return dict(zip(keys, out.tolist()))

    def step(self, iters: int) -> int:
        done = C.c_uint32(0)
        _check(A.lib().gr_fa2_session_step(self._h, int(iters), C.byref(done)))
        return int(done.value)

    def run(self) -> int:
        return self.step(self.params.max_iter)

    def positions(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1) * self.dim)
        _check(A.lib().gr_fa2_session_positions(self._h, out.ctypes.data))
        return out[: self.n * self.dim].reshape(self.n, self.dim)

    def set_positions(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float64).reshape(-1)
        assert p.size == self.n * self.dim
        _check(A.lib().gr_fa2_session_set_positions(self._h, p.ctypes.data))

    def set_params(self, params: Fa2Params, theta: Optional[float] = None):
        """New jitter tolerance, scaling ratio, gravity and seed (and opening angle) for
        the following iterations; the adaptive-speed state carries on."""
        cp = params._c()
        _check(A.lib().gr_fa2_session_set_params(self._h, C.byref(cp),
                                                 -1.0 if theta is None else float(theta)))
        self.params = params

    def record(self, iters: int, checkpoint_every: int = 20) -> int:
        """Runs up to `iters` iterations like step(), keeping a checkpoint every
        `checkpoint_every` iterations for adjoint(). Returns the iterations applied."""
        done = C.c_uint32(0)
        _check(A.lib().gr_fa2_session_record(self._h, int(iters), int(checkpoint_every),
                                             C.byref(done)))
        return int(done.value)

    def adjoint(self, pos_bar) -> dict:
        """Reverse sweep of the last record(): from d loss / d final positions (n, dim) to
        d loss / d raw edge weight, d scaling_ratio, d gravity and d start positions."""
        pb = np.ascontiguousarray(pos_bar, np.float64).reshape(-1)
        if pb.size != self.n * self.dim:
            raise ValueError("pos_bar must hold n*dim values")
        wb = np.zeros(max(self.m_raw, 1))
        par = np.zeros(2)
        sb = np.zeros(max(self.n * self.dim, 1))
        _check(A.lib().gr_fa2_session_adjoint(self._h, pb.ctypes.data if pb.size else None,
                                              wb.ctypes.data, par.ctypes.data,
                                              sb.ctypes.data))
        return {"weights": wb[: self.m_raw], "scaling_ratio": float(par[0]),
                "gravity": float(par[1]),
                "start": sb[: self.n * self.dim].reshape(self.n, self.dim)}

    def forces(self, parts: int = 7) -> np.ndarray:
        out = np.zeros(max(self.n, 1) * self.dim)
        _check(A.lib().gr_fa2_session_forces(self._h, int(parts), out.ctypes.data))
        return out[: self.n * self.dim].reshape(self.n, self.dim)

    def scalars(self) -> dict:
        out = np.zeros(7)
        _check(A.lib().gr_fa2_session_scalars(self._h, out.ctypes.data))
        keys = ("swing", "traction", "speed", "speed_efficiency", "iter_no", "iters_done",
                "done")
        return dict(zip(keys, out.tolist()))

    def graph(self):
        m = C.c_uint32(0)
        _check(A.lib().gr_fa2_session_graph(self._h, C.byref(m), None, None, None))
        lo = np.zeros(max(m.value, 1), np.uint32)
        hi = np.zeros(max(m.value, 1), np.uint32)
        mass = np.zeros(max(self.n, 1))
        _check(A.lib().gr_fa2_session_graph(self._h, C.byref(m), lo.ctypes.data,
                                            hi.ctypes.data, mass.ctypes.data))
        return lo[: m.value], hi[: m.value], mass[: self.n]

    def tree(self, bounds: bool = True) -> dict:
        nc = C.c_uint32(0)
        order = np.zeros(max(self.n, 1), np.uint32)
        _check(A.lib().gr_fa2_session_tree(self._h, C.byref(nc), 0, None, None, None, None,
                                           None, order.ctypes.data))
        c = nc.value
        cap = max(c, 1)
        b = np.zeros(cap * 2 * self.dim) if bounds else None
        sk = np.zeros(cap, np.uint32)
        st = np.zeros(cap, np.uint32)
        en = np.zeros(cap, np.uint32)
        bo = np.zeros(cap * (1 + self.dim))
        _check(A.lib().gr_fa2_session_tree(
            self._h, C.byref(nc), cap, b.ctypes.data if b is not None else None,
            sk.ctypes.data, st.ctypes.data, en.ctypes.data, bo.ctypes.data, None))
        out = {"ncells": c, "order": order[: self.n], "skip": sk[:c], "start": st[:c],
               "end": en[:c], "bodies": bo[: c * (1 + self.dim)].reshape(c, 1 + self.dim)}
        if b is not None:
            out["bounds"] = b[: c * 2 * self.dim].reshape(c, 2 * self.dim)
        return out

    def schedule(self) -> dict:
        """The session's launch schedule (it never changes a result bit)."""
        out = A.GrFa2Schedule()
        _check(A.lib().gr_fa2_session_get_schedule(self._h, C.byref(out)))
        return {f: int(getattr(out, f)) for f, _ in A.GrFa2Schedule._fields_}

    def set_schedule(self, **sched):
        cur = self.schedule()
        cur.update(sched)
        sc = A.GrFa2Schedule(*(int(cur[f]) for f, _ in A.GrFa2Schedule._fields_))
        _check(A.lib().gr_fa2_session_set_schedule(self._h, C.byref(sc)))

    def autotune(self, reps: int = 5) -> dict:
        """Times each schedule knob's candidates on this session, installs the fastest and
        caches it for the session's size class; returns it."""
        out = A.GrFa2Schedule()
        _check(A.lib().gr_fa2_session_autotune(self._h, int(reps), C.byref(out)))
        return {f: int(getattr(out, f)) for f, _ in A.GrFa2Schedule._fields_}

    def walk_stats(self) -> dict:
        """One counting run of the walk on the current tree: scan steps (summed over waves),
        lanes acting, cells accepted and leaf pairs taken."""
        out = np.zeros(4, np.uint64)
        _check(A.lib().gr_fa2_session_walk_stats(self._h, out.ctypes.data))
        return dict(zip(("steps", "acting_lanes", "accepted", "pairs"), map(int, out)))

    def time(self, stage: str, reps: int) -> float:
        ms = C.c_double(0.0)
        _check(A.lib().gr_fa2_session_time(self._h, stage.encode(), int(reps), C.byref(ms)))
        return float(ms.value)
