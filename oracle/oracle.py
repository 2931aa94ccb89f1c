"""ctypes face of the CPU oracle: TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module,
and only to check (or time, as the CPU baseline) the product path. The product library
never calls it. Build with ``make -C oracle`` (or ``build()`` below).

Parity status of each function is stated in fa2_oracle.h; in short: the dense 2D FA2 is
pinned by the reference's networkx fixture, the tree by the brute-force PR-quadtree
definition and the reference's quadtree invariants, the BH walk by theta=0 == dense.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfa2_oracle.so")

_lib = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.gro_mulberry32_next.restype = C.c_double
        L.gro_mulberry32_next.argtypes = [C.POINTER(C.c_uint32)]
        L.gro_initial_positions.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, f64p]
        L.gro_fmix64.restype = C.c_uint64
        L.gro_fmix64.argtypes = [C.c_uint64]
        L.gro_jiggle.restype = C.c_double
        L.gro_jiggle.argtypes = [C.c_uint32] * 5
        L.gro_gate_edges.restype = C.c_size_t
        L.gro_gate_edges.argtypes = [C.c_uint32, u32p, u32p, C.c_size_t]
        L.gro_fa2_new.restype = C.c_void_p
        L.gro_fa2_new.argtypes = [C.c_uint32, u32p, u32p, C.c_size_t, C.c_void_p, C.c_uint32,
                                  C.c_double, C.c_double, C.c_double, C.c_uint32, C.c_void_p,
                                  C.c_int, C.c_double]
        L.gro_fa2_free.argtypes = [C.c_void_p]
        L.gro_fa2_iterate.restype = C.c_double
        L.gro_fa2_iterate.argtypes = [C.c_void_p]
        L.gro_fa2_run.restype = C.c_uint32
        L.gro_fa2_run.argtypes = [C.c_void_p, C.c_uint32]
        L.gro_fa2_positions.argtypes = [C.c_void_p, f64p]
        L.gro_fa2_set_positions.argtypes = [C.c_void_p, f64p]
        L.gro_fa2_mass.argtypes = [C.c_void_p, f64p]
        L.gro_fa2_forces.argtypes = [C.c_void_p, C.c_uint32, f64p]
        L.gro_fa2_scalars.argtypes = [C.c_void_p, f64p]
        L.gro_fa2_simple_edges.restype = C.c_uint32
        L.gro_fa2_simple_edges.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.gro_fa2_set_theta2.argtypes = [C.c_void_p, C.c_double, C.c_int]
        L.gro_fa2_set_params.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double,
                                         C.c_uint32]
        L.gro_tree_build.restype = C.c_void_p
        L.gro_tree_build.argtypes = [C.c_uint32, C.c_uint32, f64p, f64p, C.c_void_p]
        L.gro_tree_free.argtypes = [C.c_void_p]
        L.gro_tree_ncells.restype = C.c_uint32
        L.gro_tree_ncells.argtypes = [C.c_void_p]
        L.gro_tree_cells.argtypes = [C.c_void_p, f64p, u32p, u32p, u32p]
        L.gro_tree_order.argtypes = [C.c_void_p, u32p]
        L.gro_tree_bodies.argtypes = [C.c_void_p, f64p, f64p, f64p, C.c_void_p, f64p]
        L.gro_tree_node_ids.argtypes = [C.c_void_p, u32p]
        L.gro_walk_counts.argtypes = [C.c_uint32, C.c_uint32, f64p, f64p, C.c_void_p, f64p,
                                      C.c_double, C.c_uint32, C.POINTER(C.c_uint64)]
        L.gro_gate_model.restype = C.c_size_t
        L.gro_gate_model.argtypes = [C.c_uint32, u32p, u32p, f64p, C.c_size_t]
        L.gro_d3sim_run.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                    C.c_uint32, f64p, C.c_uint32, C.c_uint32, f64p, f64p,
                                    C.c_void_p, C.c_void_p]
        L.gro_pmd3_charge_once.argtypes = [C.c_uint32, f64p, C.c_int, C.c_void_p, C.c_double,
                                           C.c_double, C.c_double, C.c_double, f64p, f64p,
                                           C.c_void_p, C.c_void_p]
        L.gro_pmd3_collide_once.argtypes = [C.c_uint32, f64p, C.c_double, C.c_uint32,
                                            C.c_uint32, f64p, C.c_void_p]
        L.gro_libm_cos.restype = C.c_double
        L.gro_libm_cos.argtypes = [C.c_double]
        L.gro_libm_sin.restype = C.c_double
        L.gro_libm_sin.argtypes = [C.c_double]
        L.gro_d3_collide.argtypes = [C.c_uint32, f64p, C.c_double, C.c_uint32, C.c_uint32, f64p]
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.gro_pcg64_state.argtypes = [C.c_uint64, u64p]
        L.gro_pcg64_raw.argtypes = [C.c_uint64, C.c_uint32, u64p]
        L.gro_pcg64_doubles.argtypes = [C.c_uint64, C.c_uint32, f64p]
        L.gro_pairwise_f32.restype = C.c_float
        L.gro_pairwise_f32.argtypes = [f32p, C.c_size_t]
        L.gro_pairwise_f64.restype = C.c_double
        L.gro_pairwise_f64.argtypes = [f64p, C.c_size_t]
        L.gro_dot3_f32.restype = C.c_float
        L.gro_dot3_f32.argtypes = [f32p, f32p]
        L.gro_cbrt_cr.restype = C.c_double
        L.gro_cbrt_cr.argtypes = [C.c_double]
        L.gro_forcesim_constants.argtypes = [C.c_uint32, f32p, C.c_double, C.c_double, f64p]
        L.gro_forcesim_run.restype = C.c_int
        L.gro_forcesim_run.argtypes = [C.c_uint32, C.c_void_p, C.c_void_p, C.c_size_t,
                                       C.c_uint32, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_uint64, f32p, f64p]
        L.gro_d3_charge_bh.argtypes = [C.c_uint32, f64p, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_uint32, C.c_uint32, f64p]
        L.gro_pm_h.restype = C.c_double
        L.gro_pm_h.argtypes = [C.c_int]
        L.gro_fft2.argtypes = [f64p, f64p, C.c_int, C.c_int]
        L.gro_pm_forces.restype = C.c_int
        L.gro_pm_forces.argtypes = [C.c_uint32, f64p, f64p, C.c_double, C.c_int, f64p, f64p]
        L.gro_pm_charge.restype = C.c_int
        L.gro_pm_charge.argtypes = [C.c_uint32, f64p, C.c_int, C.c_double, C.c_double,
                                    C.c_double, C.c_double, f64p, f64p]
        _lib = L
    return _lib


THETA2_REFERENCE = 0.8 * 0.8  # barnes_hut.rs:24


def initial_positions(n: int, seed: int, dim: int) -> np.ndarray:
    out = np.zeros(max(n, 1) * dim, np.float64)
    lib().gro_initial_positions(n, seed, dim, out)
    return out[: n * dim].reshape(n, dim)


def jiggle(seed, tick, pss, i, j) -> float:
    return lib().gro_jiggle(seed, tick, pss, i, j)


def fmix64(k: int) -> int:
    return lib().gro_fmix64(k)


def mulberry32(seed: int, count: int) -> np.ndarray:
    st = C.c_uint32(seed)
    return np.array([lib().gro_mulberry32_next(C.byref(st)) for _ in range(count)])


def gate_edges(count: int):
    """Raw (source, target) of the reference gate model seeded_model(*, count, 8)."""
    cap = 2 * count + count // 20 + 1
    s = np.zeros(cap, np.uint32)
    t = np.zeros(cap, np.uint32)
    m = lib().gro_gate_edges(count, s, t, cap)
    return s[:m].copy(), t[:m].copy()


class Fa2:
    """The reference Fa2State, dense (repulsion=0) or Barnes-Hut (repulsion=1)."""

    def __init__(self, n, source, target, *, dim=2, jitter_tolerance=1.0, scaling_ratio=2.0,
                 gravity=1.0, seed=0, start=None, weights=None, tree=False,
                 theta2=THETA2_REFERENCE):
        self.n, self.dim = int(n), int(dim)
        src = np.ascontiguousarray(source, np.uint32)
        dst = np.ascontiguousarray(target, np.uint32)
        self._keep = []
        wp = None
        if weights is not None:
            w = np.ascontiguousarray(weights, np.float64)
            self._keep.append(w)
            wp = w.ctypes.data
        sp = None
        if start is not None:
            st = np.ascontiguousarray(start, np.float64).reshape(-1)
            assert st.size == self.n * self.dim
            self._keep.append(st)
            sp = st.ctypes.data
        if src.size == 0:
            src = np.zeros(1, np.uint32)
            dst = np.zeros(1, np.uint32)
            m = 0
        else:
            m = src.size
        self._h = lib().gro_fa2_new(self.n, src, dst, m, wp, self.dim, jitter_tolerance,
                                   scaling_ratio, gravity, seed, sp, 1 if tree else 0, theta2)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().gro_fa2_free(h)
            self._h = None

    def iterate(self) -> float:
        return lib().gro_fa2_iterate(self._h)

    def run(self, max_iter: int) -> int:
        return lib().gro_fa2_run(self._h, max_iter)

    def positions(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1) * self.dim)
        lib().gro_fa2_positions(self._h, out)
        return out[: self.n * self.dim].reshape(self.n, self.dim)

    def set_positions(self, p):
        p = np.ascontiguousarray(p, np.float64).reshape(-1)
        lib().gro_fa2_set_positions(self._h, p)

    def mass(self) -> np.ndarray:
        out = np.zeros(max(self.n, 1))
        lib().gro_fa2_mass(self._h, out)
        return out[: self.n]

    def forces(self, parts: int = 7) -> np.ndarray:
        out = np.zeros(max(self.n, 1) * self.dim)
        lib().gro_fa2_forces(self._h, parts, out)
        return out[: self.n * self.dim].reshape(self.n, self.dim)

    def scalars(self) -> np.ndarray:
        out = np.zeros(5)
        lib().gro_fa2_scalars(self._h, out)
        return out

    def set_repulsion(self, tree: bool, theta2: float = THETA2_REFERENCE):
        lib().gro_fa2_set_theta2(self._h, theta2, 1 if tree else 0)

    def set_params(self, jitter_tolerance=1.0, scaling_ratio=2.0, gravity=1.0, seed=0):
        lib().gro_fa2_set_params(self._h, jitter_tolerance, scaling_ratio, gravity, seed)

    def simple_edges(self):
        m = lib().gro_fa2_simple_edges(self._h, None, None)
        lo = np.zeros(max(m, 1), np.uint32)
        hi = np.zeros(max(m, 1), np.uint32)
        lib().gro_fa2_simple_edges(self._h, lo.ctypes.data, hi.ctypes.data)
        return lo[:m], hi[:m]


class Tree:
    """The reference's preorder quadtree (dim 2) or octree (dim 3) over given points."""

    def __init__(self, pts: np.ndarray):
        pts = np.asarray(pts, np.float64)
        self.n, self.dim = pts.shape
        self._cols = [np.ascontiguousarray(pts[:, a]) for a in range(self.dim)]
        if self.n == 0:
            self._cols = [np.zeros(1) for _ in range(self.dim)]
        zs = self._cols[2].ctypes.data if self.dim == 3 else None
        self._h = lib().gro_tree_build(self.n, self.dim, self._cols[0], self._cols[1], zs)
        nc = lib().gro_tree_ncells(self._h)
        self.ncells = nc
        b = np.zeros(max(nc, 1) * 2 * self.dim)
        sk = np.zeros(max(nc, 1), np.uint32)
        s0 = np.zeros(max(nc, 1), np.uint32)
        s1 = np.zeros(max(nc, 1), np.uint32)
        lib().gro_tree_cells(self._h, b, sk, s0, s1)
        self.bounds = b[: nc * 2 * self.dim].reshape(nc, 2 * self.dim)
        self.skip, self.start, self.end = sk[:nc], s0[:nc], s1[:nc]
        o = np.zeros(max(self.n, 1), np.uint32)
        lib().gro_tree_order(self._h, o)
        self.order = o[: self.n]

    def node_ids(self) -> np.ndarray:
        """Each cell's node id: the shape slot the insertion created it as."""
        out = np.zeros(max(self.ncells, 1), np.uint32)
        lib().gro_tree_node_ids(self._h, out)
        return out[: self.ncells]

    def bodies(self, mass: np.ndarray) -> np.ndarray:
        m = np.ascontiguousarray(mass, np.float64)
        out = np.zeros(max(self.ncells, 1) * (1 + self.dim))
        zs = self._cols[2].ctypes.data if self.dim == 3 else None
        lib().gro_tree_bodies(self._h, m, self._cols[0], self._cols[1], zs, out)
        return out[: self.ncells * (1 + self.dim)].reshape(self.ncells, 1 + self.dim)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().gro_tree_free(h)
            self._h = None


def walk_counts(pts, mass, theta2: float = THETA2_REFERENCE, group: int = 64) -> dict:
    """The reference walk replayed per tree position, counted as a wave-uniform scan over
    groups of `group` consecutive positions: steps (the union of each group's visit sets,
    summed), visits, cells accepted, leaf pairs."""
    pts = np.asarray(pts, np.float64)
    n, dim = pts.shape
    cols = [np.ascontiguousarray(pts[:, a]) for a in range(dim)]
    m = np.ascontiguousarray(mass, np.float64)
    out = (C.c_uint64 * 4)()
    lib().gro_walk_counts(n, dim, cols[0], cols[1], cols[2].ctypes.data if dim == 3 else None,
                          m, float(theta2), int(group), out)
    return dict(zip(("steps", "acting_lanes", "accepted", "pairs"), (int(v) for v in out)))


def fft2(z: np.ndarray, sign: int = -1) -> np.ndarray:
    side = z.shape[0]
    re = np.ascontiguousarray(z.real, np.float64).copy()
    im = np.ascontiguousarray(z.imag, np.float64).copy()
    lib().gro_fft2(re, im, side, sign)
    return re + 1j * im


def pm_forces(xy: np.ndarray, mass: np.ndarray, k: float, side: int):
    """FA2-law particle-mesh repulsion: (forces (n,2), frame dict or None)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    fr = np.zeros(6)
    ok = lib().gro_pm_forces(n, xy.reshape(-1) if n else np.zeros(2),
                             np.ascontiguousarray(mass, np.float64) if n else np.zeros(1),
                             k, side, out, fr)
    frame = None
    if ok:
        frame = dict(h=fr[0], ox=fr[1], oy=fr[2], step=int(fr[3]), cells=int(fr[4]),
                     reach=int(fr[5]))
    return out[: 2 * n].reshape(n, 2), frame


def pm_charge(xy: np.ndarray, side: int, charge: float = -90.0, alpha: float = 1.0,
              distance_min: float = 1.0, distance_max: float = 520.0):
    """d3-law particle-mesh charge pass (the velocities it adds): (deltas (n,2), frame)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    fr = np.zeros(6)
    ok = lib().gro_pm_charge(n, xy.reshape(-1) if n else np.zeros(2), side, charge, alpha,
                             distance_min, distance_max, out, fr)
    frame = None
    if ok:
        frame = dict(h=fr[0], ox=fr[1], oy=fr[2], step=int(fr[3]), cells=int(fr[4]),
                     reach=int(fr[5]))
    return out[: 2 * n].reshape(n, 2), frame


def d3_charge_bh(xy: np.ndarray, theta: float = 0.9, charge: float = -30.0,
                 alpha: float = 1.0, distance_min: float = 1.0,
                 distance_max: float = float("inf"), seed: int = 0, tick: int = 0):
    """d3's many-body force over the quadtree (barnes_hut/charge.rs): the velocity each node
    gains, (n, 2)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    lib().gro_d3_charge_bh(n, xy.reshape(-1) if n else np.zeros(2), theta, charge, alpha,
                           distance_min, distance_max, seed, tick, out)
    return out[: 2 * n].reshape(n, 2)


# ---- layout.forceatlas2.forcesim (oracle/forcesim_oracle.c) ----

FORCESIM_SEED = 1_767_573_729


def pcg64_state(seed: int):
    """numpy's default_rng(seed) after seeding: (state, inc) as Python ints."""
    out = np.zeros(4, np.uint64)
    lib().gro_pcg64_state(seed, out)
    return (int(out[0]) << 64) | int(out[1]), (int(out[2]) << 64) | int(out[3])


def pcg64_raw(seed: int, count: int) -> np.ndarray:
    out = np.zeros(max(count, 1), np.uint64)
    lib().gro_pcg64_raw(seed, count, out)
    return out[:count]


def pcg64_doubles(seed: int, count: int) -> np.ndarray:
    out = np.zeros(max(count, 1))
    lib().gro_pcg64_doubles(seed, count, out)
    return out[:count]


def pairwise_f32(v) -> np.float32:
    v = np.ascontiguousarray(v, np.float32)
    return np.float32(lib().gro_pairwise_f32(v if v.size else np.zeros(1, np.float32), v.size))


def pairwise_f64(v) -> float:
    v = np.ascontiguousarray(v, np.float64)
    return lib().gro_pairwise_f64(v if v.size else np.zeros(1), v.size)


def dot3_f32(a, b) -> np.float32:
    return np.float32(lib().gro_dot3_f32(np.ascontiguousarray(a, np.float32),
                                         np.ascontiguousarray(b, np.float32)))


def cbrt_cr(x: float) -> float:
    return lib().gro_cbrt_cr(float(x))


def forcesim_constants(mass, scaling_ratio=2.0, gravity=1.0):
    mass = np.ascontiguousarray(mass, np.float32)
    out = np.zeros(3)
    lib().gro_forcesim_constants(mass.size, mass if mass.size else np.zeros(1, np.float32),
                                 scaling_ratio, gravity, out)
    return tuple(out)


def forcesim(n: int, source, target, iterations: int = 50, scale: float = 5.0,
             scaling_ratio: float = 2.0, gravity: float = 1.0, jitter_tolerance: float = 1.0,
             seed: int = FORCESIM_SEED):
    """layout.forceatlas2.forcesim's answer: (xyz (n, 3) f32, status, (k, repulsion, gravity)).
    status 0, or 2 for NonFinite{node.x} (the columns are then zero)."""
    s = np.ascontiguousarray(source, np.uint32)
    t = np.ascontiguousarray(target, np.uint32)
    out = np.zeros(max(n, 1) * 3, np.float32)
    cst = np.zeros(3)
    st = lib().gro_forcesim_run(n, s.ctypes.data if s.size else None,
                                t.ctypes.data if t.size else None, s.size, iterations, scale,
                                scaling_ratio, gravity, jitter_tolerance, seed, out, cst)
    return out[: 3 * n].reshape(n, 3), st, tuple(cst)


# ---- layout.force.barnes_hut, the frozen d3 simulation ----

D3_DEFAULTS = dict(charge=-90.0, theta=0.9, distance_min=1.0, distance_max=520.0,
                   link_distance=60.0, link_strength_scale=0.15, collide_radius=16.0,
                   center_strength=1.0, velocity_decay=1.0 - 0.42, alpha_decay=0.06,
                   initial_alpha=1.0)
D3_TICKS = 112


def gate_model(count: int):
    """synthetic.rs's model: (source, target, strength) of its raw edges."""
    cap = 2 * count + 8
    s = np.zeros(cap, np.uint32)
    t = np.zeros(cap, np.uint32)
    w = np.zeros(cap)
    m = lib().gro_gate_model(count, s, t, w, cap)
    return s[:m], t[:m], w[:m]


def d3sim(n: int, source, target, strength, ticks: int = D3_TICKS, seed: int = 0,
          velocities: bool = False, mesh_side: int = 0, **params):
    """The frozen layout.force.barnes_hut's f64 positions after `ticks` ticks: (x, y), and
    (x, y, vx, vy) with velocities=True. mesh_side > 0: layout.force.particle_mesh on a mesh
    of that side instead (mesh.rs side_for(n) is the stage's)."""
    p = dict(D3_DEFAULTS)
    p.update(params)
    vec = np.array([p[k] for k in ("charge", "theta", "distance_min", "distance_max",
                                   "link_distance", "link_strength_scale", "collide_radius",
                                   "center_strength", "velocity_decay", "alpha_decay",
                                   "initial_alpha")], np.float64)
    s = np.ascontiguousarray(source, np.uint32)
    t = np.ascontiguousarray(target, np.uint32)
    w = np.ascontiguousarray(strength, np.float64)
    x = np.zeros(max(n, 1))
    y = np.zeros(max(n, 1))
    vx = np.zeros(max(n, 1))
    vy = np.zeros(max(n, 1))
    lib().gro_d3sim_run(n, s.ctypes.data if s.size else None, t.ctypes.data if t.size else None,
                        w.ctypes.data if w.size else None, s.size, ticks, vec, seed,
                        int(mesh_side), x, y, vx.ctypes.data, vy.ctypes.data)
    if velocities:
        return x[:n], y[:n], vx[:n], vy[:n]
    return x[:n], y[:n]


def libm_cos(x: float) -> float:
    """libm 0.2.16's cos (fdlibm via musl), medium reduction only (NaN past 2^20 pi/2)."""
    return lib().gro_libm_cos(float(x))


def libm_sin(x: float) -> float:
    return lib().gro_libm_sin(float(x))


def golden_spiral(n: int):
    """seed.rs:1-17: (x, y) of node i at 12 sqrt(i + 1) along the golden angle."""
    x = np.array([libm_cos(i * 2.399963229728653) * (12.0 * np.sqrt(i + 1.0)) for i in range(n)])
    y = np.array([libm_sin(i * 2.399963229728653) * (12.0 * np.sqrt(i + 1.0)) for i in range(n)])
    return x, y


def pm_side_for(n: int) -> int:
    """mesh.rs:31-34: ceil(sqrt n) rounded up to a power of two, held to 128..1024."""
    root = int(np.ceil(np.sqrt(float(n))))
    side = 1
    while side < root:
        side <<= 1
    return min(max(side, 128), 1024)


def pmd3_charge(xy, side: int, charge: float = -90.0, alpha: float = 1.0,
                distance_min: float = 1.0, distance_max: float = 520.0, order=None):
    """particle_mesh/charge.rs from a fresh mesh (identity grid order unless `order`):
    (deltas (n, 2) node order, frame dict or None, kernel spectrum (re, im) or None)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    fr = np.zeros(7)
    sre = np.zeros(side * side)
    sim_ = np.zeros(side * side)
    o = None if order is None else np.ascontiguousarray(order, np.uint32)
    placed = lib().gro_pmd3_charge_once(n, xy if n else np.zeros(2), side,
                                        o.ctypes.data if o is not None else None, charge, alpha,
                                        distance_min, distance_max, out, fr, sre.ctypes.data,
                                        sim_.ctypes.data)
    frame = None
    if placed:
        frame = dict(step=int(fr[0]), h=float(fr[1]), ox=float(fr[2]), oy=float(fr[3]),
                     cells=int(fr[4]), reach=int(fr[5]), side=int(fr[6]))
    return out[: 2 * n].reshape(n, 2), frame, (sre, sim_) if placed else None


def pmd3_collide(xy, radius: float = 16.0, seed: int = 0, tick: int = 0):
    """particle_mesh/collide.rs over these positions (the hashed grid): (deltas (n, 2) node
    order, the grid order it sorted), deltas zero at radius 0."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    order = np.zeros(max(n, 1), np.uint32)
    lib().gro_pmd3_collide_once(n, xy if n else np.zeros(2), radius, seed & 0xFFFFFFFF,
                                tick & 0xFFFFFFFF, out, order.ctypes.data)
    return out[: 2 * n].reshape(n, 2), order[:n]


def d3_collide(xy, radius: float = 16.0, seed: int = 0, tick: int = 0) -> np.ndarray:
    """collide.rs node_delta of every node over the quadtree of these positions: (n, 2)."""
    xy = np.ascontiguousarray(xy, np.float64).reshape(-1, 2)
    n = xy.shape[0]
    out = np.zeros(max(n, 1) * 2)
    lib().gro_d3_collide(n, xy if n else np.zeros(2), float(radius), seed & 0xFFFFFFFF,
                         tick & 0xFFFFFFFF, out)
    return out[: 2 * n].reshape(n, 2)
