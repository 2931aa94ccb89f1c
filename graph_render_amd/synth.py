"""Seeded synthetic graphs and start positions for the benchmark configurations.

numpy's PCG64 (default_rng) drives every draw, so a (generator, seed) pair names one edge
list; digest() gives its sha256 for the record.
"""
from __future__ import annotations

import hashlib

import numpy as np


def digest(source: np.ndarray, target: np.ndarray) -> str:
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(source, np.uint32).tobytes())
    h.update(np.ascontiguousarray(target, np.uint32).tobytes())
    return h.hexdigest()


def uniform_start(n: int, dim: int, seed: int, lo: float = -100.0, hi: float = 100.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, size=(n, dim))


def barabasi_albert(n: int, m: int, seed: int):
    """Preferential attachment: node t >= m links to m distinct earlier nodes drawn with
    probability proportional to degree (the first m nodes start as targets of node m)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    repeated = []
    targets = list(range(m))
    for t in range(m, n):
        src.extend([t] * len(targets))
        dst.extend(targets)
        repeated.extend(targets)
        repeated.extend([t] * len(targets))
        chosen = set()
        while len(chosen) < m:
            chosen.add(repeated[int(rng.integers(0, len(repeated)))])
        targets = sorted(chosen)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


def erdos_renyi(n: int, mean_degree: float, seed: int):
    rng = np.random.default_rng(seed)
    m = int(round(n * mean_degree / 2))
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    return a.astype(np.uint32), b.astype(np.uint32)


def sbm_zipf(n: int, blocks: int, zipf: float, mean_degree: float, seed: int,
             inter_fraction: float = 0.1):
    """Stochastic block model with Zipf(zipf) block sizes. A (1 - inter_fraction) share of
    the expected degree is within blocks (p_in set so the mean degree comes out at
    mean_degree), the rest between uniformly drawn nodes. Edge weights lognormal(0, 0.5)."""
    rng = np.random.default_rng(seed)
    w = 1.0 / np.arange(1, blocks + 1) ** zipf
    sizes = np.floor(w / w.sum() * n).astype(np.int64)
    sizes[0] += n - sizes.sum()
    m_total = int(round(n * mean_degree / 2))
    m_in = int(round(m_total * (1.0 - inter_fraction)))
    pairs_in = (sizes * (sizes - 1) / 2.0)
    p_in = m_in / pairs_in.sum()
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    src, dst = [], []
    for b in range(blocks):
        s = int(sizes[b])
        if s < 2:
            continue
        cnt = rng.binomial(int(pairs_in[b]), p_in)
        u = rng.integers(0, s, cnt) + starts[b]
        v = rng.integers(0, s, cnt) + starts[b]
        src.append(u)
        dst.append(v)
    m_out = m_total - m_in
    src.append(rng.integers(0, n, m_out))
    dst.append(rng.integers(0, n, m_out))
    s = np.concatenate(src).astype(np.uint32)
    t = np.concatenate(dst).astype(np.uint32)
    perm = rng.permutation(s.size)
    s, t = s[perm], t[perm]
    weights = rng.lognormal(0.0, 0.5, s.size)
    return s, t, weights


def sbm_equal(n: int, blocks: int, p_in: float, p_out: float, seed: int):
    rng = np.random.default_rng(seed)
    size = n // blocks
    block = np.arange(n) // size
    src, dst = [], []
    for b in range(blocks):
        lo = b * size
        hi = n if b == blocks - 1 else lo + size
        s = hi - lo
        cnt = rng.binomial(s * (s - 1) // 2, p_in)
        src.append(rng.integers(lo, hi, cnt))
        dst.append(rng.integers(lo, hi, cnt))
    total_out = (n * (n - 1) // 2) - sum(((min(n, (b + 1) * size) - b * size) *
                                          (min(n, (b + 1) * size) - b * size - 1)) // 2
                                         for b in range(blocks))
    cnt = rng.binomial(total_out, p_out)
    a = rng.integers(0, n, cnt * 2)
    c = rng.integers(0, n, cnt * 2)
    keep = block[a] != block[c]
    src.append(a[keep][:cnt])
    dst.append(c[keep][:cnt])
    return np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32)


def random_geometric_3d(n: int, mean_degree: float, seed: int):
    """Random geometric graph in the unit cube: points uniform (seeded), an edge for every
    pair closer than r, with n*(4/3)*pi*r^3 = mean_degree. Edges sorted by (lo, hi)."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    p = rng.random((n, 3))
    r = (mean_degree / (n * 4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    pairs = cKDTree(p).query_pairs(r, output_type="ndarray").astype(np.int64)
    pairs.sort(axis=1)
    pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
    return pairs[:, 0].astype(np.uint32), pairs[:, 1].astype(np.uint32), p


def barabasi_albert_fast(n: int, m: int, rng) -> tuple:
    """Preferential attachment, vectorised: node t >= m adds m edges whose targets are
    uniform picks from the endpoint list of all earlier edges (degree-proportional), the
    first new node linking to nodes 0..m-1. Picks that repeat a target are dropped later by
    the simple-graph dedup. Endpoint picks that land on an earlier edge's target are
    resolved by pointer doubling."""
    m = max(1, min(m, n - 1)) if n > 1 else 0
    if n <= 1 or m == 0:
        return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
    E = (n - m) * m
    e = np.arange(E, dtype=np.int64)
    t = m + e // m
    first = t == m
    pool = 2 * (t - m) * m
    u = np.floor(rng.random(E) * np.maximum(pool, 1)).astype(np.int64)
    half = u // 2
    even = (u % 2) == 0
    val = np.where(first, e % m, np.where(even, m + half // m, -1)).astype(np.int64)
    ptr = np.where(~first & ~even, half, -1).astype(np.int64)
    todo = np.nonzero(ptr >= 0)[0]
    while todo.size:
        p = ptr[todo]
        got = val[p] >= 0
        jump = ptr[p[~got]]  # taken before this round clears the pointers it resolves
        val[todo[got]] = val[p[got]]
        ptr[todo[got]] = -1
        rest = todo[~got]
        ptr[rest] = jump
        todo = rest
    return t.astype(np.uint32), val.astype(np.uint32)


def batch_workload(count: int = 4096, lo: int = 50, hi: int = 20000):
    """The batched benchmark: graph b (seed b) has log-uniform node count in [lo, hi]; even b
    are preferential-attachment graphs with m uniform in 1..6, odd b Erdos-Renyi graphs with
    mean degree uniform in [4, 12]; starts uniform in [-100, 100]^2."""
    sizes, srcs, dsts, starts = [], [], [], []
    for b in range(count):
        rng = np.random.default_rng(b)
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if b % 2 == 0:
            s, t = barabasi_albert_fast(n, int(rng.integers(1, 7)), rng)
        else:
            mdeg = rng.uniform(4.0, 12.0)
            k = int(round(n * mdeg / 2))
            s = rng.integers(0, n, k).astype(np.uint32)
            t = rng.integers(0, n, k).astype(np.uint32)
        sizes.append(n)
        srcs.append(s)
        dsts.append(t)
        starts.append(rng.uniform(-100.0, 100.0, (n, 2)))
    return sizes, srcs, dsts, np.concatenate(starts)


def _mulberry32(state: int):
    """synthetic.rs:44-50: the next state and draw (integer steps, then / 2^32)."""
    state = (state + 0x6D2B79F5) & 0xFFFFFFFF
    t = ((state ^ (state >> 15)) * (1 | state)) & 0xFFFFFFFF
    t = ((t + (((t ^ (t >> 7)) * (61 | t)) & 0xFFFFFFFF)) & 0xFFFFFFFF) ^ t
    return state, (t ^ (t >> 14)) / 4294967296.0


def gate_model(count: int):
    """The reference's gate model (synthetic.rs:92-170, seeded_model at REFERENCE_DEGREE):
    preferential attachment by x*y*i draws plus 5% extra links, from Mulberry32 seeded
    0x051042; the raw edge columns and each kept edge's strength (0.4 + 0.4 * a draw)."""
    st = 0x051042
    for _ in range(count):  # the node draws: a roll, then one pick for 80% of nodes
        st, roll = _mulberry32(st)
        if roll < 0.8:
            st, _ = _mulberry32(st)
    src, dst, strength = [], [], []

    def push(a, b):
        nonlocal st
        if a != b:
            src.append(a)
            dst.append(b)
            st, u = _mulberry32(st)
            strength.append(0.4 + u * 0.4)

    for i in range(1, count):
        st, x = _mulberry32(st)
        st, y = _mulberry32(st)
        push(i, int(x * y * i))
        st, u = _mulberry32(st)
        if u < 0.5:
            st, x = _mulberry32(st)
            st, y = _mulberry32(st)
            push(i, int(x * y * i))
    for _ in range(int(count * 0.05)):
        st, u = _mulberry32(st)
        a = int(u * count)
        st, u = _mulberry32(st)
        push(a, int(u * count))
    return (np.array(src, np.uint32), np.array(dst, np.uint32), np.array(strength, np.float64))


# The named workloads of the benchmark configurations. Each returns the graph (raw edge
# columns in edge-index order), its edge weights (None: unit), the start positions (None:
# the reference's Mulberry32 starts from `seed`), the opening angle and the parameters.

def workload(name: str) -> dict:
    if name == "ba2000":  # preferential attachment, 2000 nodes, m = 5, seed 0; FP64
        s, t = barabasi_albert(2000, 5, seed=0)
        return dict(name=name, n=2000, dim=2, source=s, target=t, weights=None,
                    start=uniform_start(2000, 2, seed=0), seed=0, theta=1.2,
                    precision="fp64", iterations=100)
    if name == "sbm1m":  # 1M-node SBM, 64 Zipf(1.1) blocks, mean degree 10, seed 1; FP32
        n = 1_000_000
        s, t, w = sbm_zipf(n, 64, 1.1, 10.0, seed=1)
        return dict(name=name, n=n, dim=2, source=s, target=t, weights=w,
                    start=uniform_start(n, 2, seed=1), seed=0, theta=1.2, precision="fp32",
                    iterations=500)
    if name == "rgg250k":  # 3D random geometric graph, unit cube, mean degree 8, seed 2
        n = 250_000
        s, t, _ = random_geometric_3d(n, 8.0, 2)
        return dict(name=name, n=n, dim=3, source=s, target=t, weights=None, start=None,
                    seed=2, theta=0.8, precision="fp32", iterations=50)
    if name == "sbm10k":  # 16 equal blocks, p_in 0.01, p_out 0.0005, seed 3 (the adjoint's)
        n = 10_000
        s, t = sbm_equal(n, 16, 0.01, 0.0005, seed=3)
        return dict(name=name, n=n, dim=2, source=s, target=t, weights=None, start=None,
                    seed=0, theta=0.8, precision="fp32", iterations=200)
    raise KeyError(name)
