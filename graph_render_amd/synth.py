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
    """Random geometric graph in the unit cube: radius r with n*(4/3)*pi*r^3 = mean_degree."""
    rng = np.random.default_rng(seed)
    p = rng.random((n, 3))
    r = (mean_degree / (n * 4.0 / 3.0 * np.pi)) ** (1.0 / 3.0)
    g = max(1, int(1.0 / r))
    cell = np.minimum((p * g).astype(np.int64), g - 1)
    cid = (cell[:, 0] * g + cell[:, 1]) * g + cell[:, 2]
    order = np.argsort(cid, kind="stable")
    cid_sorted = cid[order]
    starts = np.searchsorted(cid_sorted, np.arange(g ** 3))
    ends = np.searchsorted(cid_sorted, np.arange(g ** 3), side="right")
    src, dst = [], []
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    for cx in range(g):
        for cy in range(g):
            for cz in range(g):
                a = order[starts[(cx * g + cy) * g + cz]:ends[(cx * g + cy) * g + cz]]
                if a.size == 0:
                    continue
                for dx, dy, dz in offs:
                    nx_, ny_, nz_ = cx + dx, cy + dy, cz + dz
                    if not (0 <= nx_ < g and 0 <= ny_ < g and 0 <= nz_ < g):
                        continue
                    c = (nx_ * g + ny_) * g + nz_
                    b = order[starts[c]:ends[c]]
                    if b.size == 0:
                        continue
                    d2 = ((p[a, None, :] - p[None, b, :]) ** 2).sum(-1)
                    ii, jj = np.nonzero((d2 <= r * r))
                    sel = a[ii] < b[jj]
                    src.append(a[ii][sel])
                    dst.append(b[jj][sel])
    return (np.concatenate(src).astype(np.uint32), np.concatenate(dst).astype(np.uint32), p)
