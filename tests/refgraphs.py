"""Graph families the reference's own tests use, rebuilt from their definitions."""
import numpy as np


def line(n):
    """forceatlas2/tests.rs:10-16: n{i-1} -> n{i}."""
    return np.arange(0, max(n - 1, 0), dtype=np.uint32), np.arange(1, n, dtype=np.uint32)


def bh_graph(n):
    """forceatlas2/state/barnes_hut/tests.rs:9-20: node i linked from i/2 and (7i+3)%i."""
    src, dst = [], []
    for i in range(1, n):
        for j in (i // 2, (i * 7 + 3) % i):
            src.append(j)
            dst.append(i)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


def errors(approx, exact):
    """barnes_hut/tests.rs:34-47: (per-node rms of |d|/|e|, whole-field |d|/|e|)."""
    d2 = ((approx - exact) ** 2).sum(1)
    e2 = (exact ** 2).sum(1)
    return float(np.sqrt((d2 / e2).mean())), float(np.sqrt(d2.sum() / e2.sum()))


def gate_model(oracle, seed):
    """The reference gate model seeded_model(seed, 2 + seed % 600, 8) as graph-cli's
    snapshot builds it (stage/topology.rs:19-35, synthetic.rs:92-170): node ids
    bench:db-{i % 8}:{i}, edge ids bench-e-{k}, endpoints from the oracle's restatement of
    synthetic_edges (the seed's remix redraws sources and kinds, not endpoints)."""
    n = 2 + seed % 600
    s, t = oracle.gate_edges(n)
    nodes = [f"bench:db-{i % 8}:{i}" for i in range(n)]
    edges = [f"bench-e-{k}" for k in range(len(s))]
    return n, nodes, edges, s, t


def fa2_snapshot_pins():
    """seed -> sha256 of layout.forceatlas2's snapshot (tests/golden/fa2-snapshot-sha256.txt)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "fa2-snapshot-sha256.txt")
    pins = {}
    for line in open(path):
        if line.startswith("layout.forceatlas2 "):
            _, seed, digest = line.split()
            pins[int(seed)] = digest
    return pins
