"""The oracle's frozen layout.force.barnes_hut (golden-spiral start, 112 ticks of link,
many-body over the quadtree, center, collide and integration) against the reference's own
65 golden digests (crates/graph-core/src/layout/force/session/tests/golden.rs, copied to
tests/golden/d3_barnes_hut_frozen_sha256.txt): SHA-256 of the f32 x column then y column
on the gate model of seeds 0..64 (2..66 nodes)."""
import hashlib
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    rows = []
    for line in open(os.path.join(HERE, "golden", "d3_barnes_hut_frozen_sha256.txt")):
        if line.startswith("#") or not line.strip():
            continue
        seed, digest = line.split()
        rows.append((int(seed), digest))
    return rows


def digest(x, y):
    return hashlib.sha256(np.asarray(x, np.float32).tobytes() +
                          np.asarray(y, np.float32).tobytes()).hexdigest()


def test_the_oracle_reproduces_every_golden_digest(oracle):
    rows = golden()
    assert len(rows) == 65
    bad = []
    for seed, want in rows:
        n = 2 + seed % 600
        s, t, w = oracle.gate_model(n)
        x, y = oracle.d3sim(n, s, t, w)
        if digest(x, y) != want:
            bad.append(seed)
    assert not bad, bad


@pytest.mark.parametrize("change", [dict(ticks=111), dict(theta=0.5), dict(collide_radius=40.0),
                                    dict(link_strength_scale=0.16), dict(charge=-89.0),
                                    dict(center_strength=0.99)])
def test_negative_controls_miss_the_digests(oracle, change):
    # one tick fewer, and a change to each pass's constant (many-body opening angle and
    # strength, collide radius, link strength, centering): each must move some digest
    rows = golden()[:24]
    hits = 0
    for seed, want in rows:
        n = 2 + seed % 600
        s, t, w = oracle.gate_model(n)
        x, y = oracle.d3sim(n, s, t, w, **change)
        hits += digest(x, y) == want
    assert hits < len(rows)
