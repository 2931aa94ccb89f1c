"""The benchmark workloads are the recorded graphs: each generator reproduces the edge-list
sha256 committed in tests/golden/workload_edges_sha256.txt."""
import os

import numpy as np
import pytest

from graph_render_amd import synth

PINS = {}
for line in open(os.path.join(os.path.dirname(__file__), "golden",
                              "workload_edges_sha256.txt")):
    if line.strip() and not line.startswith("#"):
        name, digest, m = line.split()
        PINS[name] = (digest, int(m))


@pytest.mark.parametrize("name", sorted(PINS))
def test_workload_edges_are_the_recorded_ones(name):
    w = synth.workload(name)
    assert (synth.digest(w["source"], w["target"]), w["source"].size) == PINS[name]
    assert w["source"].max() < w["n"] and w["target"].max() < w["n"]


def test_ba2000_is_preferential_attachment_with_m_5():
    w = synth.workload("ba2000")
    s, t = w["source"].astype(np.int64), w["target"].astype(np.int64)
    assert s.size == 5 * (2000 - 5) and (t < s).all()  # each new node links back m times
    deg = np.bincount(np.concatenate([s, t]), minlength=2000)
    assert deg[:50].mean() > 4 * deg[-500:].mean()  # early nodes gather degree
    assert np.array_equal(np.unique(s), np.arange(5, 2000))
