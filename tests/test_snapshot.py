"""The GMSN snapshot encoder (gr_snapshot_encode) against the reference's byte contract:
the pinned 84- and 92-byte examples (docs/contract/binary-layout.md, binary/tests/pinned.rs),
the refusals Snapshot::new makes, and the reference's own sha256 pins of
layout.forceatlas2 at 100 iterations (harness/golden-2d-before.txt), which the oracle's dense
FA2 must reproduce bit for bit once narrowed to f32 and encoded."""
import numpy as np
import pytest

import graph_render_amd as G
from graph_render_amd import snapshot as S
from tests.refgraphs import fa2_snapshot_pins, gate_model

PINNED_84 = bytes.fromhex(
    "474d534e" "00000000" "03000000" "00000000" "01000000" "02000000" "01000000"
    "00000000" "01000000" "03000000" "61626300" "00000000" "01000000" "65000000"
    "00000000" "01000000" "0000803f" "000020c0" "00000000" "0000003f" "00000000")


def tiny(**kw):
    return S.encode(["a", "bc"], ["e"], [0], [1], [1.0, -2.5], [0.0, 0.5], **kw)


def test_pinned_84_byte_example():
    assert len(PINNED_84) == 84
    assert tiny() == PINNED_84


def test_pinned_one_note_example():
    b = tiny(notes=[(3, 0xFFFFFFFF)])
    assert len(b) == 92
    assert b[:80] == PINNED_84[:80]
    assert b[80:] == bytes.fromhex("01000000" "03000000" "ffffffff")


def test_3d_snapshot_is_labelled_0_4_with_a_z_column():
    b = S.encode(["a", "bc"], ["e"], [0], [1], [1.0, -2.5], [0.0, 0.5], z=[2.0, -1.0])
    assert b[8] == 4 and b[14] == 1
    assert b[64:80] == PINNED_84[64:80]
    assert b[80:88] == np.array([2.0, -1.0], "<f4").tobytes()
    assert b[88:] == bytes(4)


@pytest.mark.parametrize("kw, column", [
    (dict(node_ids=["a", "a"]), "node.id"),
    (dict(edge_ids=["e"], target=[2]), "edge.target"),
    (dict(x=[float("nan"), 0.0]), "node.x"),
])
def test_refusals(kw, column):
    args = dict(node_ids=["a", "bc"], edge_ids=["e"], source=[0], target=[1],
                x=[1.0, -2.5], y=[0.0, 0.5])
    args.update(kw)
    with pytest.raises(G.StageError) as e:
        S.encode(**args)
    assert e.value.name == column


def test_notes_must_be_known_and_ascending():
    with pytest.raises(G.StageError):
        tiny(notes=[(6, 0)])
    with pytest.raises(G.StageError):
        tiny(notes=[(3, 0xFFFFFFFF), (1, 0)])
    with pytest.raises(G.StageError):
        tiny(notes=[(1, 5)])  # edge position past m


def test_the_oracle_reproduces_the_reference_fa2_snapshots(oracle):
    pins = fa2_snapshot_pins()
    assert sorted(pins) == list(range(8))
    for seed, want in pins.items():
        n, nodes, edges, s, t = gate_model(oracle, seed)
        st = oracle.Fa2(n, s, t, tree=False)  # Fa2Params::default(): 100 iterations, seed 0
        st.run(100)
        p = st.positions().astype(np.float32)
        got = S.sha256(S.encode(nodes, edges, s, t, p[:, 0], p[:, 1]))
        assert got == want, seed
