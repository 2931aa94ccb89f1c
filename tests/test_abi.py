"""CPU checks of the C-ABI library: it loads, exports every function include/*.h declares,
and its host-side rules (params wire, error reporting) behave as the reference's. No
compute is called here."""
import ctypes as C
import glob
import os
import re
import struct

import pytest

import graph_render_amd as G
from graph_render_amd import _abi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gr_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_function_is_exported():
    L = A.lib()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(L, name), name
    assert set(A.SIGNATURES) == names


def test_version():
    assert b"gfx950" in A.lib().gr_version()


def test_defaults_are_fa2_params_default():
    p = A.GrFa2Params()
    A.lib().gr_fa2_params_default(C.byref(p))
    assert (p.max_iter, p.jitter_tolerance, p.scaling_ratio, p.gravity, p.seed, p.dim) == \
        (100, 1.0, 2.0, 1.0, 0, 2)


def test_empty_wire_is_the_default():
    assert G.Fa2Params.from_wire(b"") == G.Fa2Params()


def test_wire_round_trip():
    p = G.Fa2Params(max_iter=250, jitter_tolerance=0.5, scaling_ratio=3.0, gravity=0.25,
                    seed=4294967295)
    assert G.Fa2Params.from_wire(p.to_wire()) == p


@pytest.mark.parametrize("vals,name", [
    ((0, 1, 2, 1, 0), "max_iter"),
    ((10001, 1, 2, 1, 0), "max_iter"),
    ((1.5, 1, 2, 1, 0), "max_iter"),
    ((100, -0.1, 2, 1, 0), "jitter_tolerance"),
    ((100, 1, float("nan"), 1, 0), "scaling_ratio"),
    ((100, 1, 2, float("inf"), 0), "gravity"),
    ((100, 1, 2, 1, 4294967296.0), "seed"),
    ((100, 1, 2, 1, 0.5), "seed"),
])
def test_out_of_range_is_refused_never_clamped(vals, name):
    """registry/params.rs:187-202 and graph-contract params.rs:154-170."""
    with pytest.raises(G.ParamError) as e:
        G.Fa2Params.from_wire(struct.pack("<5d", *vals))
    assert e.value.code == A.GR_ERR_PARAM_OUT_OF_RANGE and e.value.name == name
    assert "never clamped" in e.value.rule


@pytest.mark.parametrize("data", [b"\x00", b"\x00" * 39, b"\x00" * 41, b"\x00" * 48])
def test_malformed_buffers_are_refused(data):
    with pytest.raises(G.ParamError) as e:
        G.Fa2Params.from_wire(data)
    assert e.value.code == A.GR_ERR_PARAMS_MALFORMED


def test_unknown_layout_id():
    e = G.Topology(0, [], [])
    with pytest.raises(G.StageError) as err:
        G.run_params("layout.force.barnes_hut", e, b"")
    assert err.value.code == A.GR_ERR_UNKNOWN_LAYOUT


def test_registry_lists_the_three_ids():
    assert set(G.LAYOUTS) == {"layout.forceatlas2", "layout.forceatlas2.barnes_hut",
                              "layout.forceatlas2.3d"}
    assert G.find("layout.forceatlas2.barnes_hut").scale_ceiling == 250_000


def test_a_bad_dim_is_refused_before_the_device():
    with pytest.raises(G.ParamError) as e:
        G.Session(G.Topology(3, [0, 1], [1, 2]), G.Fa2Params(dim=4))
    assert e.value.name == "dim" and "2 or 3" in e.value.rule


def test_an_edge_past_n_is_refused():
    with pytest.raises(G.ParamError) as e:
        G.Session(G.Topology(3, [0, 1], [1, 3]))
    assert e.value.name == "edges"
