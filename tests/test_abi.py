"""CPU checks of the C-ABI library: it loads, exports every function include/*.h declares,
and its host-side rules (params wire, error reporting) behave as the reference's. No
compute is called here."""
import ctypes as C
import glob
import os
import re
import struct

import numpy as np
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
        G.run_params("layout.force.spring", e, b"")
    assert err.value.code == A.GR_ERR_UNKNOWN_LAYOUT


def test_registry_lists_the_six_ids():
    assert set(G.LAYOUTS) == {"layout.forceatlas2", "layout.forceatlas2.barnes_hut",
                              "layout.forceatlas2.3d", "layout.forceatlas2.forcesim",
                              "layout.force.barnes_hut", "layout.force.particle_mesh"}
    assert G.find("layout.force.particle_mesh").scale_ceiling == 800_000
    assert G.find("layout.forceatlas2.barnes_hut").scale_ceiling == 250_000
    assert G.find("layout.forceatlas2.forcesim").scale_ceiling == 2_000
    assert G.find("layout.forceatlas2.forcesim").params == ()
    # registry/force.rs:31 FORCE_CEILING; registry/layouts.rs:92-97 LayoutParams::NONE
    assert G.find("layout.force.barnes_hut").scale_ceiling == 100_000
    assert G.find("layout.force.barnes_hut").params == ()


def test_d3_layout_publishes_no_parameters():
    topo = G.Topology(4, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    with pytest.raises(G.StageError) as e:
        G.run_params("layout.force.barnes_hut", topo, b"\0" * 8)
    assert e.value.code == 21
    # of the extensions only the strength column and the opening angle apply
    with pytest.raises(G.StageError) as e:
        G.run_params("layout.force.barnes_hut", topo, b"",
                     G.Fa2Ext(start=np.zeros((4, 2))))
    assert e.value.code == A.GR_ERR_UNSUPPORTED


def test_d3_defaults_are_the_references():
    # layout/force/params.rs:59-76
    from graph_render_amd import barnes_hut as BH
    p = A.GrD3Params()
    A.lib().gr_d3_params_default(C.addressof(p))
    want = (-90.0, 0.9, 1.0, 520.0, 60.0, 0.15, 16.0, 1.0, 1.0 - 0.42, 0.06, 0.001, 1.0)
    assert tuple(getattr(p, f) for f, _ in A.GrD3Params._fields_) == want
    d = BH.ForceParams()
    assert tuple(getattr(d, f) for f, _ in A.GrD3Params._fields_) == want
    assert BH.TICKS == 112


@pytest.mark.parametrize("change", [dict(theta=float("nan")), dict(theta=-1.0),
                                    dict(alpha_decay=float("inf")),
                                    dict(distance_max=-1.0)])
def test_d3_layout_refuses_bad_parameters_before_the_device(change):
    from graph_render_amd import barnes_hut as BH
    topo = G.Topology(3, np.array([0], np.uint32), np.array([1], np.uint32))
    with pytest.raises(G.StageError) as e:
        BH.BarnesHut.run(topo, params=BH.ForceParams(**change))
    assert e.value.code == A.GR_ERR_PARAM


def test_d3_layout_refuses_a_short_strength_column():
    from graph_render_amd import barnes_hut as BH
    topo = G.Topology(3, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    with pytest.raises(G.ParamError):
        BH.BarnesHut.run(topo, strength=[1.0])


def test_forcesim_publishes_no_parameters():
    # registry/forceatlas2_forcesim.rs:67-72 registers it with LayoutParams::NONE: any byte of
    # a wire buffer is malformed, refused before the device is touched
    topo = G.Topology(4, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    with pytest.raises(G.StageError) as e:
        G.run_params("layout.forceatlas2.forcesim", topo, b"\0" * 8)
    assert e.value.code == 21


def test_forcesim_defaults_are_the_references():
    from graph_render_amd import forcesim as FS
    p = A.GrForceSimParams()
    A.lib().gr_forcesim_params_default(C.addressof(p))
    assert (p.iterations, p.scale, p.scaling_ratio, p.gravity, p.jitter_tolerance, p.seed) == \
        (50, 5.0, 2.0, 1.0, 1.0, 1_767_573_729)
    d = FS.Fa2ForceSimParams()
    assert (d.iterations, d.scale, d.seed) == (50, 5.0, 1_767_573_729)


def test_a_bad_dim_is_refused_before_the_device():
    with pytest.raises(G.ParamError) as e:
        G.Session(G.Topology(3, [0, 1], [1, 2]), G.Fa2Params(dim=4))
    assert e.value.name == "dim" and "2 or 3" in e.value.rule


def test_an_edge_past_n_is_refused():
    with pytest.raises(G.ParamError) as e:
        G.Session(G.Topology(3, [0, 1], [1, 3]))
    assert e.value.name == "edges"


# ---- StageError::Capacity through the C ABI (stage.rs:35-37, arena.rs:68-77): refused on
# the counts alone, before any array is read or any device memory is touched


def test_capacity_edges_past_the_index_space():
    src = np.zeros(4, np.uint32)
    x = (C.c_float * 4)()
    m = (1 << 31) + 1  # 2m CSR entries would not fit u32; the arrays are never read
    code = A.lib().gr_layout_run(b"layout.forceatlas2.barnes_hut", 4, src.ctypes.data,
                                 src.ctypes.data, m, None, 0, None, x, x, None)
    c, name, rule = A.last_error()
    assert code == c == A.GR_ERR_CAPACITY and name == "edges"
    assert list(x) == [0.0] * 4  # a refused run leaves no stale output


def test_capacity_nodes_past_the_index_space():
    code = A.lib().gr_layout_run(b"layout.forceatlas2", 0xFFFFFFFF, None, None, 0, None, 0,
                                 None, None, None, None)
    assert code == A.GR_ERR_CAPACITY and A.last_error()[1] == "nodes"
    counts = np.array([0x80000000, 0x80000000], np.uint32)
    edges = np.zeros(2, np.uint64)
    p = A.GrFa2Params()
    A.lib().gr_fa2_params_default(C.byref(p))
    h = A.lib().gr_fa2_batch_create(2, counts.ctypes.data, edges.ctypes.data, None, None,
                                    C.byref(p), None)
    assert not h and A.last_error()[0] == A.GR_ERR_CAPACITY
    with pytest.raises(G.CapacityError):
        G.fa2._raise()


def test_noalloc_charge_pass_refuses_sizes_past_its_array():
    from graph_render_amd import pm
    with pytest.raises(G.CapacityError) as e:
        pm.charge_pass(np.zeros((pm.MAX_NODES_NOALLOC + 1, 2)), 0, allocate=False)
    assert e.value.name == "n"
    with pytest.raises(G.CapacityError) as e:
        pm.charge_pass(np.zeros((10, 2)), 2 * pm.MAX_SIDE_NOALLOC, allocate=False)
    assert e.value.name == "side"


@pytest.mark.parametrize("kw, name", [
    (dict(charge=float("nan")), "charge"),
    (dict(alpha=float("inf")), "charge"),
    (dict(distance_min=-1.0), "charge"),
    (dict(distance_max=float("nan")), "charge"),
    (dict(theta=-0.5), "theta"),
    (dict(theta=float("nan")), "theta"),
])
def test_d3_charge_pass_refuses_bad_parameters_before_the_device(kw, name):
    from graph_render_amd import barnes_hut
    with pytest.raises(G.ParamError) as e:
        barnes_hut.charge_pass(np.zeros((4, 2)), **kw)
    assert e.value.name == name


def test_d3_charge_pass_of_no_points_is_empty():
    from graph_render_amd import barnes_hut
    assert barnes_hut.charge_pass(np.zeros((0, 2))).shape == (0, 2)
    assert barnes_hut.node_ids(np.zeros((0, 2))).shape == (0,)


# ---- the Rust binding (crates/graph-render-fa2-sys) declares what the header declares; no
# Rust toolchain runs here, so its extern block is checked as text


def _rust_externs():
    text = open(os.path.join(ROOT, "crates", "graph-render-fa2-sys", "src", "lib.rs")).read()
    block = text[text.index('extern "C" {'):]
    block = block[:block.index("\n}\n")]
    fns = {}
    for m in re.finditer(r"pub fn (gr_[a-z0-9_]+)\((.*?)\)", block, flags=re.S):
        args = [a for a in m.group(2).split(",") if a.strip()]
        fns[m.group(1)] = len(args)
    return fns


def _header_arity():
    out = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b(gr_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
            args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
            out[m.group(1)] = len(args)
    return out


def test_rust_binding_declares_every_header_function_with_its_arity():
    assert _rust_externs() == _header_arity()


def test_rust_structs_follow_the_header_field_order():
    text = open(os.path.join(ROOT, "crates", "graph-render-fa2-sys", "src", "lib.rs")).read()
    for name, fields in (("gr_fa2_params", A.GrFa2Params._fields_),
                         ("gr_fa2_ext", A.GrFa2Ext._fields_),
                         ("gr_forcesim_params", A.GrForceSimParams._fields_),
                         ("gr_d3_params", A.GrD3Params._fields_)):
        body = text[text.index(f"pub struct {name} {{"):]
        body = body[:body.index("}")]
        got = re.findall(r"pub (\w+):", body)
        assert got == [f[0] for f in fields], name


def test_schedule_select_is_a_pure_choice_from_the_candidate_sets():
    for n, m, dim, prec, ng in ((10, 20, 2, "fp64", 1), (10_000, 54_547, 2, "fp32", 1),
                                (250_000, 978_162, 3, "fp32", 1),
                                (1_000_000, 4_997_552, 2, "fp32", 1),
                                (13_543_864, 51_700_000, 2, "fp32", 4096)):
        a = G.schedule_select(n, m, dim, prec, ng)
        assert a == G.schedule_select(n, m, dim, prec, ng)
        assert a["walk_block"] in (64, 128, 256) and a["walk_xcd"] in (0, 1)
        assert a["sort_items"] in (4, 8, 16) and a["finish_split"] in (1, 2, 4, 8)
