//! Raw bindings of `libgraph_render_fa2` (`include/graph_render_fa2.h`), the MI355X
//! ForceAtlas2 / Barnes-Hut path, plus the safe pieces a graph-core `Stage` needs: the
//! status codes mapped one to one onto the reference's `StageError` variants
//! (`crates/graph-core/src/stage.rs:35-52`) and an owning session handle.
//!
//! `graph_core_glue.rs` (next to this file) is the graph-core side: the three registry ids
//! as `Stage`s over [`layout_run`].
#![allow(non_camel_case_types)]

use std::ffi::{c_char, CStr, CString};
use std::os::raw::c_int;

// ---------------------------------------------------------------- status codes

pub const GR_OK: c_int = 0;
/// StageError::Capacity: the topology does not fit the u32 index space.
pub const GR_ERR_CAPACITY: c_int = 1;
/// StageError::NonFinite{column}: the run ended with a non-finite coordinate.
pub const GR_ERR_NON_FINITE: c_int = 2;
/// StageError::Param{name, rule}: an argument outside what the stage accepts.
pub const GR_ERR_PARAM: c_int = 3;
/// StageError::Snapshot (the snapshot encoder's refusals).
pub const GR_ERR_SNAPSHOT: c_int = 4;
/// The wasm ABI's InvalidHandle.
pub const GR_ERR_INVALID_HANDLE: c_int = 10;
/// The wasm ABI's UnknownLayoutId.
pub const GR_ERR_UNKNOWN_LAYOUT: c_int = 11;
/// The wasm ABI's ParamOutOfRange (registry/params.rs ParamsError::OutOfRange).
pub const GR_ERR_PARAM_OUT_OF_RANGE: c_int = 20;
/// The wasm ABI's ParamsMalformed (ParamsError::Malformed).
pub const GR_ERR_PARAMS_MALFORMED: c_int = 21;
/// A HIP runtime failure, or no device.
pub const GR_ERR_DEVICE: c_int = 30;
/// A combination this build does not run.
pub const GR_ERR_UNSUPPORTED: c_int = 31;

pub const GR_PREC_FP64: i32 = 0;
pub const GR_PREC_FP32: i32 = 1;
pub const GR_PREC_FP32_KAHAN: i32 = 2;

pub const GR_REP_DEFAULT: i32 = 0;
pub const GR_REP_TREE: i32 = 1;
pub const GR_REP_DENSE: i32 = 2;
pub const GR_REP_PM: i32 = 3;

pub const GR_EXT_EXPORT_TREE: u32 = 1;
pub const GR_EXT_DIRECT_LAUNCH: u32 = 2;
pub const GR_EXT_BLOCK_SUMS: u32 = 4;

pub const GR_NOALLOC_MAX_NODES: u32 = 262_144;
pub const GR_NOALLOC_MAX_SIDE: u32 = 1024;

// ---------------------------------------------------------------- structs

/// `Fa2Params` (layout/forceatlas2/state.rs:20-54) field for field, C layout.
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct gr_fa2_params {
    pub max_iter: u32,
    pub jitter_tolerance: f64,
    pub scaling_ratio: f64,
    pub gravity: f64,
    pub seed: u32,
    pub dim: u32,
}

/// What the reference does not have; a null pointer means its exact semantics.
#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct gr_fa2_ext {
    pub struct_size: u32,
    pub flags: u32,
    pub theta: f64,
    pub precision: i32,
    pub repulsion: i32,
    pub edge_weights: *const f64,
    pub start: *const f64,
    pub pm_grid: u32,
    pub reserved: u32,
}

/// How the kernels are launched (never what they compute).
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub struct gr_fa2_schedule {
    pub walk_block: u32,
    pub walk_xcd: u32,
    pub sort_items: u32,
    pub finish_split: u32,
}

/// `Fa2ForceSimParams` (layout/force/forcesim.rs:104-153), field for field.
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct gr_forcesim_params {
    pub iterations: u32,
    pub scale: f64,
    pub scaling_ratio: f64,
    pub gravity: f64,
    pub jitter_tolerance: f64,
    pub seed: u64,
}

/// `ForceParams` (layout/force/params.rs:26-77), field for field.
#[repr(C)]
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct gr_d3_params {
    pub charge_strength: f64,
    pub theta: f64,
    pub distance_min: f64,
    pub distance_max: f64,
    pub link_distance: f64,
    pub link_strength_scale: f64,
    pub collide_radius: f64,
    pub center_strength: f64,
    pub velocity_decay: f64,
    pub alpha_decay: f64,
    pub alpha_min: f64,
    pub initial_alpha: f64,
}

/// An opaque resident d3 simulation (`struct gr_d3_session`).
#[repr(C)]
pub struct gr_d3_session {
    _private: [u8; 0],
}

/// An opaque device session (`struct gr_fa2_session`).
#[repr(C)]
pub struct gr_fa2_session {
    _private: [u8; 0],
}

#[link(name = "graph_render_fa2")]
extern "C" {
    pub fn gr_fa2_params_default(out: *mut gr_fa2_params);
    pub fn gr_fa2_ext_default(out: *mut gr_fa2_ext);
    pub fn gr_fa2_params_decode(bytes: *const u8, len: usize, out: *mut gr_fa2_params) -> c_int;
    pub fn gr_layout_run(layout_id: *const c_char, n: u32, source: *const u32,
                         target: *const u32, m: usize, params: *const u8, params_len: usize,
                         ext: *const gr_fa2_ext, out_x: *mut f32, out_y: *mut f32,
                         out_z: *mut f32) -> c_int;
    pub fn gr_fa2_session_create(n: u32, source: *const u32, target: *const u32, m: usize,
                                 params: *const gr_fa2_params, ext: *const gr_fa2_ext)
                                 -> *mut gr_fa2_session;
    pub fn gr_fa2_batch_create(ngraphs: u32, node_counts: *const u32, edge_counts: *const usize,
                               source: *const u32, target: *const u32,
                               params: *const gr_fa2_params, ext: *const gr_fa2_ext)
                               -> *mut gr_fa2_session;
    pub fn gr_fa2_session_release(s: *mut gr_fa2_session);
    pub fn gr_fa2_session_step(s: *mut gr_fa2_session, iters: u32, done_out: *mut u32) -> c_int;
    pub fn gr_fa2_session_set_params(s: *mut gr_fa2_session, params: *const gr_fa2_params,
                                     theta: f64) -> c_int;
    pub fn gr_fa2_session_positions(s: *mut gr_fa2_session, out: *mut f64) -> c_int;
    pub fn gr_fa2_session_set_positions(s: *mut gr_fa2_session, input: *const f64) -> c_int;
    pub fn gr_fa2_session_forces(s: *mut gr_fa2_session, parts: u32, out: *mut f64) -> c_int;
    pub fn gr_fa2_session_scalars(s: *mut gr_fa2_session, out7: *mut f64) -> c_int;
    pub fn gr_fa2_session_graph_scalars(s: *mut gr_fa2_session, graph: u32, out7: *mut f64)
                                        -> c_int;
    pub fn gr_fa2_session_graph(s: *mut gr_fa2_session, m_out: *mut u32, lo: *mut u32,
                                hi: *mut u32, mass: *mut f64) -> c_int;
    pub fn gr_fa2_session_tree(s: *mut gr_fa2_session, ncells_out: *mut u32, cap: u32,
                               bounds: *mut f64, skip: *mut u32, start: *mut u32, end: *mut u32,
                               bodies: *mut f64, order: *mut u32) -> c_int;
    pub fn gr_fa2_session_record(s: *mut gr_fa2_session, iters: u32, checkpoint_every: u32,
                                 done_out: *mut u32) -> c_int;
    pub fn gr_fa2_session_adjoint(s: *mut gr_fa2_session, pos_bar: *const f64,
                                  weight_bar: *mut f64, param_bar2: *mut f64,
                                  start_bar: *mut f64) -> c_int;
    pub fn gr_fa2_session_walk_stats(s: *mut gr_fa2_session, out4: *mut u64) -> c_int;
    pub fn gr_fa2_schedule_select(n: u32, m: usize, dim: u32, precision: i32, ngraphs: u32,
                                  out: *mut gr_fa2_schedule) -> c_int;
    pub fn gr_fa2_session_get_schedule(s: *mut gr_fa2_session, out: *mut gr_fa2_schedule)
                                       -> c_int;
    pub fn gr_fa2_session_set_schedule(s: *mut gr_fa2_session, sched: *const gr_fa2_schedule)
                                       -> c_int;
    pub fn gr_fa2_session_autotune(s: *mut gr_fa2_session, reps: u32,
                                   out: *mut gr_fa2_schedule) -> c_int;
    pub fn gr_fa2_session_time(s: *mut gr_fa2_session, stage: *const c_char, reps: u32,
                               ms_out: *mut f64) -> c_int;
    pub fn gr_snapshot_encode(n: u32, node_ids: *const *const c_char, m: u32,
                              edge_ids: *const *const c_char, source: *const u32,
                              target: *const u32, x: *const f32, y: *const f32, z: *const f32,
                              notes: u32, note_code: *const u32, note_index: *const u32,
                              out: *mut u8, cap: usize, len_out: *mut usize) -> c_int;
    pub fn gr_pm_side_for(n: u32) -> u32;
    pub fn gr_pm_charge(n: u32, xy: *const f64, side: u32, charge: f64, alpha: f64,
                        distance_min: f64, distance_max: f64, out: *mut f64,
                        frame_out: *mut f64) -> c_int;
    pub fn gr_pm_charge_noalloc(n: u32, xy: *const f64, side: u32, charge: f64, alpha: f64,
                                distance_min: f64, distance_max: f64, out: *mut f64,
                                frame_out: *mut f64) -> c_int;
    pub fn gr_d3_charge_bh(n: u32, xy: *const f64, theta: f64, charge: f64, alpha: f64,
                           distance_min: f64, distance_max: f64, seed: u32, tick: u32,
                           out: *mut f64) -> c_int;
    pub fn gr_d3_charge_bh_time(n: u32, xy: *const f64, theta: f64, charge: f64, alpha: f64,
                                distance_min: f64, distance_max: f64, seed: u32, tick: u32,
                                reps: u32, out: *mut f64, ms_out: *mut f64) -> c_int;
    pub fn gr_quadtree_node_ids(n: u32, xy: *const f64, cap: u32, ncells_out: *mut u32,
                                node_id: *mut u32) -> c_int;
    pub fn gr_d3_collide(n: u32, xy: *const f64, radius: f64, seed: u32, tick: u32,
                         out: *mut f64) -> c_int;
    pub fn gr_d3_params_default(out: *mut gr_d3_params);
    pub fn gr_pm_layout_run(n: u32, source: *const u32, target: *const u32, strength: *const f64,
                            m: usize, params: *const gr_d3_params, ticks: u32, out_x: *mut f32,
                            out_y: *mut f32, ms_ticks_out: *mut f64) -> c_int;
    pub fn gr_d3_session_with_particle_mesh(s: *mut gr_d3_session) -> c_int;
    pub fn gr_pm_d3_charge(n: u32, xy: *const f64, charge: f64, alpha: f64, distance_min: f64,
                           distance_max: f64, out: *mut f64, frame_out: *mut f64) -> c_int;
    pub fn gr_pm_d3_collide(n: u32, xy: *const f64, radius: f64, seed: u32, tick: u32,
                            out: *mut f64, order_out: *mut u32) -> c_int;
    pub fn gr_d3_session_create(n: u32, source: *const u32, target: *const u32,
                                strength: *const f64, m: usize, params: *const gr_d3_params)
                                -> *mut gr_d3_session;
    pub fn gr_d3_session_step(s: *mut gr_d3_session, ticks: u32, alpha_out: *mut f64,
                              settled_out: *mut i32) -> c_int;
    pub fn gr_d3_session_state(s: *mut gr_d3_session, x: *mut f64, y: *mut f64, vx: *mut f64,
                               vy: *mut f64) -> c_int;
    pub fn gr_d3_session_time(s: *mut gr_d3_session, reps: u32, ms_out: *mut f64,
                              cells_out: *mut u32) -> c_int;
    pub fn gr_d3_session_alpha(s: *const gr_d3_session) -> f64;
    pub fn gr_d3_session_reheat(s: *mut gr_d3_session, alpha: f64) -> c_int;
    pub fn gr_d3_session_set_alpha_target(s: *mut gr_d3_session, target: f64) -> c_int;
    pub fn gr_d3_session_release(s: *mut gr_d3_session);
    pub fn gr_d3_layout_run(n: u32, source: *const u32, target: *const u32, strength: *const f64,
                            m: usize, params: *const gr_d3_params, ticks: u32, out_x: *mut f32,
                            out_y: *mut f32, ms_ticks_out: *mut f64) -> c_int;
    pub fn gr_forcesim_params_default(out: *mut gr_forcesim_params);
    pub fn gr_forcesim_run(n: u32, source: *const u32, target: *const u32, m: usize,
                           params: *const gr_forcesim_params, out_x: *mut f32, out_y: *mut f32,
                           out_z: *mut f32) -> c_int;
    pub fn gr_forcesim_batch(ngraphs: u32, node_counts: *const u32, edge_counts: *const usize,
                             source: *const u32, target: *const u32,
                             params: *const gr_forcesim_params, out_x: *mut f32,
                             out_y: *mut f32, out_z: *mut f32, status: *mut i32) -> c_int;
    pub fn gr_forcesim_batch_time(ngraphs: u32, node_counts: *const u32,
                                  edge_counts: *const usize, source: *const u32,
                                  target: *const u32, params: *const gr_forcesim_params,
                                  out_x: *mut f32, out_y: *mut f32, out_z: *mut f32,
                                  status: *mut i32, reps: u32, ms_out: *mut f64) -> c_int;
    pub fn gr_last_error(name: *mut *const c_char, rule: *mut *const c_char) -> c_int;
    pub fn gr_version() -> *const c_char;
}

// ---------------------------------------------------------------- errors

/// The calling thread's last failure as the library reports it.
#[derive(Clone, Debug, PartialEq, Eq)]
pub struct LastError {
    pub code: c_int,
    pub name: String,
    pub rule: String,
}

pub fn last_error() -> LastError {
    let mut name: *const c_char = std::ptr::null();
    let mut rule: *const c_char = std::ptr::null();
    // SAFETY: both out-pointers are valid; the strings are the library's thread-local copies,
    // read before any other call on this thread.
    let code = unsafe { gr_last_error(&mut name, &mut rule) };
    let text = |p: *const c_char| {
        if p.is_null() {
            String::new()
        } else {
            unsafe { CStr::from_ptr(p) }.to_string_lossy().into_owned()
        }
    };
    LastError { code, name: text(name), rule: text(rule) }
}

/// A refusal in the shape of graph-core's `StageError` (stage.rs:35-52), one variant per
/// status code, with the names and rules as `&'static str` (the form StageError holds):
/// the library's strings are matched against the ones it can produce.
#[derive(Clone, Copy, Debug, PartialEq, Eq)]
pub enum Refusal {
    /// code 1: `StageError::Capacity(CapacityError { what })` (arena.rs:68-77)
    Capacity { what: &'static str },
    /// code 2: `StageError::NonFinite { column }` (forceatlas2.rs:86-88 names "node.x")
    NonFinite { column: &'static str },
    /// codes 3, 20, 21: `StageError::Param { name, rule }`; 20 and 21 carry the registry's
    /// own rules (registry/params.rs:187-202)
    Param { name: &'static str, rule: &'static str },
    /// code 4: the snapshot contract (`StageError::Snapshot`)
    Snapshot { field: &'static str },
    /// code 11: the wasm ABI's UnknownLayoutId
    UnknownLayout,
    /// codes 10, 30, 31: no StageError counterpart (a handle, the device, a combination)
    Device { code: c_int },
}

const PARAM_NAMES: [&str; 16] = ["max_iter", "jitter_tolerance", "scaling_ratio", "gravity",
                                 "seed", "dim", "params", "theta", "precision", "repulsion",
                                 "pm_grid", "edges", "edge_weights", "charge", "xy", "reps"];
const CAPACITY_WHAT: [&str; 6] = ["nodes", "edges", "tree", "n", "side", "cap"];
const SNAPSHOT_FIELDS: [&str; 6] = ["node.id", "edge.id", "note", "note.code", "note.index",
                                    "snapshot"];
const RULE_OUT_OF_RANGE: &str = "inside the range the schema publishes, never clamped";
const RULE_MALFORMED: &str = "one little-endian f64 per published parameter, and no other bytes";

fn static_of(set: &[&'static str], s: &str, other: &'static str) -> &'static str {
    set.iter().copied().find(|&x| x == s).unwrap_or(other)
}

/// Maps a status code (with the thread's last error for its name) onto [`Refusal`].
pub fn refusal(code: c_int) -> Option<Refusal> {
    if code == GR_OK {
        return None;
    }
    let e = last_error();
    Some(match code {
        GR_ERR_CAPACITY => Refusal::Capacity { what: static_of(&CAPACITY_WHAT, &e.name, "nodes") },
        GR_ERR_NON_FINITE => Refusal::NonFinite {
            column: if e.name == "edge.weight" { "edge.weight" } else { "node.x" },
        },
        GR_ERR_PARAM_OUT_OF_RANGE => Refusal::Param {
            name: static_of(&PARAM_NAMES, &e.name, "params"),
            rule: RULE_OUT_OF_RANGE,
        },
        GR_ERR_PARAMS_MALFORMED => Refusal::Param { name: "params", rule: RULE_MALFORMED },
        GR_ERR_PARAM => Refusal::Param {
            name: static_of(&PARAM_NAMES, &e.name, "params"),
            rule: if e.name == "dim" { "2 or 3 coordinates per node" } else { "see gr_last_error" },
        },
        GR_ERR_SNAPSHOT => Refusal::Snapshot { field: static_of(&SNAPSHOT_FIELDS, &e.name, "snapshot") },
        GR_ERR_UNKNOWN_LAYOUT => Refusal::UnknownLayout,
        other => Refusal::Device { code: other },
    })
}

// ---------------------------------------------------------------- safe calls

/// f32 columns of one `gr_layout_run` (Stage::run for the three ids): x, y, and z at dim 3.
pub struct Columns {
    pub dim: usize,
    pub x: Vec<f32>,
    pub y: Vec<f32>,
    pub z: Vec<f32>,
}

/// `Stage::run` of `layout_id` over the raw edge columns, with the registry's wire
/// parameters (empty = the defaults) and an optional extension (None = the reference's
/// semantics, bit for bit).
pub fn layout_run(layout_id: &str, n: u32, source: &[u32], target: &[u32], wire: &[u8],
                  ext: Option<&gr_fa2_ext>) -> Result<Columns, Refusal> {
    if source.len() != target.len() {
        return Err(Refusal::Param { name: "edges", rule: "equal source and target columns" });
    }
    let dim = if layout_id == "layout.forceatlas2.3d" { 3 } else { 2 };
    let id = CString::new(layout_id).map_err(|_| Refusal::UnknownLayout)?;
    let len = n as usize;
    let (mut x, mut y) = (vec![0f32; len], vec![0f32; len]);
    let mut z = if dim == 3 { vec![0f32; len] } else { Vec::new() };
    // SAFETY: every pointer is valid for the lengths passed; the library borrows the inputs
    // for the call and writes exactly n values into each output column.
    let code = unsafe {
        gr_layout_run(id.as_ptr(), n, source.as_ptr(), target.as_ptr(), source.len(),
                      wire.as_ptr(), wire.len(),
                      ext.map_or(std::ptr::null(), |e| e as *const gr_fa2_ext),
                      x.as_mut_ptr(), y.as_mut_ptr(),
                      if dim == 3 { z.as_mut_ptr() } else { std::ptr::null_mut() })
    };
    match refusal(code) {
        None => Ok(Columns { dim, x, y, z }),
        Some(r) => Err(r),
    }
}

/// The many-body constants one d3 charge pass reads (barnes_hut/charge.rs `Ctx`).
#[derive(Clone, Copy, Debug, PartialEq)]
pub struct D3Charge {
    pub theta: f64,
    pub charge: f64,
    pub alpha: f64,
    pub distance_min: f64,
    pub distance_max: f64,
    pub seed: u32,
    pub tick: u32,
}

/// One d3 many-body pass over the quadtree of `(x, y)` on the device (`gr_d3_charge_bh`):
/// the velocity each node gains, in node order.
pub fn d3_charge_bh(x: &[f64], y: &[f64], c: &D3Charge) -> Result<Vec<(f64, f64)>, Refusal> {
    if x.len() != y.len() {
        return Err(Refusal::Param { name: "xy", rule: "equal x and y columns" });
    }
    let n = u32::try_from(x.len()).map_err(|_| Refusal::Capacity { what: "nodes" })?;
    let xy: Vec<f64> = x.iter().zip(y).flat_map(|(&a, &b)| [a, b]).collect();
    let mut out = vec![0f64; xy.len()];
    // SAFETY: xy and out hold 2n values each; the library borrows them for the call.
    let code = unsafe {
        gr_d3_charge_bh(n, xy.as_ptr(), c.theta, c.charge, c.alpha, c.distance_min,
                        c.distance_max, c.seed, c.tick, out.as_mut_ptr())
    };
    match refusal(code) {
        None => Ok(out.chunks_exact(2).map(|d| (d[0], d[1])).collect()),
        Some(r) => Err(r),
    }
}

/// The wire form of `Fa2Params` (registry/tunable.rs:229-240 order): five little-endian f64.
pub fn wire(p: &gr_fa2_params) -> Vec<u8> {
    [f64::from(p.max_iter), p.jitter_tolerance, p.scaling_ratio, p.gravity, f64::from(p.seed)]
        .iter()
        .flat_map(|v| v.to_le_bytes())
        .collect()
}

/// An owning session: released on drop.
pub struct Session {
    raw: *mut gr_fa2_session,
    n: usize,
    dim: usize,
}

impl Session {
    pub fn new(n: u32, source: &[u32], target: &[u32], params: &gr_fa2_params,
               ext: Option<&gr_fa2_ext>) -> Result<Session, Refusal> {
        // SAFETY: inputs valid for the call; a null return reports through gr_last_error.
        let raw = unsafe {
            gr_fa2_session_create(n, source.as_ptr(), target.as_ptr(), source.len(), params,
                                  ext.map_or(std::ptr::null(), |e| e as *const gr_fa2_ext))
        };
        if raw.is_null() {
            return Err(refusal(last_error().code).unwrap_or(Refusal::Device { code: GR_ERR_DEVICE }));
        }
        Ok(Session { raw, n: n as usize, dim: params.dim as usize })
    }

    /// Up to `iters` iterations; returns the iterations applied (an early exit stops short).
    pub fn step(&mut self, iters: u32) -> Result<u32, Refusal> {
        let mut done = 0u32;
        let code = unsafe { gr_fa2_session_step(self.raw, iters, &mut done) };
        refusal(code).map_or(Ok(done), Err)
    }

    /// Row-major n * dim f64 positions.
    pub fn positions(&mut self) -> Result<Vec<f64>, Refusal> {
        let mut out = vec![0f64; self.n * self.dim];
        let code = unsafe { gr_fa2_session_positions(self.raw, out.as_mut_ptr()) };
        refusal(code).map_or(Ok(out), Err)
    }
}

impl Drop for Session {
    fn drop(&mut self) {
        // SAFETY: raw came from gr_fa2_session_create and is released once.
        unsafe { gr_fa2_session_release(self.raw) }
    }
}
