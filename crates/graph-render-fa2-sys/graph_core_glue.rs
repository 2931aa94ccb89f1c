//! The graph-core side of the device path: drop this file into graph-core as
//! `crates/graph-core/src/layout/forceatlas2/device.rs` (with `graph-render-fa2-sys` as a
//! dependency behind a `device` feature) and point registry rows at it.
//!
//! Each `Stage` here is the reference's own stage (layout/forceatlas2.rs:32-78) with its
//! `run` going through `gr_layout_run`: the same ids, parameters, wire schema and
//! `StageError` variants. The ids are kept apart from the CPU rows: per the compute-tier
//! rule (docs/decisions/compute-tiers.md, rule 5) a deployment registers these under an
//! explicitly requested id, or swaps `LAYOUTS[i].run` itself; nothing substitutes silently.

use graph_render_fa2_sys as sys;

use super::Fa2Params;
use crate::arena::CapacityError;
use crate::index::Topology;
use crate::layout::Geometry;
use crate::stage::{Stage, StageError};

fn to_sys(p: &Fa2Params) -> sys::gr_fa2_params {
    sys::gr_fa2_params {
        max_iter: p.max_iter,
        jitter_tolerance: p.jitter_tolerance,
        scaling_ratio: p.scaling_ratio,
        gravity: p.gravity,
        seed: p.seed,
        dim: p.dim as u32,
    }
}

/// One status code, one StageError variant (stage.rs:35-52). A device failure has no
/// variant of its own in StageError; it is reported as a refusal of the parameter `device`.
fn stage_error(r: sys::Refusal) -> StageError {
    match r {
        sys::Refusal::Capacity { what } => StageError::Capacity(CapacityError { what }),
        sys::Refusal::NonFinite { column } => StageError::NonFinite { column },
        sys::Refusal::Param { name, rule } => StageError::Param { name, rule },
        sys::Refusal::Snapshot { field } => StageError::Param { name: field, rule: "the snapshot contract" },
        sys::Refusal::UnknownLayout => StageError::Param { name: "layout_id", rule: "a registered layout id" },
        sys::Refusal::Device { .. } => StageError::Param {
            name: "device",
            rule: "a visible MI355X (gfx950) that runs the path; see gr_last_error",
        },
    }
}

fn run_id(id: &str, t: &Topology, p: &Fa2Params) -> Result<Geometry, StageError> {
    let edges = t.edges();
    let cols = sys::layout_run(id, t.node_count(), &edges.source, &edges.target,
                               &sys::wire(&to_sys(p)), None)
        .map_err(stage_error)?;
    Ok(Geometry::points(cols.dim, cols.x, cols.y, cols.z))
}

/// `layout.forceatlas2` on the device (the dense arm; FP64, the reference's summation
/// order, so the geometry is the CPU stage's bit for bit).
pub struct ForceAtlas2Device;

impl Stage for ForceAtlas2Device {
    type Params = Fa2Params;
    const ID: &'static str = "layout.forceatlas2";

    fn run(topology: &Topology, params: &Self::Params) -> Result<Geometry, StageError> {
        if !(2..=3).contains(&params.dim) {
            return Err(StageError::Param { name: "dim", rule: "2 or 3 coordinates per node" });
        }
        let id = if params.dim == 3 { "layout.forceatlas2.3d" } else { "layout.forceatlas2" };
        run_id(id, topology, params)
    }
}

/// `layout.forceatlas2.barnes_hut` on the device: the quadtree at theta 0.8, dim pinned at 2
/// as the CPU stage pins it (forceatlas2.rs:55-59).
pub struct ForceAtlas2BarnesHutDevice;

impl Stage for ForceAtlas2BarnesHutDevice {
    type Params = Fa2Params;
    const ID: &'static str = "layout.forceatlas2.barnes_hut";

    fn run(topology: &Topology, params: &Self::Params) -> Result<Geometry, StageError> {
        let p = Fa2Params { dim: 2, ..*params };
        run_id(Self::ID, topology, &p)
    }
}

/// `layout.forceatlas2.3d` on the device (run_3d: the dense arm with dim forced to 3).
pub fn run_3d_device(topology: &Topology, params: &Fa2Params) -> Result<Geometry, StageError> {
    let p = Fa2Params { dim: 3, ..*params };
    run_id("layout.forceatlas2.3d", topology, &p)
}

/// `layout.forceatlas2.forcesim` on the device (forcesim.rs:155-163): the same `Fa2ForceSimParams`
/// field for field; the registry row runs it with the defaults.
pub struct ForceAtlas2ForceSimDevice;

impl Stage for ForceAtlas2ForceSimDevice {
    type Params = crate::layout::force::Fa2ForceSimParams;
    const ID: &'static str = "layout.forceatlas2.forcesim";

    fn run(topology: &Topology, p: &Self::Params) -> Result<Geometry, StageError> {
        let edges = topology.edges();
        let n = topology.node_count();
        let c = sys::gr_forcesim_params {
            iterations: p.iterations,
            scale: p.scale,
            scaling_ratio: p.scaling_ratio,
            gravity: p.gravity,
            jitter_tolerance: p.jitter_tolerance,
            seed: p.seed,
        };
        let (mut x, mut y, mut z) = (vec![0f32; n as usize], vec![0f32; n as usize],
                                     vec![0f32; n as usize]);
        // SAFETY: the columns hold n values each; the edge columns are borrowed for the call.
        let code = unsafe {
            sys::gr_forcesim_run(n, edges.source.as_ptr(), edges.target.as_ptr(),
                                 edges.source.len(), &c, x.as_mut_ptr(), y.as_mut_ptr(),
                                 z.as_mut_ptr())
        };
        if let Some(r) = sys::refusal(code) {
            return Err(stage_error(r));
        }
        Ok(Geometry::in_space(
            graph_contract::geometry::NodeGeometry::Point { x, y },
            graph_contract::geometry::EdgeGeometry::Line,
            Vec::new(),
            z,
        ))
    }
}

