"""MI355X-native ForceAtlas2 / Barnes-Hut layout, a drop-in for graph-core's FA2 entries.

The compute path is libgraph_render_fa2.so (hand-written HIP for gfx950 behind a C ABI,
include/graph_render_fa2.h); this package is its Python face.
"""
from .fa2 import (  # noqa: F401
    LAYOUTS,
    CapacityError,
    DeviceError,
    Fa2Ext,
    Fa2Params,
    ForceAtlas2,
    ForceAtlas2BarnesHut,
    Geometry,
    NonFiniteError,
    ParamError,
    Session,
    StageError,
    Topology,
    find,
    run_3d,
    run_params,
    schedule_select,
)
