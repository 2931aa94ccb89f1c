"""The graph-motor snapshot's binary face (GMSN) of a layout, through gr_snapshot_encode.

A layout's output in the byte form graph-core hosts read (the reference's
docs/contract/binary-layout.md, written there by graph-contract's Snapshot::to_bytes):
Point nodes, Line edges, id tables, endpoints, f32 coordinates, notes.
"""
from __future__ import annotations

import ctypes as C
import hashlib
from typing import Optional, Sequence

import numpy as np

from . import _abi as A
from .fa2 import Geometry, _check


def _strings(ids: Sequence[str]):
    raw = [s.encode("utf-8") for s in ids]
    arr = (C.c_char_p * max(len(raw), 1))(*raw)
    return arr, raw


def encode(node_ids: Sequence[str], edge_ids: Sequence[str], source, target, x, y,
           z=None, notes: Sequence[tuple] = ()) -> bytes:
    """The snapshot's bytes: node/edge ids, edge endpoints (node positions), f32 columns
    (z makes it a 3D, format 0.4, snapshot) and (code, index) notes in ascending order."""
    n, m = len(node_ids), len(edge_ids)
    nids, keep_n = _strings(node_ids)
    eids, keep_e = _strings(edge_ids)
    src = np.ascontiguousarray(source, np.uint32)
    dst = np.ascontiguousarray(target, np.uint32)
    cols = [np.ascontiguousarray(c, np.float32) for c in (x, y)]
    zc = np.ascontiguousarray(z, np.float32) if z is not None else None
    for c in cols + ([zc] if zc is not None else []):
        if c.size != n:
            raise ValueError("every node column holds n values")
    if src.size != m or dst.size != m:
        raise ValueError("every edge column holds m values")
    codes = np.array([c for c, _ in notes], np.uint32)
    index = np.array([i for _, i in notes], np.uint32)
    size = C.c_size_t(0)
    lib = A.lib()

    def call(out, cap):
        return lib.gr_snapshot_encode(
            n, nids, m, eids, src.ctypes.data if m else None, dst.ctypes.data if m else None,
            cols[0].ctypes.data if n else None, cols[1].ctypes.data if n else None,
            zc.ctypes.data if zc is not None else None, len(notes),
            codes.ctypes.data if len(notes) else None, index.ctypes.data if len(notes) else None,
            out, cap, C.byref(size))

    _check(call(None, 0))
    buf = (C.c_uint8 * size.value)()
    _check(call(buf, size.value))
    return bytes(buf)


def of_geometry(node_ids: Sequence[str], edge_ids: Sequence[str], source, target,
                geometry: Geometry, notes: Sequence[tuple] = ()) -> bytes:
    """The snapshot of a layout's Geometry laid over its topology's ids and endpoints."""
    return encode(node_ids, edge_ids, source, target, geometry.x, geometry.y,
                  geometry.z if geometry.dim == 3 else None, notes)


def sha256(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()
