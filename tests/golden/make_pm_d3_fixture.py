"""Writes tests/golden/pm_d3_mesh1k.npz from the reference's particle-mesh fixtures
fixtures/gpu/mesh-1k-{start,settled}.gmfx (graph-cli emit-gpu-fixtures; the layout is the
reference's fixtures/gpu/README.md): for each state, the node positions, the frame the
reference placed (step, h, origin, cells, reach) and the velocity its charge pass adds from
rest at alpha 1 (delta_charge). Only inputs and expected outputs are kept; twiddles, the
spectrum and the link/collide columns are left out. Run here, where the reference is mounted:
    python tests/golden/make_pm_d3_fixture.py /root/reference
"""
import os
import struct
import sys

import numpy as np


def load(path):
    b = open(path, "rb").read()
    magic, major, minor, n, m, side, state, flags, step, pad = struct.unpack_from("<4sIIIIIIIiI", b, 0)
    assert magic == b"GMFX" and major == 1, (magic, major)
    h, ox, oy = struct.unpack_from("<ddd", b, 40)
    off = 64
    cells, reach = struct.unpack_from("<II", b, off)
    off += 8 + 16 * m
    px = np.frombuffer(b, "<f8", n, off)
    py = np.frombuffer(b, "<f8", n, off + 8 * n)
    off += 16 * n + 16 * side + 16 * side * side + 16 * n  # twiddles, spectrum, link deltas
    cx = np.frombuffer(b, "<f8", n, off)
    cy = np.frombuffer(b, "<f8", n, off + 8 * n)
    off += 16 * n + 16 * n  # charge deltas, collide deltas
    assert off == len(b)
    return {"xy": np.stack([px, py], 1), "side": side, "step": step, "h": h, "origin": (ox, oy),
            "cells": cells, "reach": reach, "delta_charge": np.stack([cx, cy], 1)}


def main(ref_root):
    out = {}
    for state in ("start", "settled"):
        f = load(os.path.join(ref_root, "fixtures", "gpu", f"mesh-1k-{state}.gmfx"))
        out[f"{state}_xy"] = f["xy"]
        out[f"{state}_delta_charge"] = f["delta_charge"]
        out[f"{state}_frame"] = np.array([f["step"], f["h"], f["origin"][0], f["origin"][1],
                                          f["cells"], f["reach"], f["side"]], np.float64)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "pm_d3_mesh1k.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
