"""Writes tests/golden/pm_d3_mesh1k.npz from the reference's particle-mesh fixtures
fixtures/gpu/mesh-1k-{start,settled}.gmfx (graph-cli emit-gpu-fixtures; the layout is the
reference's fixtures/gpu/README.md): for each state, the node positions, the frame the
reference placed (step, h, origin, cells, reach) and the velocity each pass adds from rest at
alpha 1 (delta_charge of the mesh, delta_link, delta_collide of its hashed grid at tick 0);
plus the collapsed simple edges with their strengths (shared by both states) and the
forward twiddle table of libm's cos and sin, and the SHA-256 of each state's kernel spectrum
(its re column's little-endian f64 bytes, then its im column's). Only inputs and expected
outputs are kept. Run here, where the reference is mounted:
    python tests/golden/make_pm_d3_fixture.py /root/reference
"""
import hashlib
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
    off += 8
    lo = np.frombuffer(b, "<u4", m, off)
    hi = np.frombuffer(b, "<u4", m, off + 4 * m)
    strength = np.frombuffer(b, "<f8", m, off + 8 * m)
    off += 16 * m

    def pair(count):
        nonlocal off
        a = np.frombuffer(b, "<f8", count, off)
        c = np.frombuffer(b, "<f8", count, off + 8 * count)
        off += 16 * count
        return np.stack([a, c], 1)

    xy = pair(n)
    twiddle = pair(side)
    spectrum = hashlib.sha256(b[off:off + 16 * side * side]).hexdigest()
    off += 16 * side * side
    link, charge, collide = pair(n), pair(n), pair(n)
    assert off == len(b)
    return {"xy": xy, "side": side, "step": step, "h": h, "origin": (ox, oy),
            "cells": cells, "reach": reach, "delta_charge": charge, "delta_link": link,
            "delta_collide": collide, "twiddle": twiddle, "edges": (lo, hi, strength),
            "spectrum_sha256": spectrum}


def main(ref_root):
    out = {}
    for state in ("start", "settled"):
        f = load(os.path.join(ref_root, "fixtures", "gpu", f"mesh-1k-{state}.gmfx"))
        out[f"{state}_xy"] = f["xy"]
        out[f"{state}_delta_charge"] = f["delta_charge"]
        out[f"{state}_delta_link"] = f["delta_link"]
        out[f"{state}_delta_collide"] = f["delta_collide"]
        out[f"{state}_spectrum_sha256"] = np.array(f["spectrum_sha256"])
        out["twiddle"] = f["twiddle"]
        out["edge_lo"], out["edge_hi"], out["edge_strength"] = f["edges"]
        out[f"{state}_frame"] = np.array([f["step"], f["h"], f["origin"][0], f["origin"][1],
                                          f["cells"], f["reach"], f["side"]], np.float64)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "pm_d3_mesh1k.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
