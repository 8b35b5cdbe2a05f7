"""Shared helpers of the parity tests: build a GPU scene and an oracle scene
with identical inputs, and compare fields."""
import numpy as np

from campfire_amd import scenes
from oracle.pyoracle import Oracle

FIELDS_STEP = ["T", "Y", "U", "V", "SQ"]


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def make_pair(nx, ny, dx, dt, seed=1, npieces=4, params=None, fp64=False, state=True):
    """(FluidSim on the GPU, FP32 oracle) with identical parameters, pieces,
    geometry (the GPU's, copied into the oracle) and seeded random fields."""
    from campfire_amd.fluid import FluidSim
    g = FluidSim({"nx": nx, "ny": ny, "dt": dt}, dx=dx)
    o = Oracle(nx, ny, dx, dt, fp64=fp64)
    for k, v in (params or {}).items():
        g.set_param(k, v)
        o.set_param(k, v)
    g.reset()
    o.reset()
    if npieces:
        pieces, n = scenes.pile(npieces, seed, nx * dx, spread=nx * dx / 2)
    else:
        pieces, n = np.zeros(64 * 20, np.float32), 0
    g.upload_pieces(pieces, n)
    o.upload_pieces(pieces, n)
    g.write_params()
    o.write_params()
    g.clear_geometry()
    g.encode_geometry()
    o.clear_geometry()
    o.geometry()
    if state:
        st = scenes.random_state(nx, ny, seed)
        for name in ("T", "T2", "Th"):
            put(g, o, name, st["T"])
        for name in ("Y", "Y2", "Yh"):
            put(g, o, name, st["Y"])
        for name in ("U", "U2", "Uh"):
            put(g, o, name, st["U"])
        for name in ("V", "V2", "Vh"):
            put(g, o, name, st["V"])
    return g, o, pieces, n


def put(g, o, name, arr):
    g.write_field(name, arr)
    o.write(name, arr)


def sync_geometry(g, o):
    """Copy the GPU geometry (cellG, fx, fy, faceW) into the oracle."""
    for name in ("cellG", "fx", "fy", "faceW"):
        o.write(name, g.read_field(name))
