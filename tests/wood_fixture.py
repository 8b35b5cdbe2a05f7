"""Loader and replay of the wood-coupling fixture (tests/golden/wood_coupling.npz,
written by tests/golden/make_wood_coupling.py): the per-frame pieces the
reference's CPU models produced, and the gas frame that consumed them."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "wood_coupling.npz")
SCENE = os.path.join(HERE, "golden", "default_scene.npz")


def load():
    z = np.load(PATH, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["params"] = json.loads(str(d["params"]))
    return d


def default_scene():
    z = np.load(SCENE, allow_pickle=False)
    return z["pieces"].copy(), int(z["n"])


def frame_pieces(d, f):
    """Frame f's pieces as the full 64 x 20 upload array."""
    p = np.zeros(64 * 20, np.float32)
    k = d["pieces"].shape[1]
    p[:k] = d["pieces"][f]
    return p


def replay(sim, d, frames, on_frame=None):
    """Runs `frames` app frames on `sim` (FluidSim or the oracle) from the
    fixture's pieces; returns the readbacks (acc, rad, faceW) cut to the
    fixture's lengths."""
    out = []
    for f in range(frames):
        sim.upload_pieces(frame_pieces(d, f), int(d["n"][f]))
        sim.set_param("frame", f)
        sim.write_params()
        sim.clear_geometry()
        if hasattr(sim, "encode_geometry"):
            sim.encode_geometry()
            sim.step(int(d["nsteps"][f]))
            sim.encode_radiation()
            r = sim.readback_now()
            acc, rad, fw = r["acc"], r["rad"], r["faceW"]
        else:
            sim.geometry()
            sim.step(int(d["nsteps"][f]))
            sim.radiation()
            acc, rad, fw = sim.readback()
        out.append((acc.copy(), rad[:d["rad"].shape[1]].copy(), fw[:d["faceW"].shape[1]].copy()))
        if on_frame:
            on_frame(f, out[-1])
    return out
