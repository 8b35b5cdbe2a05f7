"""The headless field dump for an offline renderer: the buffers that
shaders/render.wgsl binds from the simulation, byte for byte as the device
holds them, with a manifest of bindings, element types and counts."""
import json
import os
import sys

import numpy as np
import pytest

from campfire_amd import scenes

pytestmark = pytest.mark.gpu


def test_render_dump_matches_the_device_buffers(tmp_path):
    from campfire_amd.fluid import FluidSim, RAD_LEN
    nx, ny, dx = 96, 80, 0.004
    g = FluidSim({"nx": nx, "ny": ny, "dt": 0.0016}, dx=dx)
    pieces, n = scenes.pile(4, 3, nx * dx, spread=nx * dx / 3)
    g.reset()
    g.upload_pieces(pieces, n)
    g.write_params()
    for f in range(2):
        g.frame = f
        g.frame_all(4, True)
    g.synchronize()
    man = g.dump_render_buffers(str(tmp_path))
    print("[render dump]", json.dumps(man), flush=True)
    sys.stdout.flush()
    with open(os.path.join(str(tmp_path), "manifest.json")) as fh:
        assert json.load(fh) == man
    nC = nx * ny
    want = {"P": 1, "T": nC, "Y": nC, "SQ": nC, "pieces": 64, "U": (nx + 1) * ny, "V": nx * (ny + 1), "rad": RAD_LEN}
    bindings = {b["name"]: b for b in man["buffers"]}
    assert sorted(b["binding"] for b in man["buffers"]) == [0, 2, 3, 4, 5, 6, 7, 9]
    for name, count in want.items():
        b = bindings[name]
        assert b["elements"] == count, name
        with open(os.path.join(str(tmp_path), name + ".bin"), "rb") as fh:
            raw = fh.read()
        assert len(raw) == b["bytes"], name
        if name == "P":
            assert raw == bytes(g.packed_params())
            assert len(raw) == 160
        else:
            assert np.array_equal(np.frombuffer(raw, "<f4"), g.read_field(name)), name
    assert man["nx"] == nx and man["ny"] == ny and man["frame"] == g.info().frame
    g.close()
