"""The N-API addon and the JS FluidSim shim (campfire_amd/js): exports, the
reference's method names, and error behaviour.  The GPU test drives one app
frame from Node and checks it against the same frame through the Python
binding of the same library."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "campfire_amd", "js")
ADDON = os.path.join(JS, "campfire_addon.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the addon is not available")

# the public methods of FluidSim in js/fluid.js:26-356
REF_METHODS = ["init", "reset", "setParam", "writeParams", "uploadPieces", "clearGeometry", "encodeGeometry",
               "encodeRadiation", "encodeStep", "encodeStepProfiled", "encodeReadback", "afterSubmit",
               "readBuffer"]


def node(script, timeout=120):
    r = subprocess.run(["node", "--input-type=module", "-e", script], capture_output=True, text=True,
                       timeout=timeout, cwd=JS)
    if r.returncode != 0:
        raise AssertionError(r.stderr)
    return r.stdout


def test_addon_exports_and_errors():
    out = node("""
import { createRequire } from 'module';
const require = createRequire('%s');
const a = require('./campfire_addon.node');
const names = Object.getOwnPropertyNames(a);
let bad = null;
try { a.create(34, 34, 0.01, 0.002); } catch (e) { bad = [e.code, e.message]; }
console.log(JSON.stringify({names, bad}));
""" % (JS + "/"))
    r = json.loads(out)
    for n in ("create", "destroy", "reset", "setParam", "writeParams", "uploadPieces", "clearGeometry",
              "geometry", "step", "radiation", "encodeReadback", "takeReadback", "readField", "info"):
        assert n in r["names"], n
    assert r["bad"][0] == "-2" and "coarsest level too large" in r["bad"][1]


def test_shim_keeps_reference_interface():
    out = node("""
import * as F from './fluid_hip.mjs';
const s = new F.FluidSim(null, {nx: 160, ny: 224, dt: 0.004});
const methods = Object.getOwnPropertyNames(F.FluidSim.prototype);
console.log(JSON.stringify({methods, consts: [F.MAX_PIECES, F.PIECE_FLOATS, F.ACC_GLOBAL, F.ACC_LEN,
  F.E_SCALE, F.M_SCALE, F.G_SCALE, F.N_FLOOR, F.RAD_LEN], dx: s.dx, height: s.height}));
""")
    r = json.loads(out)
    for m in REF_METHODS:
        assert m in r["methods"], m
    assert r["consts"] == [64, 20, 512, 528, 1e6, 1e12, 1e3, 64, 320]  # js/fluid.js:5-11
    assert abs(r["dx"] - 0.7 / 160) < 1e-15 and abs(r["height"] - 0.7 / 160 * 224) < 1e-12


@pytest.mark.gpu
def test_node_frame_matches_python_binding():
    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim
    pieces, n = scenes.pile(5, 9, 128 * 0.004)
    out = node("""
import { FluidSim } from './fluid_hip.mjs';
(async () => {
const pieces = new Float32Array(%s);
const s = new FluidSim(null, {nx: 128, ny: 128, dt: 0.002, dx: 0.004});
await s.init();
s.uploadPieces(pieces, %d);
s.writeParams(); s.clearGeometry(null); s.encodeGeometry(null);
for (let k = 0; k < 3; k++) s.encodeStep(null);
s.encodeRadiation(null); s.encodeReadback(null);
const r = await new Promise(res => s.afterSubmit((x, tag) => res(x), 'tag'));
const T = await s.readBuffer('T');
let sum = 0; for (const v of T) sum += v;
console.log(JSON.stringify({acc: Array.from(r.acc.slice(512, 515)), Tsum: sum, time: s.time}));
s.destroy();
})().catch(e => { console.error(e); process.exit(1); });
""" % (json.dumps(pieces.tolist()), n), timeout=300)
    r = json.loads(out.strip().splitlines()[-1])
    g = FluidSim({"nx": 128, "ny": 128, "dt": 0.002}, dx=0.004)
    g.upload_pieces(pieces, n)
    g.write_params()
    g.clear_geometry()
    g.encode_geometry()
    g.step(3)
    g.encode_radiation()
    rb = g.readback_now()
    T = g.read_field("T")
    assert list(rb["acc"][512:515]) == r["acc"]
    assert abs(float(np.sum(T, dtype=np.float64)) - r["Tsum"]) < 1e-6 * abs(r["Tsum"])
    assert abs(r["time"] - 0.006) < 1e-12
