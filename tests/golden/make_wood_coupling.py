"""Writes the wood-coupling fixtures (run in a container that holds the
reference checkout; the GPU machine only reads the files):

  default_scene.npz   the reference app's pre-lit default scene (js/main.js:132-155)
                      as uploadPieces packs it (js/main.js:190-208), 16 pieces
  wood_coupling.npz   48 app frames of that scene in the coarse mode (160x224,
                      dx = 0.7/160, dt = 4 ms, js/config.js:7): per frame the pieces
                      the reference's rigid-body, contact and wood models produce
                      from the previous frame's readback (tests/golden/wood_driver.mjs
                      runs js/rigid.js, js/contact.js and js/wood.js under Node), the
                      gas steps of that frame, and the readback {acc, rad, faceW} of
                      the FP32 oracle's frame (geometry, the steps, radiation)

The gas side is the FP32 oracle with the app's control mapping
(js/main.js:82-95 at the defaults: 20.9 % O2 by volume, depth 0.5 m).  The
number of gas steps per frame follows js/main.js:252-257 at 60 frames/s.

    python tests/golden/make_wood_coupling.py [--ref /root/reference] [--frames 48]
"""
import argparse
import json
import math
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

MODE = {"nx": 160, "ny": 224, "dt": 0.004}
SCENE_WIDTH = 0.7


def app_params():
    """The FluidSim parameters the app sets at its control defaults (js/main.js:82-95)."""
    X = 20.9 / 100
    YO = 32 * X / (32 * X + 28.01 * (1 - X))
    return {"YOamb": YO, "viscMult": 1.0, "diffMult": 1.0, "reactB": 2.0e10, "sootFormA": 6.0e6, "charA": 6.6,
            "depth": 0.5}


def steps_per_frame(frames, dt, fps=60.0):
    """Gas steps of each frame as js/main.js:252-257 counts them at a steady frame rate."""
    acc, out, cap = 0.0, [], math.ceil(0.034 / dt)
    for _ in range(frames):
        acc += 1.0 / fps
        n = math.floor(acc / dt)
        if n > cap:
            n, acc = cap, 0.0
        else:
            acc -= n * dt
        out.append(n)
    return out


class Driver:
    def __init__(self, ref):
        self.p = subprocess.Popen(["node", os.path.join(HERE, "wood_driver.mjs"), ref], stdin=subprocess.PIPE,
                                  stdout=subprocess.PIPE, text=True)

    def ask(self, msg):
        self.p.stdin.write(json.dumps(msg) + "\n")
        self.p.stdin.flush()
        r = json.loads(self.p.stdout.readline())
        if "error" in r:
            raise RuntimeError(r["error"])
        return r

    def close(self):
        self.p.stdin.write(json.dumps({"cmd": "quit"}) + "\n")
        self.p.stdin.flush()
        self.p.wait(timeout=30)


def run_frame(o, pieces, n, frame, nsteps):
    o.upload_pieces(pieces, n)
    o.set_param("frame", frame)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    o.step(nsteps)
    o.radiation()
    return o.readback()


def main():
    from oracle.pyoracle import Oracle
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--frames", type=int, default=48)
    a = ap.parse_args()
    nx, ny, dt = MODE["nx"], MODE["ny"], MODE["dt"]
    dx = SCENE_WIDTH / nx
    drv = Driver(a.ref)
    scene = drv.ask({"cmd": "scene"})
    np.savez(os.path.join(HERE, "default_scene.npz"), pieces=np.array(scene["pieces"], np.float32),
             n=np.int32(scene["n"]))
    print("default scene: %d pieces in %d slots" % (scene["alive"], scene["n"]))
    o = Oracle(nx, ny, dx, dt)
    for k, v in app_params().items():
        o.set_param(k, v)
    o.reset()
    nsteps = steps_per_frame(a.frames, dt)
    P, N, ACC, RAD, FW, HRR = [], [], [], [], [], []
    rb = None
    for f in range(a.frames):
        # the readback of the previous frame arrives before this frame's CPU work
        r = drv.ask({"cmd": "frame", "simDt": nsteps[f] * dt, "readback": rb,
                     "rbDt": nsteps[f - 1] * dt if f > 0 else 0.0})
        pieces = np.array(r["pieces"], np.float32)
        acc, rad, fw = run_frame(o, pieces, r["n"], f, nsteps[f])
        rb = {"acc": acc.tolist(), "rad": rad.tolist(), "faceW": fw.tolist()}
        hrr = acc[512] / 1e3 / (nsteps[f] * dt)
        P.append(pieces); N.append(r["n"]); ACC.append(acc); RAD.append(rad); FW.append(fw); HRR.append(hrr)
        print("frame %2d: %d steps, %d pieces, HRR %.1f W/m, Tmax %d K" % (f, nsteps[f], r["alive"], hrr, acc[514]),
              flush=True)
    drv.close()
    o.close()
    nmax = max(N)
    ntgt = 4 * nmax + 64
    np.savez_compressed(os.path.join(HERE, "wood_coupling.npz"),
                        nx=np.int32(nx), ny=np.int32(ny), dx=np.float64(dx), dt=np.float64(dt),
                        params=np.array(json.dumps(app_params())),
                        pieces=np.stack(P)[:, :20 * nmax], n=np.array(N, np.int32),
                        nsteps=np.array(nsteps, np.int32), acc=np.stack(ACC),
                        rad=np.stack(RAD)[:, :4 * ntgt], faceW=np.stack(FW)[:, :4 * nmax])


if __name__ == "__main__":
    main()
