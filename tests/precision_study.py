"""Precision study of one 1024^2 step: each multigrid schedule against an
FP64 oracle step from the same state.

The scene (the seeded 12-piece pile) is warmed up for `warm` steps on the GPU
with the reference schedule; its state is then copied into the FP64 oracle,
which takes one step with V-cycles to ||r||/||rhs|| < 1e-10 and an exact
coarsest solve.  From the same state the GPU takes one step with each
schedule, and the per-field relative L2 error against the FP64 step is
reported for U, V, T, the four Y species, p and the dilatation S.

    python tests/precision_study.py --size 1024 --warm 2000
"""
import argparse
import json
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from campfire_amd import scenes  # noqa: E402

STATE = ["T", "T2", "Th", "Y", "Y2", "Yh", "U", "U2", "Uh", "V", "V2", "Vh", "SQ", "cellG", "fx", "fy", "faceW",
         "p", "acc", "rad"]

# At 1024^2 (7 levels) 120 red-black sweeps leave the 16x16 coarsest level far
# from solved, and the V-cycle then converges at ~0.95 per cycle: that variant
# runs 80 cycles and reports where it gets.  With the direct FP64 coarsest
# solve the cycle converges at ~0.7 and reaches 1e-6 in about 40.
SCHEDULES = {
    "fp32_reference": {"nu": 2, "vcycles": 1, "tol": 0.0, "coarse_iters": 120},
    "fp32_dc_120sweeps": {"nu": 2, "vcycles": 80, "tol": 1e-6, "coarse_iters": 120, "refine": 1},
    "fp32_dc_direct_coarse": {"nu": 2, "vcycles": 80, "tol": 1e-6, "coarse_iters": 0, "refine": 1},
    "fp16_dc_direct_coarse": {"nu": 2, "vcycles": 80, "tol": 1e-6, "coarse_iters": 0, "refine": 1,
                              "smoother_fp16": 1},
    # the V-cycle as the preconditioner of flexible CG: the same target with
    # the cheap coarsest sweeps
    "fp32_fcg_120sweeps": {"nu": 2, "vcycles": 80, "tol": 1e-6, "coarse_iters": 120, "refine": 1, "krylov": 1},
    "fp16_fcg_24sweeps": {"nu": 3, "vcycles": 80, "tol": 1e-6, "coarse_iters": 24, "refine": 1, "smoother_fp16": 1,
                          "krylov": 1},
}


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def split(fields):
    """The compared quantities from a field dict (Y as 4 species, S from SQ)."""
    Y = fields["Y"].reshape(-1, 4)
    return {"U": fields["U"], "V": fields["V"], "T": fields["T"], "Y_F": Y[:, 0], "Y_O": Y[:, 1], "Y_P": Y[:, 2],
            "Y_S": Y[:, 3], "p": fields["p"], "S": fields["SQ"].reshape(-1, 2)[:, 0]}


def study(size=1024, warm=2000, seed=1, log=None):
    from oracle.pyoracle import Oracle
    from campfire_amd.fluid import FluidSim
    log = log or (lambda *a: None)
    cfg = scenes.config_pile_1024(seed)
    if size != 1024:
        h = cfg["dx"] * 1024 / size
        cfg.update(nx=size, ny=size, dx=h, dt=0.5 * h / 2.5)
        cfg["pieces"], cfg["n"] = scenes.pile(12, seed, size * h)
    g = FluidSim({"nx": cfg["nx"], "ny": cfg["ny"], "dt": cfg["dt"]}, dx=cfg["dx"])
    scenes.apply(g, cfg)
    t0 = time.perf_counter()
    for f in range(warm // 8):
        g.frame = f
        g.frame_all(8, True)
        g.readback_now()
    g.write_params()
    g.synchronize()
    log("[precision] %dx%d warmed %d steps in %.1f s" % (cfg["nx"], cfg["ny"], warm, time.perf_counter() - t0))
    snap = {k: g.read_field(k) for k in STATE}
    out = {"grid": [cfg["nx"], cfg["ny"]], "warm_steps": warm, "errors": {}, "vcycles": {}, "resid": {},
           "compat": {}, "ms_per_step": {}}
    # the FP64 step
    o = Oracle(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dt"], fp64=True)
    scenes.apply(o, cfg)
    for k, v in snap.items():
        if k in ("cellG", "faceW", "acc", "rad"):   # ids and integers: as stored
            o.write(k, v)
        else:
            o.write64(k, v.astype(np.float64))
    o.write_params()
    o.set_schedule(nu=2, vcycles=200, tol=1e-10, coarse_iters=0)
    t0 = time.perf_counter()
    o.step(1)
    ref = {k: o.read64(k) for k in ("T", "Y", "U", "V", "p", "SQ")}
    log("[precision] FP64 oracle step: %d V-cycles, residual %.2e, %.1f s"
        % (o.last_vcycles, o.last_resid, time.perf_counter() - t0))
    R = split(ref)
    # the FP32 oracle with the reference schedule: FP32 arithmetic alone
    o32 = Oracle(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dt"])
    scenes.apply(o32, cfg)
    for k, v in snap.items():
        o32.write(k, v)
    o32.write_params()
    o32.step(1)
    F = split({k: o32.read(k) for k in ("T", "Y", "U", "V", "p", "SQ")})
    out["errors"]["oracle_fp32_reference"] = {k: rel(F[k], R[k]) for k in R}
    # the GPU schedules
    def restore(sch):
        # the fields, and (through a schedule without refine) p64 from p
        for k, v in snap.items():
            g.write_field(k, v)
        g.set_schedule()
        g.set_schedule(**sch)
        g.write_params()

    for name, sch in SCHEDULES.items():
        restore(sch)
        g.step(1)
        info = g.info()
        G = split({k: g.read_field(k) for k in ("T", "Y", "U", "V", "p", "SQ")})
        out["errors"][name] = {k: rel(G[k], R[k]) for k in R}
        if name == "fp32_reference":
            # the GPU's reference step against the FP32 oracle's, directly
            out["gpu_vs_oracle_fp32"] = {k: rel(G[k], F[k]) for k in F}
            log("[precision] GPU vs FP32 oracle, reference step: %s"
                % " ".join("%s %.1e" % (k, v) for k, v in out["gpu_vs_oracle_fp32"].items()))
        out["vcycles"][name] = int(info.last_vcycles)
        out["resid"][name] = float(info.last_resid) if sch.get("tol", 0) > 0 else None
        src, outf = g.compat_sums()
        out["compat"][name] = {"source": src, "outflow": outf, "rel_gap": abs(src - outf) / max(abs(src), 1e-300)}
        # device time of the same step from the same state (median of 5 tries)
        lat = []
        for _ in range(5):
            restore(sch)
            lat.append(float(g.step_latencies(1)[0]))
        out["ms_per_step"][name] = float(np.median(lat))
        log("[precision] %-28s %s  %d cycles  %.3f ms" % (
            name, " ".join("%s %.1e" % (k, v) for k, v in out["errors"][name].items()), out["vcycles"][name],
            out["ms_per_step"][name]))
    g.set_schedule()
    g.close()
    o.close()
    o32.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--warm", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args()
    res = study(a.size, a.warm, a.seed, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
