"""Benchmark: Mcell-steps/s of the 1024x1024 reacting gas step on one MI355X.

Workload: 1024x1024 MAC grid, h = 1.25 mm (1.28 m square domain), 300 K air
with Y_O = 0.233, 12 sticks (2.5 cm) and logs (5 cm) in a seeded pile with
pre-lit faces, a 1200 K ignition disc of r = 3 cm; dt = 0.5 h / 2.5 m/s; the
reference schedule (one V-cycle, nu = 2, 120-sweep coarsest, 8 chemistry
substeps).  Synthetic inputs from the seeded scene builder; all fields
resident in HBM when timing starts.

A "step" is one gas step (js/fluid.js:225-253).  W untimed warm-up steps
(after a burn-in of a few frames so the fields are reacting), then K timed
steps replayed from the captured step graph, bracketed by stream syncs.

Besides the throughput line it measures, live on the sim's stream with HIP
events: per-kernel average durations (roofline of the dominant kernel
against the 8 TB/s HBM peak with its algorithmic bytes per launch), and per-
step latency p50/p99.  The CPU baseline is the FP32 oracle (OpenMP) on a
bounded sample of the same workload, timed on this host.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s at 1024² reacting step; batched scenes/s + p99 step latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(tag, nx, ny, levels):
    """Algorithmic HBM bytes of one step's launches of `tag` (FP32; each array
    read once and written once per launch; faces counted as cells)."""
    nC = nx * ny
    per_cell = {"advect_fwd": 56, "advect_corr": 84, "diffuse": 48, "diffuse_sources": 72,
                "sources": 68, "forces": 28, "coef_rhs": 40, "project": 32}
    if tag in per_cell:
        return per_cell[tag] * nC
    fine = [a * b for a, b in levels]
    if tag == "relax":       # p all (4) + rhs half (2) + ax, ay (8) + p half written (2), per half-sweep
        return None
    if tag == "coef_restrict":
        return sum(10 * f for f in fine[:-1])
    if tag == "restrict_res":
        return sum(18 * f for f in fine[:-1])
    if tag == "prolong":
        return sum(17 * f for f in fine[:-1])
    return None


def relax_bytes(levels, nu, vcycles):
    fine = [a * b for a, b in levels]
    # per level: 2 nu sweeps down + 2 nu up, each sweep = 2 half-sweeps of 16 B/cell
    return vcycles * sum(16 * f * 4 * nu for f in fine[:-1])


def cpu_baseline(cfg, budget_s=12.0, max_steps=30):
    from campfire_amd import scenes
    from oracle.pyoracle import Oracle
    o = Oracle(cfg["nx"], cfg["ny"], cfg["dx"], cfg["dt"])
    scenes.apply(o, cfg)
    o.clear_geometry()
    o.geometry()
    o.step(1)
    t0 = time.perf_counter()
    n = 0
    while n < max_steps and time.perf_counter() - t0 < budget_s:
        o.step(1)
        n += 1
    t = time.perf_counter() - t0
    o.close()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    return {"value": round(cfg["nx"] * cfg["ny"] * n / t / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port",
            "sample": "%d steps of the same %dx%d workload in the FP32 C oracle (OpenMP, %d threads), %.1f s"
                      % (n, cfg["nx"], cfg["ny"], threads, t)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--burnin-frames", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs one scene per GPU on a single GPU; --gpus %d is not supported "
              "(scenes shard across GPUs with no communication; see the design notes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim

    cfg = scenes.config_pile_1024(args.seed)
    if args.size != 1024:
        h = cfg["dx"] * 1024 / args.size
        cfg.update(nx=args.size, ny=args.size, dx=h, dt=0.5 * h / 2.5)
        cfg["pieces"], cfg["n"] = scenes.pile(12, args.seed, args.size * h)
    nx, ny = cfg["nx"], cfg["ny"]
    log("[bench] %dx%d h=%.3g mm dt=%.3g ms" % (nx, ny, cfg["dx"] * 1e3, cfg["dt"] * 1e3))
    sim = FluidSim({"nx": nx, "ny": ny, "dt": cfg["dt"]}, dx=cfg["dx"])
    scenes.apply(sim, cfg)
    levels = sim.levels
    # burn-in: app frames (geometry, 8 steps, radiation) so the scene is reacting
    for f in range(args.burnin_frames):
        sim.frame = f
        sim.frame_all(8, True)
        sim.readback_now()
    sim.write_params()
    sim.clear_geometry()
    sim.encode_geometry()
    sim.step(args.warmup)
    sim.synchronize()
    log("[bench] burn-in done, timing %d steps" % args.steps)

    # timed region: K steps from the captured graph
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    t = time.perf_counter() - t0
    ms_step = t * 1e3 / args.steps
    value = nx * ny * args.steps / t / 1e6
    log("[bench] %.4f ms/step  %.1f Mcell-steps/s" % (ms_step, value))

    # per-step latency (one graph replay per step, events around it)
    lat = sim.step_latencies(args.steps) if hasattr(sim, "step_latencies") else None

    # per-kernel durations with HIP events on the sim's stream
    prof = sim.profile_kernels(args.profile_steps)
    sched = {"nu": 2, "vcycles": 1}
    rows = []
    for tag, (ms, launches) in prof.items():
        b = kernel_bytes(tag, nx, ny, levels)
        if tag == "relax":
            b = relax_bytes(levels, sched["nu"], sched["vcycles"])
        ms_per_step = ms / args.profile_steps
        gbs = (b / (ms_per_step * 1e-3) / 1e9) if b else None
        rows.append((tag, ms_per_step, launches // args.profile_steps, b, gbs))
    rows.sort(key=lambda r: -r[1])
    for r in rows:
        log("[bench]   %-16s %8.4f ms/step  %3d launches  %s" % (r[0], r[1], r[2],
            ("%.0f GB/s" % r[4]) if r[4] else ""))
    dom = next(r for r in rows if r[3])
    per_launch_ms = dom[1] / max(dom[2], 1)
    bytes_per_launch = dom[3] / max(dom[2], 1)
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    roof = {"kernel": dom[0], "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(per_launch_ms * 1e3, 2)}
    stage_sum = sum(r[1] for r in rows)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%dx%d reacting gas step, %d-piece seeded pile, h=%.3g mm, dt=%.3g ms, "
                               "reference schedule (1 V-cycle, nu=2, 120-sweep coarsest)"
                               % (nx, ny, cfg["n"], cfg["dx"] * 1e3, cfg["dt"] * 1e3),
                   "grid": [nx, ny], "levels": len(levels), "diff_sub": sim.diff_sub, "seed": args.seed},
        "roofline": roof,
        "kernels_ms_per_step": {r[0]: round(r[1], 5) for r in rows},
        "profiled_ms_per_step": round(stage_sum, 5),
    }
    if lat is not None:
        out["step_latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 5),
                                  "p99": round(float(np.percentile(lat, 99)), 5)}
    sim.close()
    if not args.no_cpu:
        log("[bench] CPU baseline (FP32 oracle)")
        out["cpu_baseline"] = cpu_baseline(cfg)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
