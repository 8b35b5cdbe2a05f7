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
step latency p50/p99.  Before any of that (the process must not have touched
the GPU when it starts a profiler), two rocprofv3 --pmc runs of a short
no-graph copy of the workload give the HBM bytes each kernel fetches and
writes, calibrated on a 256 MiB device copy.  The CPU baseline is the FP32
oracle (OpenMP) on a bounded sample of the same workload, timed on this host.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size N] [--no-cpu] [--no-traffic]
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
PLAIN_CAP = 400  # V-cycle cap of the plain per-scene FP32 form of the equal-accuracy comparison
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(tag, nx, ny, levels, launches):
    """Algorithmic HBM bytes of one step's `launches` launches of `tag`: each
    array the kernel needs read once and each output written once (FP32;
    faces counted as cells; float4 Y = 16 B, float2 SQ = 8 B per cell).
    Halo re-reads, LDS staging and partial sums are not counted."""
    nC = nx * ny
    per_cell = {
        "advect": 56,            # T, Y, U, V read; T2, Y2, U2, V2 written
        "advect_fwd": 56,        # T, Y, U, V read; Th, Yh, Uh, Vh written
        "advect_corr": 84,       # T, Y, U, V, Th, Yh, Uh, Vh read; T2, Y2, U2, V2 written
        "diffuse": 48,           # T, Y, fx, fy read; T, Y written
        "diffuse_sources": 72,   # diffusion + sources: T, Y, fx, fy, cellG read; T, Y, SQ written
        "sources": 68,
        "forces": 28,
        "coef_rhs": 40,
        "project": 32,           # U, V, p (x2 neighbours), fx, fy read; U, V written
    }
    if tag in per_cell:
        return per_cell[tag] * nC
    fine = [a * b for a, b in levels]
    base, _, lev = tag.partition(".")
    lev = int(lev) if lev else None
    if tag == "forces_coef":
        # T, U2, V2, fx, fy (4 B), SQ (8 B) read; U, V, ax0, ay0, rhs0 written;
        # ax, ay of min(L-1, 5) coarse levels written
        casc = min(len(levels) - 1, 5)
        return 48 * nC + sum(8 * f for f in fine[1:1 + casc])
    if base == "mg_down" and lev is not None:
        # p (level 0 only: a level above starts from zero unread), rhs, ax, ay
        # read; p, coarse rhs, coarse p written
        return (22 if lev == 0 else 18) * fine[lev]
    if base == "mg_up" and lev is not None:     # p, coarse p, rhs, ax, ay read; p written
        return 21 * fine[lev]
    if base == "mg_up_project" and lev is not None:
        # the level-0 up leg (21 B/cell) and the projection without its p
        # reads: U, V, fx, fy, T read, U, V written
        return 21 * fine[0] + 28 * nC
    if base == "mg_down2" and lev is not None:
        # levels lev, lev+1 in one launch: fine rhs, ax, ay read, q written
        # (the fine p starts from zero unread); coarse ax, ay read, q, rhs
        # written; next level's rhs, p written
        return 16 * fine[lev] + 16 * fine[lev + 1] + 8 * fine[lev + 2]
    if base == "mg_up2" and lev is not None:
        # fine and coarse q, rhs, ax, ay read, p written; next level's p read
        return 20 * fine[lev] + 20 * fine[lev + 1] + 4 * fine[lev + 2]
    if base == "coef_restrict" and lev is not None:   # fine ax, ay read; coarse ax, ay written
        return 8 * fine[lev - 1] + 8 * fine[lev]
    if tag == "restrict_res":
        return sum(18 * f for f in fine[:-1])
    if tag == "prolong":
        return sum(17 * f for f in fine[:-1])
    return None     # mg_bottom: one workgroup on <= 5.4k cells, latency-bound; relax: see relax_bytes


def relax_bytes(levels, nu, vcycles):
    fine = [a * b for a, b in levels]
    # per level: 2 nu sweeps down + 2 nu up, each sweep = 2 half-sweeps of 16 B/cell
    return vcycles * sum(16 * f * 4 * nu for f in fine[:-1])


# kernel symbol (base name, template argument) behind each profile tag
KERNEL_SYMBOL = {"diffuse_sources": ("k_sources", "true"), "sources": ("k_sources", "false")}
COPY_BYTES = 256 << 20   # the calibration copy: 256 MiB read, 256 MiB written


def _symbol(kname):
    import re
    m = re.search(r"cf::(k_\w+)(?:<([^>]*)>)?", kname)
    if not m:
        return None
    return m.group(1), (m.group(2) or "").split(",")[0].strip()


def pmc_child(args):
    """Body of a counter run: the library's streaming copy of 256 MiB (the
    calibration kernel, k_bw_copy), then a few steps of the workload
    dispatched one kernel at a time."""
    os.environ["CF_NO_GRAPH"] = "1"
    import ctypes as C
    from campfire_amd import _lib
    g = C.c_double(0)
    _lib.check(_lib.lib().cf_copy_bandwidth(COPY_BYTES, 3, C.byref(g), None), "copy")
    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim
    cfg = bench_config(args)
    sim = FluidSim({"nx": cfg["nx"], "ny": cfg["ny"], "dt": cfg["dt"]}, dx=cfg["dx"])
    scenes.apply(sim, cfg)
    sim.frame_all(8, True)
    sim.readback_now()
    sim.step(4)
    sim.synchronize()
    sim.close()
    log("[pmc-child] done")


def measure_traffic(args):
    """HBM bytes per launch of every gas-step kernel from two rocprofv3 --pmc
    runs (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by the
    ratio of the known bytes of a plain 256 MiB copy to what the same counter
    reads for it.  Runs before this process touches the GPU."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="cf_pmc_", dir="/tmp")
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "pmc", "--",
               sys.executable, os.path.abspath(__file__), "--pmc-child", "--size", str(args.size),
               "--seed", str(args.seed)]
        log("[bench] counter run %s" % counter)
        env = dict(os.environ, TMPDIR="/tmp")
        p = subprocess.Popen(cmd, cwd="/tmp", env=env, start_new_session=True,
                             stdout=subprocess.DEVNULL, stderr=sys.stderr)
        try:
            rc = p.wait(timeout=240)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return None, "%s counter run timed out" % counter
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None, "%s counter run wrote no csv (rc %d)" % (counter, rc)
        rows = []
        for f in files:
            with open(f, newline="") as fh:
                rows += [r for r in csv.DictReader(fh) if r.get("Counter_Name") == counter]
        shutil.rmtree(d, ignore_errors=True)
        per = {}
        copy = []
        for r in rows:
            v = float(r["Counter_Value"])
            sym = _symbol(r["Kernel_Name"])
            if sym is not None and sym[0] == "k_bw_copy":
                copy.append((v, r["Kernel_Name"]))
            elif sym is not None:
                per.setdefault(sym + (int(r.get("Grid_Size") or 0),), []).append(v)
        out[counter] = (per, copy)
    # calibration: the library's 256 MiB copy (k_bw_copy), median over its launches
    fetch_copy = [v for v, n in out["FETCH_SIZE"][1]]
    write_copy = [v for v, n in out["WRITE_SIZE"][1]]
    name = out["FETCH_SIZE"][1][0][1] if out["FETCH_SIZE"][1] else ""
    if not fetch_copy or not write_copy or min(fetch_copy) <= 0 or min(write_copy) <= 0:
        return None, "copy calibration kernel not found"
    f_scale = COPY_BYTES / (float(np.median(fetch_copy)) * 1024.0)
    w_scale = COPY_BYTES / (float(np.median(write_copy)) * 1024.0)
    per = {}
    for sym in set(out["FETCH_SIZE"][0]) | set(out["WRITE_SIZE"][0]):
        fv = out["FETCH_SIZE"][0].get(sym)
        wv = out["WRITE_SIZE"][0].get(sym)
        if fv and wv:
            per[sym] = {"fetch": float(np.mean(fv)) * 1024.0 * f_scale, "write": float(np.mean(wv)) * 1024.0 * w_scale,
                        "dispatches": len(fv)}
    cal = {"copy_kernel": name[:80], "copy_bytes": COPY_BYTES, "fetch_factor": round(f_scale, 4),
           "write_factor": round(w_scale, 4)}
    for k in sorted(per, key=lambda k: -(per[k]["fetch"] + per[k]["write"])):
        log("[bench]   traffic %-16s grid %9d  fetch %10.0f  write %10.0f  B/launch" %
            (k[0] + ("<%s>" % k[1] if k[1] else ""), k[2], per[k]["fetch"], per[k]["write"]))
    return (per, cal), None


def step_bytes(n_sub=1, nu=2, vcycles=1):
    """Algorithmic HBM bytes of one gas step per cell (FP32, each array read
    and written once per global-synchronisation phase, faces as cells):
    advect forward 60 + advect correct 84 + sources 68 + forces 28 + coef and
    rhs0 40 + coefficient restriction over the levels 13.3 + project 32, then
    48 per diffusion substep and (40 nu + 35) * 4/3 per V-cycle (4/3: the
    levels' geometric sum).  526.7 B at the reference schedule."""
    return 325.3 + 48.0 * n_sub + vcycles * (40.0 * nu + 35.0) * 4.0 / 3.0


def batched_1024(args, n):
    """n scenes of the 1024^2 workload (the same pile and disc, seeds
    args.seed .. args.seed + n - 1) stepped together by one ScenePool at the
    reference schedule: Mcell-steps/s, the pool step's latency, and the
    step's rate in algorithmic bytes (step_bytes) against the 8 TB/s HBM
    peak.  The n scenes' state (about 150 MB each) is larger than the
    256 MB Infinity Cache."""
    from campfire_amd import scenes
    from campfire_amd.pool import ScenePool
    pool = ScenePool(200 << 30)
    views = []
    for q in range(n):
        a2 = argparse.Namespace(**vars(args))
        a2.seed = args.seed + q
        c = bench_config(a2)
        slot, sim = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
        scenes.apply(sim, c)
        views.append(sim)
    for f in range(2):
        pool.frame(8, True)
    tiles = pool_tiles(args, pool, views)
    pool.step(args.warmup)            # after the tuner, so the timed steps start on a busy device
    pool.synchronize()
    info = pool.info()
    cells = int(info.cells)
    K = args.batch_steps
    t0 = time.perf_counter()
    pool.step(K)
    pool.synchronize()
    tb = time.perf_counter() - t0
    lat = pool.step_latencies(K)
    pool.close()
    rate = cells * K / tb
    gbs = rate * step_bytes() / 1e9
    res = {"scenes": n, "grid": [1024, 1024], "cells": cells, "hbm_bytes": int(info.used), "steps": K,
           "mcell_steps_per_s": round(rate / 1e6, 1), "ms_per_step": round(tb * 1e3 / K, 4),
           "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                               "p99": round(float(np.percentile(lat, 99)), 4)},
           "launches_per_step": int(info.launches_per_step), "leg_tiles": tiles,
           "step_bytes_per_cell": round(step_bytes(), 1), "achieved_gbs": round(gbs, 1),
           "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    log("[bench] batch of %d x 1024^2: %.1f Mcell-steps/s, %.3f ms per pool step (p99 %.3f), %.0f GB/s at %.1f B per "
        "cell-step = %.3f of the HBM peak" % (n, res["mcell_steps_per_s"], res["ms_per_step"],
                                             res["step_latency_ms"]["p99"], gbs, step_bytes(), gbs / HBM_PEAK_GBS))
    return res


def pool_tiles(args, pool, views):
    """The pool's leg tile per depth: the autotuner's pick from a snapshot of
    the scenes as they are (left as they were), or the size rule with
    --no-tune."""
    if not args.no_tune:
        from campfire_amd.autotune import tune_pool_leg_tiles
        tune_pool_leg_tiles(pool, views, log=log)
        pool.synchronize()
    return pool.leg_tiles()


def batched(args):
    """The mixed-size batch: args.batch scenes (sizes drawn from 256^2, 512^2,
    1024^2) stepped together by one ScenePool, against the same scenes stepped
    one after another with their own step graphs (the per-scene form).  Both
    give bitwise the same fields."""
    from campfire_amd import scenes
    from campfire_amd.pool import ScenePool
    cfgs = scenes.config_batch(args.seed, args.batch)
    pool = ScenePool(200 << 30)
    sims = []
    for c in cfgs:
        slot, sim = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
        scenes.apply(sim, c)
        sims.append(sim)
    pool.frame(8, True)          # geometry and a first burning frame
    tiles = pool_tiles(args, pool, sims)
    pool.step(args.warmup)            # after the tuner, so the timed steps start on a busy device
    pool.synchronize()
    info = pool.info()
    cells = int(info.cells)
    K = args.batch_steps
    log("[bench] batch: %d scenes, %.1f M cells, %.2f GB, %d launches per step"
        % (len(cfgs), cells / 1e6, info.used / 1e9, info.launches_per_step))
    t0 = time.perf_counter()
    pool.step(K)
    pool.synchronize()
    tb = time.perf_counter() - t0
    lat = pool.step_latencies(K)
    # per-scene form: each scene alone, one after another
    for sim in sims:
        sim.step(2)              # capture its own graph
    pool.synchronize()
    t0 = time.perf_counter()
    for sim in sims:
        sim.step(K)
    pool.synchronize()
    ts = time.perf_counter() - t0
    pool.close()
    n = len(cfgs)
    res = {"scenes": n, "sizes": {str(s): sum(1 for c in cfgs if c["nx"] == s) for s in (256, 512, 1024)},
           "cells": cells, "hbm_bytes": int(info.used), "steps": K,
           "batched_mcell_steps_per_s": round(cells * K / tb / 1e6, 1),
           "batched_scene_steps_per_s": round(n * K / tb, 1),
           "batched_ms_per_step": round(tb * 1e3 / K, 4),
           "per_scene_mcell_steps_per_s": round(cells * K / ts / 1e6, 1),
           "speedup_vs_per_scene": round(ts / tb, 3),
           "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                               "p99": round(float(np.percentile(lat, 99)), 4)},
           "launches_per_step": int(info.launches_per_step), "leg_tiles": tiles}
    log("[bench] batch: %.1f Mcell-steps/s batched, %.1f per scene (x%.2f); %.1f scene-steps/s; p99 %.3f ms"
        % (res["batched_mcell_steps_per_s"], res["per_scene_mcell_steps_per_s"], res["speedup_vs_per_scene"],
           res["batched_scene_steps_per_s"], res["step_latency_ms"]["p99"]))
    return res


def converged_batch(args):
    """The same mixed batch at equal accuracy: every scene's V-cycles run to
    ||r||/||rhs|| < tol (capped at `cap` cycles per step) in
      (a) the plain per-scene FP32 form: one FluidSim per scene, stepped one
          scene after another, FP32 V-cycles with the reference smoother
          settings (nu 2, 120 coarsest sweeps), capped at PLAIN_CAP cycles,
          which it never reaches;
      (b) the batched mixed-precision form: one ScenePool stepping every
          scene per launch, FP16 smoother with FP64 defect correction, nu 3
          and 24 coarsest sweeps;
      (c) as (b) with the V-cycle preconditioning flexible conjugate
          gradients at level 0 (FP64 pressure, residual and direction), at
          the autotuner's pick for these scenes.
    Steps are timed one at a time; each scene's V-cycles and achieved
    residual after every step are recorded for every form, so the accuracies
    are compared, not assumed."""
    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim
    from campfire_amd.pool import ScenePool
    tol, cap, K = 1e-4, 40, args.conv_steps
    cfgs = scenes.config_batch(args.seed, args.batch)
    cells = sum(c["nx"] * c["ny"] for c in cfgs)

    def batched(sched):
        pool = ScenePool(200 << 30)
        pool.set_schedule(vcycles=cap, tol=tol, refine=1, smoother_fp16=1, **sched)
        views = []
        for c in cfgs:
            slot, sim = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
            scenes.apply(sim, c)
            views.append(sim)
        pool.frame(8, True)
        pool.step(2)
        pool.synchronize()
        tiles = pool_tiles(args, pool, views)
        pool.step(2)
        pool.synchronize()
        t, rb, cb = 0.0, [], []
        for q in range(K):
            t0 = time.perf_counter()
            pool.step(1)
            pool.synchronize()
            t += time.perf_counter() - t0
            rb += [float(v.info().last_resid) for v in views]
            cb += [int(v.info().last_vcycles) for v in views]
        lat = pool.step_latencies(K)
        pool.close()
        return {"mcell_steps_per_s": round(cells * K / t / 1e6, 1),
                "schedule": "nu %d, %s coarsest, FP16 smoother, FP64 %s"
                            % (sched["nu"], "%d-sweep" % sched["coarse_iters"] if sched["coarse_iters"] else "direct",
                               "flexible CG" if sched.get("krylov") else "defect correction"),
                "resid_max": max(rb), "cycles_mean": round(float(np.mean(cb)), 2),
                "steps_at_cap": int(sum(1 for c in cb if c >= cap)), "leg_tiles": tiles,
                "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                                    "p99": round(float(np.percentile(lat, 99)), 4)}}, t

    dcb, tdc = batched({"nu": 3, "coarse_iters": 24})
    kry, tkr = batched({"nu": args.fcg_nu, "coarse_iters": args.fcg_coarse, "krylov": 1})

    def per_scene(sched, label, cap=cap):
        # one FluidSim per scene, stepped one scene after another, FP32 smoother
        sims = []
        for c in cfgs:
            sim = FluidSim({"nx": c["nx"], "ny": c["ny"], "dt": c["dt"]}, dx=c["dx"])
            scenes.apply(sim, c)
            sim.set_schedule(vcycles=cap, tol=tol, **sched)
            sim.frame_all(8, True)
            sim.step(2)
            sims.append(sim)
        for sim in sims:
            sim.synchronize()
        ts, ra, ca = 0.0, [], []
        for q in range(K):
            for sim in sims:
                t0 = time.perf_counter()
                sim.step(1)
                sim.synchronize()
                ts += time.perf_counter() - t0
                ra.append(float(sim.info().last_resid))
                ca.append(int(sim.info().last_vcycles))
        for sim in sims:
            sim.close()
        return {"mcell_steps_per_s": round(cells * K / ts / 1e6, 1), "schedule": label, "cycle_cap": cap,
                "resid_max": max(ra), "cycles_mean": round(float(np.mean(ca)), 2),
                "cycles_max": int(max(ca)), "steps_at_cap": int(sum(1 for c in ca if c >= cap))}, ts

    # (a) the same solver as the batched FCG form, per scene and in FP32
    same, tsame = per_scene({"nu": args.fcg_nu, "coarse_iters": args.fcg_coarse, "refine": 1, "krylov": 1},
                            "nu %d, %d-sweep coarsest, FP32 smoother, FP64 flexible CG" % (args.fcg_nu, args.fcg_coarse))
    # (b) the reference's smoother settings with plain FP32 V-cycles; a few
    # scene-steps need more than 40 of them, so this form gets a cap it never
    # reaches and meets the tolerance in every scene-step, as the others do
    plain, tplain = per_scene({"nu": 2, "coarse_iters": 120}, "nu 2, 120 coarsest sweeps, FP32 V-cycles",
                              cap=PLAIN_CAP)
    res = {"tol": tol, "cycle_cap": cap, "steps": K, "scenes": len(cfgs), "cells": cells,
           "batched_mixed_fcg": kry, "batched_mixed": dcb,
           "per_scene_fp32_fcg": same, "per_scene_fp32": plain,
           # batching + the FP16 smoother alone: the same solver on both sides
           "speedup": round(tsame / tkr, 3),
           # against the plain per-scene FP32 V-cycles, the reference's solver, at the same tolerance
           "speedup_vs_plain_vcycles": round(tplain / tkr, 3),
           "speedup_defect_correction_vs_plain": round(tplain / tdc, 3)}
    for name, r in (("batched mixed + FCG", kry), ("batched mixed", dcb), ("per-scene FP32 + FCG", same),
                    ("per-scene FP32 V-cycles", plain)):
        log("[bench] equal accuracy (tol %.0e): %s %.1f Mcell-steps/s (resid max %.2e, %.2f cycles, %d scene-steps "
            "at the cap)" % (tol, name, r["mcell_steps_per_s"], r["resid_max"], r["cycles_mean"], r["steps_at_cap"]))
    log("[bench] equal accuracy: x%.2f against the same solver per scene, x%.2f against plain V-cycles; "
        "batched FCG p99 %.2f ms" % (res["speedup"], res["speedup_vs_plain_vcycles"], kry["step_latency_ms"]["p99"]))
    return res


def plain_256(args, steps=300):
    """The 256^2 plain-method workload (CFL dt on the device, V-cycles until
    ||r||/||rhs|| < 1e-6 with the FP64 pressure) over `steps` steps after 20
    warm-up steps: its own schedule (defect correction, nu 2, 120 coarsest
    sweeps) against the V-cycle as FCG preconditioner (same smoother) and FCG
    with the FP16 smoother (nu 3, 24 coarsest sweeps).  Every step's cycle
    count and residual are recorded."""
    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim
    variants = {"defect_correction": {}, "fcg": {"krylov": 1},
                "fcg_fp16": {"krylov": 1, "smoother_fp16": 1, "nu": 3, "coarse_iters": 24}}
    out = {}
    for name, extra in variants.items():
        cfg = scenes.config_plain_256(args.seed)
        cfg["schedule"] = dict(cfg["schedule"], **extra)
        g = FluidSim({"nx": cfg["nx"], "ny": cfg["ny"], "dt": cfg["dt"]}, dx=cfg["dx"])
        scenes.apply(g, cfg)
        g.clear_geometry()
        g.encode_geometry()
        g.step(20)
        g.synchronize()
        cyc, res = [], []
        t = 0.0
        for q in range(steps):
            t0 = time.perf_counter()
            g.step(1)
            g.synchronize()
            t += time.perf_counter() - t0
            i = g.info()
            cyc.append(int(i.last_vcycles))
            res.append(float(i.last_resid))
        g.close()
        out[name] = {"mcell_steps_per_s": round(256 * 256 * steps / t / 1e6, 2),
                     "ms_per_step": round(t * 1e3 / steps, 4),
                     "cycles_mean": round(float(np.mean(cyc)), 2), "cycles_max": int(max(cyc)),
                     "resid_max": max(res), "steps": steps}
        log("[bench] 256^2 plain run, %-18s %.3f ms/step (%.1f Mcell-steps/s), %.1f cycles (max %d), resid max %.2e"
            % (name, out[name]["ms_per_step"], out[name]["mcell_steps_per_s"], out[name]["cycles_mean"],
               out[name]["cycles_max"], out[name]["resid_max"]))
    return out


def _steps_per_frame(frames, dt, fps=60.0):
    """Gas steps per app frame at a steady frame rate (js/main.js:252-257)."""
    import math
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


def reference_modes(args, frames=60):
    """The reference's three modes (js/config.js:6-10; dx = 0.7 m / nx) on the
    reference's own default scene (js/main.js:132-155, packed by
    tests/golden/make_wood_coupling.py), stepped as the app does: per frame
    the parameters, geometry, the frame's gas steps at 60 frames/s of
    simulated time, radiation and a readback that the host waits for.  Against
    the real-time rate each mode implies (steps of dt per second)."""
    import json as _json
    from campfire_amd.fluid import FluidSim
    z = np.load(os.path.join(ROOT, "tests", "golden", "default_scene.npz"), allow_pickle=False)
    pieces, n = z["pieces"], int(z["n"])
    params = _json.loads(str(np.load(os.path.join(ROOT, "tests", "golden", "wood_coupling.npz"),
                                     allow_pickle=False)["params"]))
    out = {}
    for name, (nx, ny, dt) in (("coarse", (160, 224, 0.004)), ("balanced", (224, 320, 0.003)),
                               ("fine", (352, 512, 0.002))):
        sim = FluidSim({"nx": nx, "ny": ny, "dt": dt}, dx=0.7 / nx)
        for k, v in params.items():
            sim.set_param(k, v)
        sim.reset()
        sim.upload_pieces(pieces, n)
        per = _steps_per_frame(frames + 10, dt)
        for f in range(10):
            sim.frame = f
            sim.frame_all(per[f], True)
            sim.readback_now()
        sim.synchronize()
        t0 = time.perf_counter()
        steps = 0
        for f in range(10, 10 + frames):
            sim.frame = f
            sim.frame_all(per[f], True)
            sim.readback_now()
            steps += per[f]
        t = time.perf_counter() - t0
        sim.close()
        rt = nx * ny / dt / 1e6
        out[name] = {"grid": [nx, ny], "dt_ms": dt * 1e3, "frames": frames, "steps": steps,
                     "ms_per_frame": round(t * 1e3 / frames, 4),
                     "mcell_steps_per_s": round(nx * ny * steps / t / 1e6, 1),
                     "real_time_mcell_steps_per_s": round(rt, 2),
                     "x_real_time": round(nx * ny * steps / t / 1e6 / rt, 1)}
        log("[bench] %-8s %dx%d: %.3f ms per app frame (%.1f steps), %.1f Mcell-steps/s = %.1fx real time"
            % (name, nx, ny, out[name]["ms_per_frame"], steps / frames, out[name]["mcell_steps_per_s"],
               out[name]["x_real_time"]))
    return out


def serving(args):
    """The serving workload: Poisson arrivals in wall-clock time of scene
    requests at 256^2, 512^2 and 1024^2, each for 500-5000 steps, admitted
    into one pool by HBM budget and stepped together (continuous batching).
    Reports p50/p99 of the batched step's device time, aggregate cell-steps/s
    over the wall time, queueing delay and rebuild times."""
    from campfire_amd.serve import make_requests, serve
    reqs = make_requests(args.serve_requests, args.serve_rate, args.seed, steps=(500, 5000))
    budget = int(args.serve_budget_gb * (1 << 30))
    stats, _ = serve(reqs, budget)
    stats["config"] = {"requests": args.serve_requests, "rate_per_s": args.serve_rate,
                       "sizes": [256, 512, 1024], "steps_per_request": [500, 5000],
                       "budget_gib": args.serve_budget_gb, "seed": args.seed, "clock": "wall"}
    log("[bench] serving: %d requests over %.1f s of arrivals, %.1f s wall: %.1f Mcell-steps/s, step p50/p99 "
        "%.3f/%.3f ms, queue p50/p99 %.2f/%.2f s, %d rebuilds p50 %.2f ms, up to %d live"
        % (stats["requests"], stats["arrival_span_s"], stats["wall_s"], stats["mcell_steps_per_s"],
           stats["step_latency_ms"]["p50"], stats["step_latency_ms"]["p99"], stats["queue_s"]["p50"],
           stats["queue_s"]["p99"], stats["rebuild_ms"]["count"], stats["rebuild_ms"]["p50"] or 0, stats["max_live"]))
    return stats


def flame1d_leg():
    """The 1D flame table of tests/flame1d.mjs (B = 2e10, seven dx/F rows) in
    one GPU launch (FP64), the C port of the script on one host core, and the
    reference script's own wall time under Node (recorded with its output in
    tests/golden/flame1d.json)."""
    import json as _json
    from campfire_amd.flame1d import flame_speeds
    from oracle.pyoracle import flame_speed as port_speed
    cases = []
    for dx in [5e-5, 2e-4, 1e-3, 3.125e-3]:
        for F in ([1, max(1, 3 * dx / 5e-4)] if dx > 5e-5 else [1]):
            cases.append(dict(B=2e10, Ta=15000.0, dx=dx, F=F, L=max(0.03, 60 * dx), tEnd=3.0))
    flame_speeds(cases[-1:], fp64=True)
    t0 = time.perf_counter()
    gpu = flame_speeds(cases, fp64=True)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    port = [port_speed(c["B"], c["Ta"], c["dx"], c["F"], 1.0, c["L"], c["tEnd"]) for c in cases]
    tp = time.perf_counter() - t0
    gold = _json.load(open(os.path.join(ROOT, "tests", "golden", "flame1d.json")))["runs"][0]
    res = {"cases": len(cases), "gpu_s": round(tg, 4), "port_s": round(tp, 4), "port_cores": 1,
           "script_s": round(gold["wall_s"], 3), "script_where": gold.get("source", "") + ", builder's container",
           "SL_readme_case": round(gpu[0]["SL"], 5), "SL_table_gpu": [round(r["SL"], 4) for r in gpu],
           "SL_table_port": [round(r["SL"], 4) for r in port], "SL_table_script": [r["SL"] for r in gold["rows"]]}
    log("[bench] flame1d: 7 cases GPU %.3f s, C port %.3f s (1 core), script %.2f s (Node); S_L %.5f m/s"
        % (tg, tp, gold["wall_s"], gpu[0]["SL"]))
    return res


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_config(args):
    from campfire_amd import scenes
    cfg = scenes.config_pile_1024(args.seed)
    if args.size != 1024:
        h = cfg["dx"] * 1024 / args.size
        cfg.update(nx=args.size, ny=args.size, dx=h, dt=0.5 * h / 2.5)
        cfg["pieces"], cfg["n"] = scenes.pile(12, args.seed, args.size * h)
    return cfg


def cpu_baseline(cfg, budget_s=12.0, max_steps=200):
    """The FP32 oracle (OpenMP) on a bounded sample of the workload `cfg`."""
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
            "kind": "port", "cpu": _cpu_model(),
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
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc runs")
    ap.add_argument("--batch", type=int, default=64, help="scenes of the batched measurement (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=40)
    ap.add_argument("--batch1024", type=int, default=16, help="1024^2 scenes of the one-size batch (0: skip)")
    ap.add_argument("--fcg-nu", type=int, default=3, help="smoother sweeps of the batched FCG form")
    ap.add_argument("--fcg-coarse", type=int, default=24, help="coarsest sweeps of the batched FCG form (0: direct)")
    ap.add_argument("--conv-steps", type=int, default=6,
                    help="steps of the equal-accuracy batch comparison (0: skip)")
    ap.add_argument("--plain-steps", type=int, default=300,
                    help="timed steps of the 256^2 plain-method workload (0: skip)")
    ap.add_argument("--converged-only", action="store_true",
                    help="run only the equal-accuracy batch comparison and print its JSON (for profiling)")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests of the serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0, help="request arrivals per second")
    ap.add_argument("--serve-budget-gb", type=float, default=2.0, help="HBM budget of the serving pool")
    ap.add_argument("--no-modes", action="store_true", help="skip the reference-mode frames")
    ap.add_argument("--no-tune", action="store_true", help="keep the default leg tiles (no autotuning)")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.converged_only:
        print(json.dumps(converged_batch(args)), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs one scene per GPU on a single GPU; --gpus %d is not supported "
              "(scenes shard across GPUs with no communication; see the design notes)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    traffic, traffic_err = (None, "disabled") if args.no_traffic else measure_traffic(args)
    if traffic_err:
        log("[bench] traffic: %s" % traffic_err)

    from campfire_amd import scenes
    from campfire_amd.fluid import FluidSim

    cfg = bench_config(args)
    nx, ny = cfg["nx"], cfg["ny"]
    log("[bench] %dx%d h=%.3g mm dt=%.3g ms" % (nx, ny, cfg["dx"] * 1e3, cfg["dt"] * 1e3))
    # the leg tiles of this grid, by the autotuner (results do not depend on them)
    tiles = None
    if not args.no_tune:
        from campfire_amd.autotune import tune_leg_tiles
        tiles, _ = tune_leg_tiles(nx, ny, cfg["dx"], cfg["dt"], log=log)
    sim = FluidSim({"nx": nx, "ny": ny, "dt": cfg["dt"]}, dx=cfg["dx"])
    scenes.apply(sim, cfg)
    if tiles:
        sim.set_leg_tiles(tiles)
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
        b = kernel_bytes(tag, nx, ny, levels, launches // args.profile_steps)
        if tag == "relax":
            b = relax_bytes(levels, sched["nu"], sched["vcycles"])
        ms_per_step = ms / args.profile_steps
        gbs = (b / (ms_per_step * 1e-3) / 1e9) if b else None
        rows.append((tag, ms_per_step, launches // args.profile_steps, b, gbs))
    rows.sort(key=lambda r: -r[1])
    for r in rows:
        log("[bench]   %-16s %8.4f ms/step  %3d launches  %s" % (r[0], r[1], r[2],
            ("%.0f GB/s" % r[4]) if r[4] else ""))
    # the roofline kernel: the HBM-streaming launch site with the most time per step
    dom = next(r for r in rows if r[3])
    per_launch_ms = dom[1] / max(dom[2], 1)
    bytes_per_launch = dom[3] / max(dom[2], 1)
    achieved = bytes_per_launch / (per_launch_ms * 1e-3) / 1e9
    tr = None
    if traffic:
        base, _, lev = dom[0].partition(".")
        sym = KERNEL_SYMBOL.get(base, ("k_" + base, None))
        hits = {k: v for k, v in traffic[0].items() if k[0] == sym[0] and (sym[1] is None or k[1] == sym[1])}
        grids = sorted({k[2] for k in hits}, reverse=True)
        if grids:
            g = grids[int(lev)] if lev and int(lev) < len(grids) else grids[0]
            h = [v for k, v in hits.items() if k[2] == g]
            tr = int(sum((v["fetch"] + v["write"]) * v["dispatches"] for v in h) / sum(v["dispatches"] for v in h))
    roof = {"kernel": dom[0], "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr,
            "bytes_per_launch": int(bytes_per_launch), "avg_launch_us": round(per_launch_ms * 1e3, 2)}
    # the practical peak: the library's streaming copy of 1 GiB (2 GiB moved per launch)
    import ctypes as C
    from campfire_amd import _lib
    cg, cms = C.c_double(0), C.c_double(0)
    if _lib.lib().cf_copy_bandwidth(1 << 30, 10, C.byref(cg), C.byref(cms)) == 0:
        roof["copy_gbs"] = round(cg.value, 1)
        roof["frac_of_copy"] = round(achieved / cg.value, 4)
        log("[bench] copy kernel (k_bw_copy, 1 GiB): %.1f GB/s; %s at %.3f of it" % (cg.value, dom[0],
                                                                                 achieved / cg.value))
    if traffic:
        roof["traffic_calibration"] = traffic[1]
    elif traffic_err:
        roof["traffic_note"] = traffic_err
    lat_bound = {r[0]: round(r[1] / max(r[2], 1) * 1e3, 2) for r in rows if not r[3]}
    stage_sum = sum(r[1] for r in rows)

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%dx%d reacting gas step, %d-piece seeded pile, h=%.3g mm, dt=%.3g ms, "
                               "reference schedule (1 V-cycle, nu=2, 120-sweep coarsest)"
                               % (nx, ny, cfg["n"], cfg["dx"] * 1e3, cfg["dt"] * 1e3),
                   "grid": [nx, ny], "levels": len(levels), "diff_sub": sim.diff_sub, "seed": args.seed,
                   "leg_tiles": sim.leg_tiles()},
        "roofline": roof,
        # the whole step at the per-cell-step byte model, against the same peak
        "step_roofline": {"bytes_per_cell": round(step_bytes(sim.diff_sub), 1),
                          "achieved": round(value * 1e6 * step_bytes(sim.diff_sub) / 1e9, 1),
                          "frac": round(value * 1e6 * step_bytes(sim.diff_sub) / 1e9 / HBM_PEAK_GBS, 4)},
        "kernels_ms_per_step": {r[0]: round(r[1], 5) for r in rows},
        "kernels_gbs": {r[0]: round(r[4], 1) for r in rows if r[4]},
        "profiled_ms_per_step": round(stage_sum, 5),
        "latency_bound_us_per_launch": lat_bound,
    }
    if lat is not None:
        out["step_latency_ms"] = {"p50": round(float(np.percentile(lat, 50)), 5),
                                  "p99": round(float(np.percentile(lat, 99)), 5)}
    # the per-frame kernels of the same scene (geometry of 12 pieces; radiation
    # of 4 x 12 faces + 64 floor probes)
    geo_ms, rad_ms = sim.profile_frame(10)
    out["per_frame_ms"] = {"geometry": round(geo_ms, 5), "radiation": round(rad_ms, 5)}
    log("[bench] per frame: geometry %.4f ms, radiation %.4f ms" % (geo_ms, rad_ms))
    sim.close()
    if args.plain_steps > 0:
        out["plain_256"] = plain_256(args, args.plain_steps)
    if args.batch1024 > 0:
        out["batched_1024"] = batched_1024(args, args.batch1024)
    if args.batch > 0:
        out["batched"] = batched(args)
        if args.conv_steps > 0:
            out["converged_batch"] = converged_batch(args)
    if not args.no_modes:
        out["reference_modes"] = reference_modes(args)
    if args.serve_requests > 0:
        out["serving"] = serving(args)
    if not args.no_cpu:
        out["flame1d"] = flame1d_leg()
        log("[bench] CPU baseline (FP32 oracle)")
        out["cpu_baseline"] = cpu_baseline(cfg)
        # the same FP32 oracle at the smaller grids of the workload family
        sizes = {}
        for n in (256, 512):
            a2 = argparse.Namespace(**vars(args))
            a2.size = n
            sizes["%dx%d" % (n, n)] = cpu_baseline(bench_config(a2), budget_s=6.0, max_steps=2000)
        out["cpu_baseline_sizes"] = sizes
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
