"""Autotuners: the multigrid schedule for residual-controlled stepping, and
the launch shape (tile code: square tiles of edge 16, 32 or 64, or the wide
64-cell tiles, code 48) of the tiled multigrid legs per level.

The leg tiles do not change results (a tile with a halo of 2 nu + 1 cells
reproduces the global red-black half-sweeps exactly), only speed, which
depends on the level's size, the scenes launched with it and the device:

    tiles, table = tune_leg_tiles(1024, 1024, 0.00125, 0.00025)
    sim.set_leg_tiles(tiles)
    tiles, table = tune_pool_leg_tiles(pool, views)   # a pool's depths, in place

Schedule autotuner:

For a mix of scene sizes it builds a probe pool (one seeded burning scene
per size), warms it up, snapshots every scene on the device, and then, from
the same snapshot, times a few batched steps under each candidate schedule
(smoother sweeps nu, coarsest solve: direct FP64 or a number of red-black
sweeps, the V-cycle alone on the defect or as the preconditioner of
flexible conjugate gradients).  The fastest schedule whose every step reaches the tolerance wins.
The choice depends on the hardware and the mix, which is why it is measured
rather than fixed.

    best, table = tune(sizes=(256, 512, 1024), tol=1e-6)
    pool.set_schedule(**best)
"""
import time

import numpy as np

from . import scenes
from .fluid import FluidSim
from .pool import ScenePool, scene_bytes

CANDIDATES = [{"nu": nu, "coarse_iters": ci, "krylov": kr} for kr in (0, 1) for nu in (1, 2, 3) for ci in (0, 24, 120)]


def probe_config(n, seed, width=1.28):
    h = width / n
    pieces, k = scenes.pile(8, seed, width)
    cx = float(np.mean(pieces.reshape(64, 20)[:k, 0]))
    return {"nx": n, "ny": n, "dx": h, "dt": 0.5 * h / 2.5, "pieces": pieces, "n": k,
            "params": scenes.air_params(), "disc": (cx, 0.06, 0.03, 1200.0)}


def tune(sizes=(256, 512, 1024), tol=1e-6, refine=1, cfl=0.5, umin=1.0, vcycles=80, warm=24, steps=6, seed=1,
         candidates=None, smoother_fp16=0, log=None):
    """Returns (best schedule dict, [(schedule, ms per step, converged, V-cycles per scene)])."""
    log = log or (lambda *a: None)
    candidates = candidates or CANDIDATES
    base = {"vcycles": vcycles, "tol": tol, "refine": refine, "cfl": cfl, "umin": umin, "smoother_fp16": smoother_fp16}
    cfgs = [probe_config(n, seed + q) for q, n in enumerate(sizes)]
    pool = ScenePool(sum(2 * scene_bytes(c["nx"], c["ny"], c["dx"]) for c in cfgs) + (64 << 20))
    pool.set_schedule(nu=2, coarse_iters=120, **dict(base, smoother_fp16=0))
    probes, twins = [], []
    for c in cfgs:
        slot, sim = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
        scenes.apply(sim, c)
        probes.append(sim)
    pool.frame(warm, True)
    pool.synchronize()
    for c, sim in zip(cfgs, probes):
        t = FluidSim({"nx": c["nx"], "ny": c["ny"], "dt": c["dt"]}, dx=c["dx"])
        t.copy_state_from(sim)
        twins.append(t)
    table = []
    for cand in candidates:
        sched = dict(base, **cand)
        pool.set_schedule(**sched)
        for sim, t in zip(probes, twins):
            sim.copy_state_from(t)
        pool.step(1)                       # builds and captures the batched step
        for sim, t in zip(probes, twins):
            sim.copy_state_from(t)
        ok = True
        cycles = []
        t0 = time.perf_counter()
        elapsed = 0.0
        for q in range(steps):
            t0 = time.perf_counter()
            pool.step(1)
            pool.synchronize()
            elapsed += time.perf_counter() - t0
            infos = [sim.info() for sim in probes]
            ok = ok and all(i.last_resid < tol for i in infos)
            cycles.append([int(i.last_vcycles) for i in infos])
        ms = elapsed * 1e3 / steps
        table.append((sched, ms, ok, cycles))
        log("[autotune] nu=%d coarse=%-3s %-6s %.3f ms/step %s cycles %s"
            % (sched["nu"], sched["coarse_iters"] if sched["coarse_iters"] else "LU",
               "FCG" if sched.get("krylov") else "defect", ms, "ok" if ok else "NOT CONVERGED", cycles[-1]))
    for t in twins:
        t.close()
    pool.close()
    good = [r for r in table if r[2]]
    best = min(good, key=lambda r: r[1])[0] if good else None
    return best, table


TILES = (16, 32, 48, 64)


def _timed_steps(restore, step, sync, steps, reps=2):
    """ms per step of `steps` steps launched back to back from the restored
    state (the throughput the benchmarks measure), the better of `reps` runs.
    The first replay of a rebuilt step runs slower than the next ones, so the
    caller builds and runs the step once before."""
    best = float("inf")
    for _ in range(reps):
        restore()
        sync()
        t0 = time.perf_counter()
        step(steps)
        sync()
        best = min(best, (time.perf_counter() - t0) * 1e3 / steps)
    return best


def _step_ms(sim, twin, steps):
    sim.copy_state_from(twin)
    sim.step(1)                       # captures the step graph of the current tiles
    sim.synchronize()
    return _timed_steps(lambda: sim.copy_state_from(twin), sim.step, sim.synchronize, steps)


def tune_pool_leg_tiles(pool, views, steps=10, log=None):
    """The leg tile code of each depth of a pool's tiled levels, for the
    pool's scenes and schedule as they are: depth by depth from the finest,
    each candidate is timed over `steps` pool steps from the same device
    snapshot of every scene (launched back to back, the better of two runs),
    and a change is kept only when it beats the best by more than 1 %.  The scenes are left in their
    snapshot state and the pool with the step of the chosen tiles built (and
    run once, so its first replay is not in the caller's timing).  Returns (tiles,
    [(tiles tried, ms per step)])."""
    log = log or (lambda *a: None)
    twins = []
    for v in views:
        t = FluidSim({"nx": v.nx, "ny": v.ny, "dt": v.dt}, dx=v.dx)
        t.copy_state_from(v)
        twins.append(t)

    def restore():
        pool.synchronize()            # no pool step may still write the scenes
        for v, t in zip(views, twins):
            v.copy_state_from(t)

    def timed(tiles):
        pool.set_leg_tiles(tiles)
        restore()
        pool.step(1)                  # rebuilds and captures the batched step
        return _timed_steps(restore, pool.step, pool.synchronize, steps)

    pool.set_leg_tiles([])
    tiles = pool.leg_tiles()
    timed(tiles)                      # the device's first timed steps run slow: discarded
    best_ms = timed(tiles)
    table = [(list(tiles), best_ms)]
    log("[autotune] pool leg tiles %s (size rule on each depth's tile count): %.3f ms/step" % (tiles, best_ms))
    for d in range(len(tiles)):
        for t in TILES:
            if t == tiles[d]:
                continue
            trial = list(tiles)
            trial[d] = t
            ms = timed(trial)
            table.append((trial, ms))
            log("[autotune] pool leg tiles %s: %.3f ms/step" % (trial, ms))
            if ms < 0.99 * best_ms:
                best_ms, tiles = ms, trial
    pool.set_leg_tiles(tiles)
    pool.step(1)                      # the step of the chosen tiles, built, captured and run once
    pool.synchronize()
    for v, t in zip(views, twins):
        v.copy_state_from(t)
        t.close()
    return tiles, table


def tune_leg_tiles(nx, ny, dx, dt, steps=20, warm=16, seed=1, log=None):
    """The leg tile code of each tiled level of an nx x ny scene (reference
    schedule), chosen level by level from the finest: each candidate is timed
    over `steps` steps from the same device snapshot of a burning probe scene
    (launched back to back, the better of two runs).  Returns (tiles, [(tiles
    tried, ms per step)])."""
    log = log or (lambda *a: None)
    cfg = probe_config(nx, seed, width=nx * dx)
    cfg.update(ny=ny, dx=dx, dt=dt)
    sim = FluidSim({"nx": nx, "ny": ny, "dt": dt}, dx=dx)
    scenes.apply(sim, cfg)
    for f in range(max(1, warm // 8)):
        sim.frame = f
        sim.frame_all(8, True)
    sim.synchronize()
    twin = FluidSim({"nx": nx, "ny": ny, "dt": dt}, dx=dx)
    twin.copy_state_from(sim)
    tiles = sim.leg_tiles()
    _step_ms(sim, twin, steps)        # the device's first timed steps run slow: discarded
    table = [(list(tiles), _step_ms(sim, twin, steps))]
    log("[autotune] %dx%d leg tiles %s (size rule): %.4f ms/step" % (nx, ny, tiles, table[0][1]))
    best_ms = table[0][1]
    for l in range(len(tiles)):
        if tiles[l] == 0:  # the level's legs run in a two-level pair, whose tile is fixed
            continue
        for t in TILES:
            if t == tiles[l]:
                continue
            trial = list(tiles)
            trial[l] = t
            sim.set_leg_tiles(trial)
            ms = _step_ms(sim, twin, steps)
            table.append((trial, ms))
            log("[autotune] %dx%d leg tiles %s: %.4f ms/step" % (nx, ny, trial, ms))
            if ms < 0.99 * best_ms:  # a change must beat the noise of the timing
                best_ms, tiles = ms, trial
        sim.set_leg_tiles(tiles)
    twin.close()
    sim.close()
    return tiles, table
