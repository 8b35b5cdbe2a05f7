"""Serving loop: continuous batching of gas-scene requests on one GPU.

Requests arrive by a seeded Poisson process (times in pool steps), each for a
scene of 256^2, 512^2 or 1024^2 cells and a number of steps.  Every round the
loop
  1. admits waiting requests, in arrival order, while their device memory
     fits the pool's HBM budget (a request that does not fit waits for
     memory to free up; later, smaller ones may go ahead of it),
  2. advances every live scene together, one batched step graph per step,
     for as many steps as the nearest finishing request still needs (<= 8),
  3. releases the requests that are done, after taking their results.
Requests that arrive while others run join at the next round and share
every launch with them.  A scene's results are bitwise what it gives alone.

    python -m campfire_amd.serve --requests 24 --rate 0.05 --budget-gb 8
"""
import argparse
import json
import sys
import time

import numpy as np

from . import scenes
from .pool import ScenePool, scene_bytes


def make_requests(n, rate, seed, sizes=(256, 512, 1024), steps=(500, 5000), width=1.28):
    """n requests: Poisson arrivals at `rate` per pool step, sizes and step
    counts drawn from the seeded PCG32 stream."""
    g = scenes.PCG32(seed)
    t = 0.0
    out = []
    for q in range(n):
        t += -np.log(max(1.0 - g.uniform(0.0, 1.0), 1e-12)) / rate
        nx = sizes[min(int(g.uniform(0, len(sizes))), len(sizes) - 1)]
        k = int(steps[0] + g.uniform(0.0, 1.0) * (steps[1] - steps[0]))
        h = width / nx
        npc = 2 + min(int(g.uniform(0, 11)), 10)
        pieces, npieces = scenes.pile(npc, seed * 7919 + q, width)
        cx = float(np.mean(pieces.reshape(64, 20)[:npieces, 0]))
        out.append({"id": q, "arrive": t, "steps": k,
                    "cfg": {"nx": nx, "ny": nx, "dx": h, "dt": 0.5 * h / 2.5, "pieces": pieces, "n": npieces,
                            "params": scenes.air_params(), "disc": (cx, 0.06, 0.03, 1200.0)}})
    return out


def result_of(sim):
    """What a finished request returns: its final fields and accumulators."""
    return {"T": sim.read_field("T"), "U": sim.read_field("U"), "V": sim.read_field("V"), "Y": sim.read_field("Y"),
            "acc": sim.readback_now()["acc"]}


def serve(requests, budget, schedule=None, keep_results=False, max_round=8, log=None):
    """Run the requests to completion; returns (stats, results by request id)."""
    pool = ScenePool(budget)
    if schedule:
        pool.set_schedule(**schedule)
    pending = sorted(requests, key=lambda r: r["arrive"])
    waiting, live, results, done = [], {}, {}, []
    now = 0                      # pool steps taken so far
    step_ms = []                 # wall time per step, per round
    cell_steps = 0
    t_start = time.perf_counter()
    while pending or waiting or live:
        while pending and pending[0]["arrive"] <= now:
            waiting.append(pending.pop(0))
        # 1. admission by memory (arrival order, smaller requests may pass a blocked one)
        for r in list(waiting):
            c = r["cfg"]
            got = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
            if got is None:
                continue
            slot, sim = got
            scenes.apply(sim, c)
            sim.clear_geometry()
            sim.encode_geometry()
            live[slot] = {"req": r, "sim": sim, "left": r["steps"], "admitted": now,
                          "t_admit": time.perf_counter()}
            waiting.remove(r)
        if not live:
            if waiting:
                raise ValueError("request %d needs more than the pool's HBM budget" % waiting[0]["id"])
            now = int(np.ceil(pending[0]["arrive"])) if pending else now
            continue
        # 2. advance every live scene together
        k = min(max_round, min(v["left"] for v in live.values()))
        if pending:
            k = max(1, min(k, int(np.ceil(pending[0]["arrive"] - now))))
        t0 = time.perf_counter()
        pool.step(k)
        pool.synchronize()
        dt = time.perf_counter() - t0
        step_ms.extend([dt * 1e3 / k] * k)
        cells = sum(v["sim"].nx * v["sim"].ny for v in live.values())
        cell_steps += cells * k
        now += k
        # 3. finished requests leave
        for slot in [s for s, v in live.items() if v["left"] - k <= 0]:
            v = live.pop(slot)
            if keep_results:
                results[v["req"]["id"]] = result_of(v["sim"])
            done.append({"id": v["req"]["id"], "n": v["sim"].nx, "steps": v["req"]["steps"],
                         "queued_steps": v["admitted"] - v["req"]["arrive"],
                         "wall_s": time.perf_counter() - v["t_admit"]})
            pool.release(slot)
        for v in live.values():
            v["left"] -= k
        if log and len(done) % 8 == 0 and done:
            log("[serve] step %d: %d live, %d waiting, %d done" % (now, len(live), len(waiting), len(done)))
    wall = time.perf_counter() - t_start
    pool.close()
    lat = np.array(step_ms)
    stats = {"requests": len(requests), "pool_steps": now, "wall_s": round(wall, 3),
             "cell_steps_per_s": round(cell_steps / wall, 1),
             "mcell_steps_per_s": round(cell_steps / wall / 1e6, 1),
             "step_latency_ms": {"p50": round(float(np.percentile(lat, 50)), 4),
                                 "p99": round(float(np.percentile(lat, 99)), 4)},
             "mean_queued_steps": round(float(np.mean([d["queued_steps"] for d in done])), 1)}
    return stats, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=0.05, help="arrivals per pool step")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=8.0)
    ap.add_argument("--min-steps", type=int, default=500)
    ap.add_argument("--max-steps", type=int, default=5000)
    a = ap.parse_args()
    reqs = make_requests(a.requests, a.rate, a.seed, steps=(a.min_steps, a.max_steps))
    stats, _ = serve(reqs, int(a.budget_gb * (1 << 30)), log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(stats))


if __name__ == "__main__":
    main()
