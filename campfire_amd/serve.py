"""Serving loop: continuous batching of gas-scene requests on one GPU.

Requests arrive by a seeded Poisson process in wall-clock time (`rate`
requests per second from the start of the session), each for a scene of
256^2, 512^2 or 1024^2 cells and a number of steps.  Each round the loop
  1. takes in the requests whose arrival time has passed,
  2. admits waiting requests, in arrival order, while their device memory
     fits the pool's HBM budget (a request that does not fit waits for memory
     to free up; later, smaller ones may go ahead of it), and rebuilds the
     batched step for the new scene set at once (cf_pool_prepare), timed,
  3. advances every live scene together for as many steps as the nearest
     finishing request still needs, at most `max_round` (the steps are queued
     back to back; the host waits once, at the end of the round),
  4. releases the requests that are done, after taking their results.
Requests that arrive while others run join at the next round and share every
launch with them.  A scene's results are bitwise what it gives alone.

Reported: p50/p99 of the device time of one batched step (every live scene
advancing one step), aggregate cell-steps/s over the session's wall time, the
queueing delay from arrival to admission and the time of the rebuilds.

    python -m campfire_amd.serve --requests 24 --rate 4 --budget-gb 2
"""
import argparse
import json
import sys
import time

import numpy as np

from . import scenes
from .pool import ScenePool


def make_requests(n, rate, seed, sizes=(256, 512, 1024), steps=(500, 5000), width=1.28):
    """n requests: Poisson arrivals at `rate` per second (arrival times in
    seconds), sizes and step counts drawn from the seeded PCG32 stream."""
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


def _pct(a, q):
    return round(float(np.percentile(a, q)), 4) if len(a) else None


def serve(requests, budget, schedule=None, keep_results=False, max_round=64, log=None):
    """Run the requests to completion; returns (stats, results by request id)."""
    pool = ScenePool(budget)
    if schedule:
        pool.set_schedule(**schedule)
    pending = sorted(requests, key=lambda r: r["arrive"])
    waiting, live, results, done = [], {}, {}, []
    step_ms, rebuild_ms, queue_s = [], [], []
    cell_steps, steps, max_live = 0, 0, 0
    t_start = time.perf_counter()
    clock = lambda: time.perf_counter() - t_start
    while pending or waiting or live:
        now = clock()
        while pending and pending[0]["arrive"] <= now:
            waiting.append(pending.pop(0))
        # admission by memory (arrival order; smaller requests may pass a blocked one)
        changed = False
        for r in list(waiting):
            c = r["cfg"]
            got = pool.admit(c["nx"], c["ny"], c["dx"], c["dt"])
            if got is None:
                continue
            slot, sim = got
            scenes.apply(sim, c)
            sim.clear_geometry()
            sim.encode_geometry()
            live[slot] = {"req": r, "sim": sim, "left": r["steps"], "t_admit": clock()}
            queue_s.append(clock() - r["arrive"])
            waiting.remove(r)
            changed = True
        if not live:
            if waiting:
                raise ValueError("request %d needs more than the pool's HBM budget" % waiting[0]["id"])
            if pending:   # idle until the next arrival
                time.sleep(max(0.0, pending[0]["arrive"] - clock()))
            continue
        if changed:
            t0 = time.perf_counter()
            pool.prepare()
            rebuild_ms.append((time.perf_counter() - t0) * 1e3)
        max_live = max(max_live, len(live))
        # advance every live scene together
        k = min(max_round, min(v["left"] for v in live.values()))
        lat = pool.step_latencies(k)
        step_ms.extend(float(x) for x in lat)
        cells = sum(v["sim"].nx * v["sim"].ny for v in live.values())
        cell_steps += cells * k
        steps += k
        # finished requests leave
        released = False
        for slot in [s for s, v in live.items() if v["left"] - k <= 0]:
            v = live.pop(slot)
            if keep_results:
                results[v["req"]["id"]] = result_of(v["sim"])
            done.append({"id": v["req"]["id"], "n": v["sim"].nx, "steps": v["req"]["steps"],
                         "wall_s": clock() - v["t_admit"]})
            pool.release(slot)
            released = True
        for v in live.values():
            v["left"] -= k
        if released and live:
            t0 = time.perf_counter()
            pool.prepare()
            rebuild_ms.append((time.perf_counter() - t0) * 1e3)
        if log and released:
            log("[serve] t %.2f s: %d live, %d waiting, %d done" % (clock(), len(live), len(waiting), len(done)))
    wall = clock()
    pool.close()
    stats = {"requests": len(requests), "pool_steps": steps, "wall_s": round(wall, 3),
             "arrival_span_s": round(max(r["arrive"] for r in requests), 3),
             "cell_steps": int(cell_steps), "max_live": max_live,
             "mcell_steps_per_s": round(cell_steps / wall / 1e6, 1),
             "step_latency_ms": {"p50": _pct(step_ms, 50), "p99": _pct(step_ms, 99)},
             "queue_s": {"p50": _pct(queue_s, 50), "p99": _pct(queue_s, 99)},
             "rebuild_ms": {"p50": _pct(rebuild_ms, 50), "p99": _pct(rebuild_ms, 99), "count": len(rebuild_ms)},
             "device_busy_frac": round(sum(step_ms) / 1e3 / wall, 3)}
    return stats, results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=24)
    ap.add_argument("--rate", type=float, default=4.0, help="arrivals per second")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--budget-gb", type=float, default=2.0)
    ap.add_argument("--min-steps", type=int, default=500)
    ap.add_argument("--max-steps", type=int, default=5000)
    a = ap.parse_args()
    reqs = make_requests(a.requests, a.rate, a.seed, steps=(a.min_steps, a.max_steps))
    stats, _ = serve(reqs, int(a.budget_gb * (1 << 30)), log=lambda *x: print(*x, file=sys.stderr, flush=True))
    print(json.dumps(stats))


if __name__ == "__main__":
    main()
