"""The serving loop: Poisson arrivals, admission by HBM budget, continuous
batching; every request's result is bitwise the same as the request run
alone."""
import sys

import numpy as np
import pytest

from campfire_amd import scenes

pytestmark = pytest.mark.gpu


def _flush(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_served_results_equal_alone_runs():
    from campfire_amd.fluid import FluidSim
    from campfire_amd.pool import scene_bytes
    from campfire_amd.serve import make_requests, serve
    reqs = make_requests(8, rate=0.08, seed=3, sizes=(128, 256), steps=(16, 64), width=0.64)
    # room for about three scenes: later arrivals wait for memory
    budget = 3 * scene_bytes(256, 256, 0.0025) + scene_bytes(128, 128, 0.005)
    stats, results = serve(reqs, budget, keep_results=True)
    _flush("[serve]", stats)
    assert stats["requests"] == 8 and len(results) == 8
    assert stats["mean_queued_steps"] >= 0
    for r in reqs:
        c = r["cfg"]
        a = FluidSim({"nx": c["nx"], "ny": c["ny"], "dt": c["dt"]}, dx=c["dx"])
        scenes.apply(a, c)
        a.clear_geometry()
        a.encode_geometry()
        a.step(r["steps"])
        a.synchronize()
        got = results[r["id"]]
        for name in ("T", "U", "V", "Y"):
            assert np.array_equal(a.read_field(name), got[name]), (r["id"], name)
        assert np.array_equal(a.readback_now()["acc"], got["acc"]), r["id"]
        a.close()
    _flush("[serve] all 8 requests identical to their alone runs")


def test_serving_statistics_at_scale():
    """A short mixed-size serving session at the real sizes (256^2 to 1024^2):
    p50/p99 step latency and aggregate cell-steps/s."""
    from campfire_amd.serve import make_requests, serve
    reqs = make_requests(16, rate=0.2, seed=5, steps=(40, 200))
    stats, _ = serve(reqs, 4 << 30)
    _flush("[serve] 16 requests (40-200 steps, 256^2..1024^2, 4 GiB budget):", stats)
    assert stats["step_latency_ms"]["p99"] >= stats["step_latency_ms"]["p50"] > 0
    assert stats["mcell_steps_per_s"] > 0
