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
    reqs = make_requests(10, rate=300.0, seed=3, sizes=(128, 256, 512, 1024), steps=(16, 64), width=0.64)
    _flush("[serve] sizes", [r["cfg"]["nx"] for r in reqs], "steps", [r["steps"] for r in reqs])
    assert {512, 1024} <= {r["cfg"]["nx"] for r in reqs}
    # room for a 1024^2, a 512^2 and two 256^2 scenes: later arrivals wait for memory
    budget = (scene_bytes(1024, 1024, 0.000625) + scene_bytes(512, 512, 0.00125)
              + 2 * scene_bytes(256, 256, 0.0025))
    stats, results = serve(reqs, budget, keep_results=True)
    _flush("[serve]", stats)
    assert stats["requests"] == 10 and len(results) == 10
    assert stats["max_live"] >= 2
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
    _flush("[serve] all 10 requests identical to their alone runs")


def test_serving_statistics_at_scale():
    """A short mixed-size serving session at the real sizes (256^2 to 1024^2,
    Poisson arrivals at 40/s, 200-800 steps each, a 1 GiB budget that holds
    about six 1024^2 scenes): requests overlap on the device, some wait for
    memory, and the aggregate rate is that of batched stepping, well above one
    scene at a time."""
    from campfire_amd.serve import make_requests, serve
    reqs = make_requests(16, rate=40.0, seed=5, steps=(200, 800))
    stats, _ = serve(reqs, 1 << 30)
    _flush("[serve] 16 requests (200-800 steps, 256^2..1024^2, 1 GiB budget):", stats)
    assert stats["step_latency_ms"]["p99"] >= stats["step_latency_ms"]["p50"] > 0
    assert stats["max_live"] >= 3
    assert stats["queue_s"]["p99"] > 0
    assert stats["device_busy_frac"] > 0.5
    assert stats["mcell_steps_per_s"] > 2000
