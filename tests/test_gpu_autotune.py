"""The schedule autotuner: every candidate is timed from the same device
snapshot, and the winner is the fastest one that converges."""
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_autotune_picks_fastest_converging_schedule():
    from campfire_amd.autotune import tune
    best, table = tune(sizes=(128, 256), steps=4, warm=16, log=lambda *a: print(*a, flush=True))
    sys.stdout.flush()
    assert best is not None
    ok = [r for r in table if r[2]]
    assert best == min(ok, key=lambda r: r[1])[0]
    # the snapshot makes the candidates comparable: the same schedule twice
    # gives the same V-cycle counts
    again, t2 = tune(sizes=(128, 256), steps=4, warm=16, candidates=[{k: best[k] for k in ("nu", "coarse_iters", "krylov")}] * 2)
    assert t2[0][3] == t2[1][3]


def test_leg_tiles_change_speed_not_results():
    """Every leg tile choice gives bitwise the same fields; the tuner keeps
    the fastest choice it timed."""
    import numpy as np
    from campfire_amd import scenes
    from campfire_amd.autotune import probe_config, tune_leg_tiles
    from campfire_amd.fluid import FluidSim
    cfg = probe_config(256, 3)
    out = []
    for tiles in (None, [64, 64], [32, 16], [48, 48]):   # 256^2 has two tiled levels
        g = FluidSim({"nx": 256, "ny": 256, "dt": cfg["dt"]}, dx=cfg["dx"])
        scenes.apply(g, cfg)
        if tiles:
            g.set_leg_tiles(tiles)
            # a level whose legs run in a two-level pair reports 0
            assert all(a in (b, 0) for a, b in zip(g.leg_tiles(), tiles)) and g.leg_tiles()[0] == tiles[0]
        g.frame_all(8, True)
        g.step(8)
        out.append({k: g.read_field(k) for k in ("T", "U", "V", "p", "Y")})
        print("[tiles]", tiles or g.leg_tiles(), flush=True)
        g.close()
    for o in out[1:]:
        for k in o:
            assert np.array_equal(o[k], out[0][k]), k
    tiles, table = tune_leg_tiles(256, 256, cfg["dx"], cfg["dt"], steps=10, log=lambda *a: print(*a, flush=True))
    best = min(ms for _, ms in table)
    chosen = [ms for t, ms in table if t == tiles]
    assert chosen and min(chosen) <= 1.02 * best / 0.99
