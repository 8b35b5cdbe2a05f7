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
