"""Launch fusions of the multigrid against the launches they replace, every
field and every level's p bitwise the same after burning steps:
- the two-level legs (k_mg_down2 / k_mg_up2: the legs of levels 1-2, 3-4, ...
  two per launch) against the one-level legs, for both pair tiles, nu = 1, 2,
  3, square and oblong grids with partial tiles, and a residual-controlled
  FCG schedule;
- the projection inside the last V-cycle's level-0 up leg
  (k_mg_up_project) against the up leg followed by k_project, for one and two
  V-cycles per step and a grid whose bottom kernel holds every level.
The fused V-cycle against the oracle's per-dispatch V-cycle is in
test_gpu_parity.py; this file pins that the fusions change no bit of it."""
import os
import sys

import numpy as np
import pytest

from campfire_amd import scenes

pytestmark = pytest.mark.gpu


def _flush(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scene(g, nx, ny, dx, seed):
    pieces, n = scenes.pile(8, seed, nx * dx, spread=nx * dx / 3)
    g.reset()
    g.upload_pieces(pieces, n)
    st = scenes.random_state(nx, ny, seed, vmax=1.5)
    for name in ("T", "Y", "U", "V"):
        g.write_field(name, st[name])
    g.write_params()
    g.clear_geometry()
    g.encode_geometry()


def _run(nx, ny, pairs, tile, sched, steps, seed=3, up_project=True):
    from campfire_amd.fluid import FluidSim
    old = {k: os.environ.get(k) for k in ("CF_LEG_PAIRS", "CF_PAIR_TILE", "CF_UP_PROJECT")}
    os.environ["CF_LEG_PAIRS"] = "1" if pairs else "0"
    os.environ["CF_PAIR_TILE"] = str(tile)
    os.environ["CF_UP_PROJECT"] = "1" if up_project else "0"
    try:
        dx = 1.28 / max(nx, ny)
        g = FluidSim({"nx": nx, "ny": ny, "dt": 0.4 * dx / 2.5}, dx=dx)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    if sched:
        g.set_schedule(**sched)
    _scene(g, nx, ny, dx, seed)
    g.step(steps)
    g.synchronize()
    out = {k: g.read_field(k) for k in ("T", "Y", "U", "V")}
    for l in range(len(g.levels)):
        out["p" if l == 0 else "p%d" % l] = g.read_field("p" if l == 0 else "p%d" % l)
    tiles = g.leg_tiles()
    g.close()
    return out, tiles


CASES = [
    # (nx, ny, nu, pair tile): 1024^2 pairs levels 1-2 and 3-4 (the 64^2 level
    # leaves the bottom kernel); 160x224 and 224x320 pair 1-2 with partial tiles
    (1024, 1024, 2, 0), (1024, 1024, 2, 16), (512, 512, 1, 8), (512, 512, 3, 16),
    (160, 224, 2, 8), (224, 320, 3, 8), (256, 256, 1, 16), (352, 512, 2, 16),
]


@pytest.mark.parametrize("nx,ny,nu,tile", CASES)
def test_pairs_equal_single_legs(nx, ny, nu, tile):
    _flush("[pairs] %dx%d nu=%d pair tile %s" % (nx, ny, nu, tile or "rule"))
    sched = {"nu": nu, "vcycles": 1, "coarse_iters": 120}
    a, ta = _run(nx, ny, True, tile, sched, 6)
    b, tb = _run(nx, ny, False, tile, sched, 6)
    _flush("  leg tiles with pairs %s, without %s" % (ta, tb))
    assert 0 in ta, "no level ran in a pair"
    for k in b:
        assert np.isfinite(b[k]).all(), k
        assert np.array_equal(a[k], b[k]), (k, float(np.max(np.abs(a[k] - b[k]))))


def test_pairs_equal_single_legs_residual_controlled():
    """FCG with the V-cycle as preconditioner and FP64 defect correction
    (tol 1e-6, at most 30 cycles): the same cycle counts and bits."""
    sched = {"nu": 2, "vcycles": 30, "tol": 1e-6, "coarse_iters": 24, "refine": 1, "krylov": 1}
    _flush("[pairs] 512x512 FCG to 1e-6")
    a, _ = _run(512, 512, True, 0, sched, 4)
    b, _ = _run(512, 512, False, 0, sched, 4)
    for k in b:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("nx,ny,nu,vcycles", [(1024, 1024, 2, 1), (160, 224, 3, 2), (96, 80, 1, 1), (72, 48, 2, 1)])
def test_projection_in_the_up_leg_equals_its_own_launch(nx, ny, nu, vcycles):
    """72x48 has three levels, all in the bottom kernel: no level-0 leg, so
    the projection keeps its own launch there."""
    _flush("[fused projection] %dx%d nu=%d, %d V-cycles" % (nx, ny, nu, vcycles))
    sched = {"nu": nu, "vcycles": vcycles, "coarse_iters": 120}
    a, _ = _run(nx, ny, True, 0, sched, 5, up_project=True)
    b, _ = _run(nx, ny, True, 0, sched, 5, up_project=False)
    for k in b:
        assert np.isfinite(b[k]).all(), k
        assert np.array_equal(a[k], b[k]), (k, float(np.max(np.abs(a[k] - b[k]))))
