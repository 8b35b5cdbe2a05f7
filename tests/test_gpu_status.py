"""Status reporting and pool bookkeeping through the C ABI:
- a NaN written into T or U is reported as CF_ENONFINITE by the status call,
  and the flag clears once read;
- a residual target a single V-cycle cannot meet is reported as CF_ENOCONV;
- a pool schedule that one scene rejects changes no scene's schedule, and a
  scene of a pool refuses a schedule of its own;
- a slot released and re-admitted with a scene of another size gives the
  same frame as that scene alone (the per-frame tables are rebuilt)."""
import sys

import numpy as np
import pytest

from campfire_amd import _lib as L
from campfire_amd import scenes

pytestmark = pytest.mark.gpu


def _flush(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _scene(sim, nx, ny, dx, seed, npieces=3):
    pieces, n = scenes.pile(npieces, seed, nx * dx, spread=nx * dx / 3)
    sim.reset()
    sim.upload_pieces(pieces, n)
    st = scenes.random_state(nx, ny, seed, vmax=1.5)
    for name in ("T", "T2", "Th"):
        sim.write_field(name, st["T"])
    for name in ("Y", "Y2", "Yh"):
        sim.write_field(name, st["Y"])
    for name in ("U", "U2", "Uh"):
        sim.write_field(name, st["U"])
    for name in ("V", "V2", "Vh"):
        sim.write_field(name, st["V"])
    sim.write_params()
    sim.clear_geometry()  # the face fractions: without them every face is closed
    sim.encode_geometry()


def test_nan_in_T_is_reported_then_cleared():
    from campfire_amd.fluid import FluidSim
    _flush("[status] 96x80: NaN written into T, one step")
    g = FluidSim({"nx": 96, "ny": 80, "dt": 0.0016}, dx=0.004)
    _scene(g, 96, 80, 0.004, 3)
    g.step(1)
    g.status()  # a healthy step
    T = g.read_field("T")
    T[40 * 96 + 50] = np.nan
    g.write_field("T", T)
    g.step(1)
    with pytest.raises(L.CampfireError) as e:
        g.status()
    assert e.value.code == L.CF_ENONFINITE
    assert "T or Y" in str(e.value)
    _scene(g, 96, 80, 0.004, 3)  # healthy fields again; the flag was cleared by the read
    g.step(1)
    g.status()
    g.close()


def test_nan_in_U_is_reported_by_the_cfl_pass():
    from campfire_amd.fluid import FluidSim
    _flush("[status] 96x80: NaN written into U, CFL step")
    g = FluidSim({"nx": 96, "ny": 80, "dt": 0.0016}, dx=0.004)
    g.set_schedule(cfl=0.5, umin=1.0)
    _scene(g, 96, 80, 0.004, 4)
    U = g.read_field("U")
    U[10 * 97 + 30] = np.nan
    g.write_field("U", U)
    g.step(1)
    with pytest.raises(L.CampfireError) as e:
        g.status()
    assert e.value.code == L.CF_ENONFINITE
    assert "U or V" in str(e.value)
    g.close()


def test_missed_residual_target_is_reported():
    from campfire_amd.fluid import FluidSim
    _flush("[status] 128x96: one V-cycle against tol 1e-12")
    g = FluidSim({"nx": 128, "ny": 96, "dt": 0.0016}, dx=0.004)
    _scene(g, 128, 96, 0.004, 5)
    g.set_schedule(nu=2, vcycles=1, tol=1e-12, coarse_iters=120)
    g.step(1)
    with pytest.raises(L.CampfireError) as e:
        g.status()
    assert e.value.code == L.CF_ENOCONV
    # the same scene converges with enough defect-corrected cycles
    g.set_schedule(nu=2, vcycles=60, tol=1e-6, coarse_iters=120, refine=1)
    g.step(1)
    g.status()
    g.close()


def test_pool_schedule_is_all_or_nothing():
    from campfire_amd.pool import ScenePool
    _flush("[status] pool of 96x80 and 136x4 (a 68-wide coarsest level)")
    pool = ScenePool(1 << 30)
    s0, a = pool.admit(96, 80, 0.004, 0.0016)
    s1, b = pool.admit(136, 4, 0.004, 0.0016)
    before = (a.get_schedule(), b.get_schedule())
    # the direct coarsest solve needs a coarsest level <= 32 wide: the second scene refuses
    with pytest.raises(L.CampfireError) as e:
        pool.set_schedule(nu=2, vcycles=1, coarse_iters=0)
    assert e.value.code == L.CF_EINVAL
    assert (a.get_schedule(), b.get_schedule()) == before
    # a scene of a pool takes the pool's schedule only
    with pytest.raises(L.CampfireError):
        a.set_schedule(nu=3)
    a.set_schedule(**before[0])  # the pool's own schedule is accepted as a no-op
    pool.close()


def test_pool_slot_reuse_rebuilds_frame_tables():
    from campfire_amd.fluid import FluidSim
    from campfire_amd.pool import ScenePool
    _flush("[status] pool: frame, release, re-admit another size into the slot, frame")
    pool = ScenePool(1 << 30)
    s0, a = pool.admit(128, 96, 0.004, 0.0016)
    s1, b = pool.admit(96, 80, 0.004, 0.0016)
    _scene(a, 128, 96, 0.004, 6, npieces=4)
    _scene(b, 96, 80, 0.004, 7, npieces=4)
    pool.frame(2, True)
    pool.synchronize()
    pool.release(s0)
    s2, c = pool.admit(160, 224, 0.0035, 0.0007)
    assert s2 == s0
    _scene(c, 160, 224, 0.0035, 8, npieces=4)
    pool.frame(2, True)
    pool.synchronize()
    pool.status()
    alone = FluidSim({"nx": 160, "ny": 224, "dt": 0.0007}, dx=0.0035)
    _scene(alone, 160, 224, 0.0035, 8, npieces=4)
    alone.frame_all(2, True)
    alone.synchronize()
    for name in ("T", "Y", "U", "V", "p", "cellG", "fx", "fy", "faceW", "rad"):
        x, y = c.read_field(name), alone.read_field(name)
        assert np.array_equal(x, y), name
    alone.close()
    pool.close()


def test_pool_status_names_the_failing_slot():
    from campfire_amd.pool import ScenePool
    _flush("[status] pool: NaN in the second scene")
    pool = ScenePool(1 << 30)
    s0, a = pool.admit(96, 80, 0.004, 0.0016)
    s1, b = pool.admit(64, 64, 0.004, 0.0016)
    _scene(a, 96, 80, 0.004, 9)
    _scene(b, 64, 64, 0.004, 10)
    T = b.read_field("T")
    T[5] = np.inf
    b.write_field("T", T)
    pool.step(1)
    with pytest.raises(L.CampfireError) as e:
        pool.status()
    assert e.value.code == L.CF_ENONFINITE
    assert ("slot %d" % s1) in str(e.value)
    pool.close()
