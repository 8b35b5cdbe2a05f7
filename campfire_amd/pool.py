"""ScenePool — many gas scenes on one GPU, stepped together.

    pool = ScenePool(hbm_budget=8 << 30)
    slot, sim = pool.admit(nx=512, ny=512, dx=2.5e-3, dt=5e-4)   # sim: a FluidSim view
    scenes.apply(sim, cfg)                                       # params, pieces, fields as usual
    pool.frame(8)                                                # every scene: geometry, 8 steps, radiation
    pool.release(slot)

Each phase of a step is one launch over the tiles of every live scene, so the
latency-bound launches of small scenes overlap with the large ones.  A
scene's results are the same alone and in any mix (the same tiles and the
same arithmetic).  No counterpart in the reference, which runs one FluidSim.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .fluid import FluidSim


def scene_bytes(nx, ny, dx=1e-3):
    """Device memory a scene of this grid needs (0 if the grid is not supported)."""
    return int(L.lib().cf_scene_bytes(int(nx), int(ny), float(dx)))


class ScenePool:
    def __init__(self, hbm_budget):
        h = C.c_void_p()
        L.check(L.lib().cf_pool_create(int(hbm_budget), C.byref(h)), "pool")
        self._h = h
        self._scenes = {}

    def close(self):
        if self._h:
            for v in self._scenes.values():
                v._h = None
            self._scenes.clear()
            L.lib().cf_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_schedule(self, nu=2, vcycles=1, tol=0.0, coarse_iters=120, smoother_fp16=0, cfl=0.0, umin=1.0, refine=0,
                     krylov=0):
        s = L.CfSchedule(nu, vcycles, tol, coarse_iters, smoother_fp16, cfl, umin, refine, krylov)
        L.check(L.lib().cf_pool_set_schedule(self._h, C.byref(s)), "pool schedule")

    def set_leg_tiles(self, tiles):
        """Leg tile code (16, 32, 64, or 48 for the wide tiles; 0 = the size rule on the depth's tile
        count) of each depth of the tiled multigrid levels; speed only."""
        a = np.asarray(tiles, np.int32)
        L.check(L.lib().cf_pool_set_leg_tiles(self._h, a.ctypes.data, int(a.size)), "pool set_leg_tiles")

    def leg_tiles(self):
        """The tile of each tiled depth for the live scenes."""
        a = np.zeros(8, np.int32)
        n = L.lib().cf_pool_get_leg_tiles(self._h, a.ctypes.data, int(a.size))
        L.check(0 if n >= 0 else n, "pool leg_tiles")
        return [int(x) for x in a[:n]]

    def admit(self, nx, ny, dx, dt):
        """(slot, FluidSim view), or None when the scene does not fit the budget."""
        d = L.CfSceneDesc(int(nx), int(ny), float(dx), float(dt))
        slot = C.c_int(-1)
        rc = L.lib().cf_pool_admit(self._h, C.byref(d), C.byref(slot))
        if rc == L.CF_EAGAIN:
            return None
        L.check(rc, "admit")
        h = C.c_void_p()
        L.check(L.lib().cf_pool_scene(self._h, slot.value, C.byref(h)), "pool scene")
        sim = FluidSim.borrow(h, {"nx": nx, "ny": ny, "dt": dt}, dx)
        self._scenes[slot.value] = sim
        return slot.value, sim

    def scene(self, slot):
        return self._scenes[slot]

    def release(self, slot):
        L.check(L.lib().cf_pool_release(self._h, int(slot)), "release")
        sim = self._scenes.pop(slot, None)
        if sim is not None:
            sim._h = None

    def prepare(self):
        """Builds the batched step of the current scenes now (tables, graphs)."""
        L.check(L.lib().cf_pool_prepare(self._h), "pool prepare")

    def step(self, n=1):
        L.check(L.lib().cf_pool_step(self._h, int(n)), "pool step")

    def frame(self, nsteps, with_radiation=True):
        L.check(L.lib().cf_pool_frame(self._h, int(nsteps), int(bool(with_radiation))), "pool frame")

    def synchronize(self):
        L.check(L.lib().cf_pool_synchronize(self._h), "pool sync")

    def status(self):
        """cf_pool_status: raises CampfireError naming the first failing slot
        (CF_ENONFINITE before CF_ENOCONV); every scene's flag is cleared."""
        slot = C.c_int(-1)
        L.check(L.lib().cf_pool_status(self._h, C.byref(slot)), "pool status")

    def info(self):
        i = L.CfPoolInfo()
        L.check(L.lib().cf_pool_get_info(self._h, C.byref(i)), "pool info")
        return i

    def step_latencies(self, n):
        ms = np.zeros(int(n), np.float32)
        L.check(L.lib().cf_pool_step_latencies(self._h, int(n), ms.ctypes.data), "pool latencies")
        return ms
