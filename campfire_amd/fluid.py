"""FluidSim — Python mirror of campfire's `FluidSim` (js/fluid.js:26-356)
over the C ABI of libcampfire_hip.so.

Method names follow the reference (camel case kept as snake case), and so do
argument meaning and error behaviour: construction with an unsupported grid
raises 'coarsest level too large' like js/fluid.js:91; the encode* methods
take an ignored pass/encoder argument so call sites read like
js/main.js:272-283; readback goes through a 3-slot ring and is skipped when
the slot is busy (js/fluid.js:317).
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import (ACC_GLOBAL, ACC_LEN, E_SCALE, G_SCALE, M_SCALE, MAX_PIECES, N_FLOOR,  # noqa: F401
                   PARAM_NAMES, PIECE_FLOATS, RAD_LEN, CampfireError, check)

SCENE_WIDTH = 0.7  # js/config.js:3
STAGES = ["advect", "diffuse", "sources", "forces", "coef/rhs", "vcycle", "project"]


class FluidSim:
    """One gas scene on the GPU.

    mode: {"nx", "ny", "dt"} as js/config.js MODES; dx defaults to
    SCENE_WIDTH / nx as js/fluid.js:30 derives it.
    """

    def __init__(self, mode, dx=None, init=True):
        self.nx, self.ny = int(mode["nx"]), int(mode["ny"])
        self.dt = float(mode["dt"])
        self.dx = float(dx) if dx is not None else SCENE_WIDTH / self.nx
        self.height = self.dx * self.ny
        self._h = None
        self._owned = True
        self._pending = []
        if init:
            self.init()

    @classmethod
    def borrow(cls, handle, mode, dx):
        """A view of a scene owned elsewhere (a pool slot): closing it does not
        free the scene."""
        self = cls(mode, dx, init=False)
        self._h = handle
        self._owned = False
        return self

    # ---- lifetime ----------------------------------------------------------
    def init(self):
        h = C.c_void_p()
        rc = L.lib().cf_sim_create(self.nx, self.ny, self.dx, self.dt, C.byref(h))
        if rc == L.CF_EBADSIZE:
            raise CampfireError(rc, "coarsest level too large")
        check(rc, "init")
        self._h = h
        return self

    def close(self):
        if self._h and self._owned:
            L.lib().cf_sim_destroy(self._h)
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

    # ---- state ---------------------------------------------------------------
    def info(self):
        i = L.CfSimInfo()
        check(L.lib().cf_sim_get_info(self._h, C.byref(i)), "info")
        return i

    @property
    def time(self):
        return self.info().time

    @time.setter
    def time(self, v):
        self.set_param("time", v)

    @property
    def frame(self):
        return int(self.info().frame)

    @frame.setter
    def frame(self, v):
        self.set_param("frame", v)

    @property
    def diff_sub(self):
        return self.info().diff_sub

    @property
    def levels(self):
        i = self.info()
        return [(i.level_nx[l], i.level_ny[l]) for l in range(i.nlevels)]

    @property
    def params(self):
        out = {}
        v = C.c_double()
        for k in PARAM_NAMES:
            check(L.lib().cf_sim_get_param(self._h, k.encode(), C.byref(v)), k)
            out[k] = v.value
        return out

    def packed_params(self):
        P = L.CfParams()
        check(L.lib().cf_sim_get_packed_params(self._h, C.byref(P)))
        return P

    # ---- FluidSim methods -----------------------------------------------------
    def reset(self):  # js/fluid.js:162-176
        check(L.lib().cf_sim_reset(self._h), "reset")

    def set_param(self, k, v):  # js/fluid.js:178
        check(L.lib().cf_sim_set_param(self._h, k.encode(), float(v)), "setParam " + k)

    setParam = set_param

    def write_params(self):  # js/fluid.js:180-199
        check(L.lib().cf_sim_write_params(self._h), "writeParams")

    def set_schedule(self, nu=2, vcycles=1, tol=0.0, coarse_iters=120, smoother_fp16=0, cfl=0.0, umin=1.0, refine=0,
                     krylov=0):
        """Multigrid schedule (reference: nu 2, 1 V-cycle, 120 coarsest sweeps).
        tol > 0 runs up to `vcycles` V-cycles until ||r||/||rhs|| < tol; cfl > 0
        takes dt = cfl*dx/max(max|u|, umin) each step, on the device; refine = 1
        keeps the level-0 pressure and residual in FP64 (defect correction);
        smoother_fp16 = 1 (with refine) stores p and the face coefficients of
        the tiled multigrid levels in FP16, arithmetic staying FP32; krylov = 1
        (with refine) runs flexible conjugate gradients at level 0, each
        V-cycle preconditioning one iteration."""
        s = L.CfSchedule(nu, vcycles, tol, coarse_iters, smoother_fp16, cfl, umin, refine, krylov)
        check(L.lib().cf_sim_set_schedule(self._h, C.byref(s)), "set_schedule")

    def get_schedule(self):
        """The schedule as a dict of set_schedule's keyword arguments."""
        s = L.CfSchedule()
        check(L.lib().cf_sim_get_schedule(self._h, C.byref(s)), "get_schedule")
        return {n: getattr(s, n) for n, _ in L.CfSchedule._fields_}

    def set_leg_tiles(self, tiles):
        """Tile code (16, 32 or 64: square tiles of that edge; 48: the wide tiles,
        64 cells wide with 48 or 32 rows; 0 = default rule) of the tiled multigrid legs
        of each level from level 0; bitwise the same results for any choice."""
        a = np.ascontiguousarray(tiles, dtype=np.int32)
        check(L.lib().cf_sim_set_leg_tiles(self._h, a.ctypes.data, int(a.size)), "set_leg_tiles")

    def leg_tiles(self):
        """The tile codes in use, one per tiled level; 0 for a level whose
        legs run in a two-level pair (their tile is not set per level)."""
        a = np.zeros(16, np.int32)
        n = check(L.lib().cf_sim_get_leg_tiles(self._h, a.ctypes.data, 16), "leg_tiles")
        return [int(x) for x in a[:n]]

    def upload_pieces(self, data, n):  # js/fluid.js:201-204
        a = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
        if a.size != MAX_PIECES * PIECE_FLOATS:
            raise ValueError("pieces must be Float32Array(64*20)")
        check(L.lib().cf_sim_upload_pieces(self._h, a.ctypes.data, int(n)), "uploadPieces")

    def clear_geometry(self, encoder=None):  # js/fluid.js:213
        check(L.lib().cf_sim_clear_geometry(self._h), "clearGeometry")

    def encode_geometry(self, pass_=None):  # js/fluid.js:215-218
        check(L.lib().cf_sim_geometry(self._h), "encodeGeometry")

    def encode_radiation(self, pass_=None):  # js/fluid.js:220-222
        check(L.lib().cf_sim_radiation(self._h), "encodeRadiation")

    def encode_step(self, pass_=None, n=1):  # js/fluid.js:225-253
        check(L.lib().cf_sim_step(self._h, int(n)), "encodeStep")

    def step(self, n=1):
        self.encode_step(None, n)

    def run_stage(self, stage):
        check(L.lib().cf_sim_run_stage(self._h, int(stage)), "stage %d" % stage)

    def encode_step_profiled(self, n=1):  # js/fluid.js:256-292
        ms = np.zeros(7, np.float32)
        check(L.lib().cf_sim_profile_steps(self._h, int(n), ms.ctypes.data), "profile")
        return dict(zip(STAGES, (ms / n).tolist()))

    def profile_kernels(self, n=1):
        """{kernel: (total_ms, launches)} over n production steps."""
        mx = 64
        names = C.create_string_buffer(32 * mx)
        ms = np.zeros(mx, np.float32)
        la = np.zeros(mx, np.int32)
        cnt = C.c_int(0)
        check(L.lib().cf_sim_profile_kernels(self._h, int(n), mx, names, ms.ctypes.data, la.ctypes.data,
                                              C.byref(cnt)), "profile_kernels")
        raw = names.raw
        out = {}
        for q in range(cnt.value):
            nm = raw[32 * q:32 * (q + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[q]), int(la[q]))
        return out

    def step_latencies(self, n):
        """Device time of each of n steps (ms), events around every step."""
        ms = np.zeros(int(n), np.float32)
        check(L.lib().cf_sim_step_latencies(self._h, int(n), ms.ctypes.data), "step_latencies")
        return ms

    def synchronize(self):
        check(L.lib().cf_sim_synchronize(self._h), "synchronize")

    def profile_frame(self, reps=10):
        """Mean device milliseconds of the per-frame geometry and radiation."""
        g, r = C.c_float(0), C.c_float(0)
        check(L.lib().cf_sim_profile_frame(self._h, int(reps), C.byref(g), C.byref(r)), "profile_frame")
        return float(g.value), float(r.value)

    def status(self):
        """Waits for the enqueued steps; raises CampfireError with CF_ENONFINITE
        (a NaN/Inf reached the fields; the flag is then cleared) or CF_ENOCONV
        (a residual target was missed by the last step)."""
        check(L.lib().cf_sim_status(self._h), "status")

    def frame_all(self, nsteps, with_radiation=True):
        """One app frame as js/main.js:266-283 encodes it."""
        check(L.lib().cf_sim_frame(self._h, int(nsteps), int(bool(with_radiation))), "frame")

    # ---- readback (js/fluid.js:315-342) -----------------------------------------
    def encode_readback(self, encoder=None):
        rc = L.lib().cf_sim_encode_readback(self._h)
        if rc == L.CF_EAGAIN:
            self._pending_stage = -1
            return -1
        check(rc, "encodeReadback")
        self._pending_stage = rc
        return rc

    def after_submit(self, cb, tag=None):
        """Queue cb({acc, rad, faceW}, tag) for when the pending readback completes;
        callbacks run from poll() (or flush()), like a resolved mapAsync."""
        i = getattr(self, "_pending_stage", -1)
        if i < 0:
            return
        self._pending.append((i, cb, tag))
        self._pending_stage = -1

    def _take(self, slot, wait):
        acc = np.zeros(ACC_LEN, np.int32)
        rad = np.zeros(RAD_LEN * 4, np.float32)
        fw = np.zeros(MAX_PIECES * 4, np.int32)
        rc = L.lib().cf_sim_take_readback(self._h, slot, int(wait), acc.ctypes.data, rad.ctypes.data,
                                          fw.ctypes.data)
        if rc == L.CF_ENOTREADY:
            return None
        check(rc, "readback")
        return {"acc": acc, "rad": rad, "faceW": fw}

    def poll(self, wait=False):
        keep = []
        for slot, cb, tag in self._pending:
            r = self._take(slot, wait)
            if r is None:
                keep.append((slot, cb, tag))
            elif cb is not None:
                cb(r, tag)
        self._pending = keep

    def flush(self):
        self.poll(wait=True)

    def readback_now(self):
        """encodeReadback + wait: {acc, rad, faceW}; acc is cleared on the device."""
        slot = L.lib().cf_sim_encode_readback(self._h)
        check(slot, "encodeReadback")
        return self._take(slot, True)

    # ---- buffers ------------------------------------------------------------------
    def field_bytes(self, name):
        n = C.c_size_t()
        check(L.lib().cf_sim_field_bytes(self._h, name.encode(), C.byref(n)), "field " + name)
        return n.value

    def read_field(self, name):
        """Typed copy of a device buffer (int32 for acc/faceW, float32 otherwise)."""
        nb = self.field_bytes(name)
        dt = np.int32 if name in ("acc", "faceW") else np.float32
        a = np.empty(nb // 4, dt)
        check(L.lib().cf_sim_read_field(self._h, name.encode(), a.ctypes.data, a.nbytes), "read " + name)
        return a

    # the buffers shaders/render.wgsl:20-29 binds that the simulation owns,
    # as (binding, buffer, WGSL element type, bytes per element); bindings 1
    # (the renderer's camera uniform) and 8 (the wood model's state) belong to
    # the app and are not dumped
    RENDER_BUFFERS = [(0, "P", "Params (uniform)", 160), (2, "T", "f32", 4), (3, "Y", "vec4f", 16),
                      (4, "SQ", "vec2f", 8), (5, "pieces", "Piece", 4 * PIECE_FLOATS), (6, "U", "f32", 4),
                      (7, "V", "f32", 4), (9, "rad", "vec4f", 16)]

    def dump_render_buffers(self, path):
        """Headless field dump for an offline renderer: every buffer that
        shaders/render.wgsl:20-29 binds from the simulation, written as raw
        little-endian bytes in the layout the shader reads (js/render.js:37
        binds the simulation's own buffers), to <path>/<name>.bin, with
        <path>/manifest.json naming the binding, element type, element count
        and the grid.  Returns the manifest."""
        import json
        import os
        os.makedirs(path, exist_ok=True)
        info = self.info()
        man = {"nx": self.nx, "ny": self.ny, "dx": float(self.dx), "time": float(info.time),
               "frame": int(info.frame), "buffers": []}
        for binding, name, wtype, esize in self.RENDER_BUFFERS:
            data = bytes(self.packed_params()) if name == "P" else self.read_field(name).astype("<f4").tobytes()
            with open(os.path.join(path, name + ".bin"), "wb") as fh:
                fh.write(data)
            man["buffers"].append({"binding": binding, "name": name, "type": wtype, "bytes": len(data),
                                   "elements": len(data) // esize})
        with open(os.path.join(path, "manifest.json"), "w") as fh:
            json.dump(man, fh, indent=1)
        return man

    def read_buffer(self, name):  # js/fluid.js:345-355: the raw buffer as Float32Array
        return self.read_field(name).view(np.float32)

    readBuffer = read_buffer

    def read_p64(self):
        """The FP64 level-0 pressure of defect correction (refine = 1)."""
        a = np.empty(self.nx * self.ny, np.float64)
        check(L.lib().cf_sim_read_field(self._h, b"p64", a.ctypes.data, a.nbytes), "read p64")
        return a

    def read_direction(self):
        """The stored FP32 direction of the Krylov iteration (krylov = 1)."""
        a = np.empty(self.nx * self.ny, np.float32)
        check(L.lib().cf_sim_read_field(self._h, b"d32", a.ctypes.data, a.nbytes), "read d32")
        return a

    def write_field(self, name, arr):
        nb = self.field_bytes(name)
        dt = np.int32 if name in ("acc", "faceW") else np.float32
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        if a.nbytes != nb:
            raise ValueError("%s: %d bytes expected, got %d" % (name, nb, a.nbytes))
        check(L.lib().cf_sim_write_field(self._h, name.encode(), a.ctypes.data, a.nbytes), "write " + name)

    def copy_state_from(self, other):
        """Every device buffer and parameter of `other` (same grid) into this scene."""
        check(L.lib().cf_sim_copy_state(self._h, other._h), "copy_state")

    def compat_sums(self):
        """(net source sum S dx^2, net open-boundary outflow), FP64 sums on the device."""
        a, b = C.c_double(), C.c_double()
        check(L.lib().cf_sim_compat_sums(self._h, C.byref(a), C.byref(b)), "compat_sums")
        return a.value, b.value

    def div_error(self):
        """(max, mean) of |div(f u) - S| (H.divErr, dev/gas_test.html:101-111)."""
        mx, mean = C.c_float(), C.c_float()
        check(L.lib().cf_sim_div_error(self._h, C.byref(mx), C.byref(mean)), "divErr")
        return mx.value, mean.value
