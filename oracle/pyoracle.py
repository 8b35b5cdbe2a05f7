"""ctypes loader of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Imported by tests/, the smoke check and the benchmark's CPU-baseline leg, as
the checker.  `Oracle(nx, ny, dx, dt, fp64=False)` mirrors the cf_sim C ABI of
the product library (same field names, layouts, stage numbering), so a test
can run both on identical inputs and compare.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PARAM_NAMES = [
    "nx", "ny", "dx", "dt", "Tamb", "rhoAmb", "YOamb", "grav",
    "cp", "Pr", "viscMult", "diffMult", "reactB", "reactTa", "dHc", "sO2",
    "sootFormA", "sootFormTa", "sootOxA", "sootOxTa",
    "kappaSootC", "kappaGas", "exchAlpha", "depth",
    "charA", "charTa", "nPieces", "frame",
    "heaterX0", "heaterX1", "heaterY1", "heaterT",
    "burnerX0", "burnerX1", "burnerFlux", "burnerT",
    "sootDensity", "sigma", "charO2", "time"]


class CfParams(C.Structure):
    _fields_ = [(n, C.c_int32 if n in ("nx", "ny", "nPieces") else C.c_uint32 if n == "frame" else C.c_float)
                for n in PARAM_NAMES]


class CfSchedule(C.Structure):
    _fields_ = [("nu", C.c_int32), ("vcycles", C.c_int32), ("tol", C.c_float),
                ("coarse_iters", C.c_int32), ("smoother_fp16", C.c_int32), ("cfl", C.c_float), ("umin", C.c_float), ("refine", C.c_int32), ("krylov", C.c_int32)]


class FlameResult(C.Structure):
    _fields_ = [("SL", C.c_double), ("Tb", C.c_double), ("steps", C.c_int)]


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_LIBS = {}


def lib(path=None):
    """The oracle library (liboracle.so beside this file), or the build at
    `path` (a mutant of the oracle, built by the mutation tests)."""
    global _LIB
    if path is not None:
        if path not in _LIBS:
            _LIBS[path] = _bind(C.CDLL(path))
        return _LIBS[path]
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(C.CDLL(path))
    return _LIB


def _bind(L):
    if True:
        for p in ("cfo32_", "cfo64_"):
            f = getattr(L, p + "create"); f.restype = C.c_void_p
            f.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int)]
            getattr(L, p + "destroy").argtypes = [C.c_void_p]
            getattr(L, p + "set_param").argtypes = [C.c_void_p, C.c_char_p, C.c_double]
            for n in ("write_params", "reset", "diff_sub", "nlevels", "last_vcycles"):
                getattr(L, p + n).argtypes = [C.c_void_p]
            for n in ("time", "last_resid"):
                getattr(L, p + n).argtypes = [C.c_void_p]
                getattr(L, p + n).restype = C.c_double
            getattr(L, p + "get_packed").argtypes = [C.c_void_p, C.POINTER(CfParams)]
            getattr(L, p + "set_schedule").argtypes = [C.c_void_p, C.POINTER(CfSchedule)]
            getattr(L, p + "upload_pieces").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
            for n in ("clear_geometry", "geometry", "radiation"):
                getattr(L, p + n).argtypes = [C.c_void_p]
                getattr(L, p + n).restype = None
            getattr(L, p + "run_stage").argtypes = [C.c_void_p, C.c_int]
            getattr(L, p + "run_kernel").argtypes = [C.c_void_p, C.c_char_p, C.c_int]
            getattr(L, p + "step").argtypes = [C.c_void_p, C.c_int]
            getattr(L, p + "readback").argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            for n in ("read_field", "write_field"):
                getattr(L, p + n).argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
            for n in ("read_field64", "write_field64"):
                getattr(L, p + n).argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_size_t]
            getattr(L, p + "closure").argtypes = [C.c_void_p, C.c_char_p, C.c_double]
            getattr(L, p + "closure").restype = C.c_double
        L.cfo_f2h.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.cfo_gpu_sum.argtypes = [C.c_void_p, C.c_int]
        L.cfo_gpu_sum.restype = C.c_double
        L.cfo_f2h.restype = None
        L.cfo_h2f.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.cfo_h2f.restype = None
        L.cfo_flame_speed.restype = FlameResult
        L.cfo_flame_speed.argtypes = [C.c_double] * 7
        for n in ("cfo_flame_speed_ex", "cfo_flame_speed32_ex"):
            getattr(L, n).restype = FlameResult
            getattr(L, n).argtypes = [C.c_double] * 9
    return L


def f2h(a):
    """float32 -> IEEE binary16 bits (uint16) by the oracle's conversion."""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    h = np.zeros(a.size, np.uint16)
    lib().cfo_f2h(a.ctypes.data, h.ctypes.data, a.size)
    return h


def h2f(h):
    """IEEE binary16 bits (uint16) -> float32 by the oracle's conversion."""
    h = np.ascontiguousarray(h, dtype=np.uint16).reshape(-1)
    a = np.zeros(h.size, np.float32)
    lib().cfo_h2f(h.ctypes.data, a.ctypes.data, h.size)
    return a


def gpu_sum(v):
    """Sum of a float64 array in the GPU's fixed reduction order."""
    a = np.ascontiguousarray(v, dtype=np.float64)
    return float(lib().cfo_gpu_sum(a.ctypes.data, a.size))


def flame_speed(B, Ta, dx, F=1.0, phi=1.0, L=0.12, tEnd=0.25, Tamb=None, Tign=None, fp32=False):
    """tests/flame1d.mjs flameSpeed() restated in C (oracle/flame1d.c).  Tamb
    and Tign default to the script's 293.15 K and 2000 K; fp32 runs the
    single-precision build of the same algorithm."""
    if Tamb is None and Tign is None and not fp32:
        r = lib().cfo_flame_speed(B, Ta, dx, F, phi, L, tEnd)
    else:
        fn = lib().cfo_flame_speed32_ex if fp32 else lib().cfo_flame_speed_ex
        r = fn(B, Ta, dx, F, phi, L, tEnd, 293.15 if Tamb is None else Tamb, 2000.0 if Tign is None else Tign)
    return {"SL": r.SL, "Tb": r.Tb, "steps": r.steps}


class OracleError(RuntimeError):
    pass


class Oracle:
    """One scene in the CPU oracle (FP32 by default, FP64 with fp64=True)."""

    def __init__(self, nx, ny, dx, dt, fp64=False, lib_path=None):
        self.L = lib(lib_path)
        self.p = "cfo64_" if fp64 else "cfo32_"
        self.fp64 = fp64
        self.nx, self.ny, self.dx, self.dt = nx, ny, dx, dt
        err = C.c_int(0)
        self.h = self._f("create")(nx, ny, dx, dt, C.byref(err))
        if not self.h:
            raise OracleError("create failed: %d" % err.value)

    def _f(self, n):
        return getattr(self.L, self.p + n)

    def _chk(self, rc, what):
        if rc not in (0, None):
            raise OracleError("%s -> %d" % (what, rc))
        return rc

    def close(self):
        if self.h:
            self._f("destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def closure(self, which, x):
        """rhoOf / muOf / rhoDOf / invRho / planck580 evaluated by the oracle."""
        return self._f("closure")(self.h, which.encode(), float(x))

    # FluidSim surface -------------------------------------------------------
    def set_param(self, k, v):
        self._chk(self._f("set_param")(self.h, k.encode(), float(v)), "set_param " + k)

    def write_params(self):
        self._chk(self._f("write_params")(self.h), "write_params")

    def packed_params(self):
        P = CfParams()
        self._f("get_packed")(self.h, C.byref(P))
        return P

    @property
    def diff_sub(self):
        return self._f("diff_sub")(self.h)

    @property
    def nlevels(self):
        return self._f("nlevels")(self.h)

    @property
    def last_vcycles(self):
        return self._f("last_vcycles")(self.h)

    @property
    def last_resid(self):
        return self._f("last_resid")(self.h)

    @property
    def time(self):
        return self._f("time")(self.h)

    @property
    def dt_now(self):
        """dt of the parameter block (the CFL step writes it each step)."""
        return float(self.packed_params().dt)

    def set_schedule(self, nu=2, vcycles=1, tol=0.0, coarse_iters=120, smoother_fp16=0, cfl=0.0, umin=1.0, refine=0,
                     krylov=0):
        s = CfSchedule(nu, vcycles, tol, coarse_iters, smoother_fp16, cfl, umin, refine, krylov)
        self._chk(self._f("set_schedule")(self.h, C.byref(s)), "set_schedule")

    def reset(self):
        self._chk(self._f("reset")(self.h), "reset")

    def upload_pieces(self, pieces, n):
        a = np.ascontiguousarray(pieces, dtype=np.float32).reshape(-1)
        assert a.size == 64 * 20
        self._chk(self._f("upload_pieces")(self.h, a.ctypes.data, n), "upload_pieces")

    def clear_geometry(self):
        self._f("clear_geometry")(self.h)

    def geometry(self):
        self._f("geometry")(self.h)

    def radiation(self):
        self._f("radiation")(self.h)

    def run_stage(self, st):
        return self._f("run_stage")(self.h, st)

    def run_kernel(self, name, arg=0):
        self._chk(self._f("run_kernel")(self.h, name.encode(), arg), "run_kernel " + name)

    def step(self, n=1):
        self._chk(self._f("step")(self.h, n), "step")

    def readback(self):
        acc = np.zeros(528, np.int32)
        rad = np.zeros(320 * 4, np.float32)
        fw = np.zeros(256, np.int32)
        self._f("readback")(self.h, acc.ctypes.data, rad.ctypes.data, fw.ctypes.data)
        return acc, rad, fw

    def field_size(self, name):
        nx, ny = self.nx, self.ny
        nC, nU, nV = nx * ny, (nx + 1) * ny, nx * (ny + 1)
        tab = {"U": nU, "Uh": nU, "U2": nU, "V": nV, "Vh": nV, "V2": nV, "T": nC, "Th": nC, "T2": nC,
               "Y": 4 * nC, "Yh": 4 * nC, "Y2": 4 * nC, "SQ": 2 * nC, "cellG": 2 * nC, "fx": nU, "fy": nV,
               "acc": 528, "faceW": 256, "pieces": 1280, "rad": 1280, "p64": nC,
               "d32": nC}
        if name in tab:
            return tab[name]
        lx, ly = nx, ny
        for pre in ("rhs", "ax", "ay", "p"):
            if name.startswith(pre):
                rest = name[len(pre):]
                lev = int(rest) if rest else 0
                for _ in range(lev):
                    lx //= 2
                    ly //= 2
                return {"p": lx * ly, "rhs": lx * ly, "ax": (lx + 1) * ly, "ay": lx * (ly + 1)}[pre]
        raise KeyError(name)

    def read(self, name):
        n = self.field_size(name)
        dt = np.int32 if name in ("acc", "faceW") else np.float32
        a = np.zeros(n, dt)
        self._chk(self._f("read_field")(self.h, name.encode(), a.ctypes.data, a.nbytes), "read " + name)
        return a

    def read64(self, name):
        n = self.field_size(name) if name != "sdf" else self.nx * self.ny
        a = np.zeros(n, np.float64)
        self._chk(self._f("read_field64")(self.h, name.encode(), a.ctypes.data, n), "read64 " + name)
        return a

    def write(self, name, arr):
        n = self.field_size(name)
        dt = np.int32 if name in ("acc", "faceW") else np.float32
        a = np.ascontiguousarray(arr, dtype=dt).reshape(-1)
        assert a.size == n, (name, a.size, n)
        self._chk(self._f("write_field")(self.h, name.encode(), a.ctypes.data, a.nbytes), "write " + name)

    def write64(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        self._chk(self._f("write_field64")(self.h, name.encode(), a.ctypes.data, a.size), "write64 " + name)
