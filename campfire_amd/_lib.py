"""ctypes binding of libcampfire_hip.so (the C ABI of include/campfire_hip.h).

There is no fallback: if the HIP library has not been built, importing the
binding raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcampfire_hip.so")

CF_OK, CF_EINVAL, CF_EBADSIZE, CF_ENOMEM, CF_EDEVICE = 0, -1, -2, -3, -4
CF_EAGAIN, CF_ENOTREADY, CF_ENOCONV, CF_ENONFINITE = -5, -6, -7, -8
ERR_NAMES = {CF_EINVAL: "CF_EINVAL", CF_EBADSIZE: "CF_EBADSIZE", CF_ENOMEM: "CF_ENOMEM",
             CF_EDEVICE: "CF_EDEVICE", CF_EAGAIN: "CF_EAGAIN", CF_ENOTREADY: "CF_ENOTREADY",
             CF_ENOCONV: "CF_ENOCONV", CF_ENONFINITE: "CF_ENONFINITE"}

MAX_PIECES, PIECE_FLOATS, ACC_GLOBAL, ACC_LEN = 64, 20, 512, 528
E_SCALE, M_SCALE, G_SCALE, N_FLOOR, RAD_LEN = 1e6, 1e12, 1e3, 64, 320

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


class CfSimInfo(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dt", C.c_double),
                ("height", C.c_double), ("time", C.c_double), ("frame", C.c_uint32),
                ("diff_sub", C.c_int32), ("nlevels", C.c_int32),
                ("level_nx", C.c_int32 * 16), ("level_ny", C.c_int32 * 16),
                ("last_vcycles", C.c_int32), ("last_resid", C.c_float)]


class CfFlameCase(C.Structure):
    _fields_ = [("B", C.c_double), ("Ta", C.c_double), ("dx", C.c_double), ("F", C.c_double),
                ("phi", C.c_double), ("L", C.c_double), ("tEnd", C.c_double),
                ("Tamb", C.c_double), ("Tign", C.c_double), ("max_steps", C.c_int32), ("pad_", C.c_int32)]


class CfFlameResult(C.Structure):
    _fields_ = [("SL", C.c_double), ("Tb", C.c_double), ("steps", C.c_int32), ("n", C.c_int32)]


class CfSceneDesc(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("dx", C.c_double), ("dt", C.c_double)]


class CfPoolInfo(C.Structure):
    _fields_ = [("nscenes", C.c_int32), ("nslots", C.c_int32), ("budget", C.c_uint64), ("used", C.c_uint64),
                ("cells", C.c_uint64), ("steps", C.c_int64), ("launches_per_step", C.c_int32), ("pad_", C.c_int32)]


class CampfireError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%s)" % (msg, ERR_NAMES.get(code, code)))
        self.code = code


_lib = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libcampfire_hip.so is not built (run __graft_entry__.build() or "
                          "python -m campfire_amd.build)")
    L = C.CDLL(LIB_PATH)
    p, i, d, f, vp, cp, sz = C.c_void_p, C.c_int, C.c_double, C.c_float, C.c_void_p, C.c_char_p, C.c_size_t
    sig = {
        "cf_sim_create": ([i, i, d, d, C.POINTER(p)], i),
        "cf_sim_destroy": ([p], None),
        "cf_sim_reset": ([p], i),
        "cf_sim_set_param": ([p, cp, d], i),
        "cf_sim_get_param": ([p, cp, C.POINTER(d)], i),
        "cf_sim_write_params": ([p], i),
        "cf_sim_get_packed_params": ([p, C.POINTER(CfParams)], i),
        "cf_param_name": ([i], cp),
        "cf_sim_set_schedule": ([p, C.POINTER(CfSchedule)], i),
        "cf_sim_get_schedule": ([p, C.POINTER(CfSchedule)], i),
        "cf_sim_set_leg_tiles": ([p, vp, i], i),
        "cf_sim_get_leg_tiles": ([p, vp, i], i),
        "cf_sim_upload_pieces": ([p, vp, i], i),
        "cf_sim_clear_geometry": ([p], i),
        "cf_sim_geometry": ([p], i),
        "cf_sim_step": ([p, i], i),
        "cf_sim_radiation": ([p], i),
        "cf_sim_frame": ([p, i, i], i),
        "cf_sim_run_stage": ([p, i], i),
        "cf_sim_profile_steps": ([p, i, vp], i),
        "cf_sim_profile_kernels": ([p, i, i, vp, vp, vp, C.POINTER(i)], i),
        "cf_sim_step_latencies": ([p, i, vp], i),
        "cf_sim_profile_frame": ([p, i, C.POINTER(f), C.POINTER(f)], i),
        "cf_sim_synchronize": ([p], i),
        "cf_sim_status": ([p], i),
        "cf_sim_encode_readback": ([p], i),
        "cf_sim_take_readback": ([p, i, i, vp, vp, vp], i),
        "cf_sim_field_bytes": ([p, cp, C.POINTER(sz)], i),
        "cf_sim_read_field": ([p, cp, vp, sz], i),
        "cf_sim_write_field": ([p, cp, vp, sz], i),
        "cf_sim_get_info": ([p, C.POINTER(CfSimInfo)], i),
        "cf_sim_div_error": ([p, C.POINTER(f), C.POINTER(f)], i),
        "cf_sim_copy_state": ([p, p], i),
        "cf_sim_compat_sums": ([p, C.POINTER(d), C.POINTER(d)], i),
        "cf_flame1d_run": ([C.POINTER(CfFlameCase), i, i, C.POINTER(CfFlameResult)], i),
        "cf_flame1d_run_static": ([C.POINTER(CfFlameCase), i, i, C.POINTER(CfFlameResult), vp], i),
        "cf_scene_bytes": ([i, i, C.c_double], C.c_uint64),
        "cf_copy_bandwidth": ([C.c_uint64, i, C.POINTER(d), C.POINTER(d)], i),
        "cf_pool_create": ([C.c_uint64, C.POINTER(p)], i),
        "cf_pool_destroy": ([p], None),
        "cf_pool_set_schedule": ([p, C.POINTER(CfSchedule)], i),
        "cf_pool_set_leg_tiles": ([p, vp, i], i),
        "cf_pool_get_leg_tiles": ([p, vp, i], i),
        "cf_pool_admit": ([p, C.POINTER(CfSceneDesc), C.POINTER(i)], i),
        "cf_pool_scene": ([p, i, C.POINTER(p)], i),
        "cf_pool_release": ([p, i], i),
        "cf_pool_prepare": ([p], i),
        "cf_pool_step": ([p, i], i),
        "cf_pool_frame": ([p, i, i], i),
        "cf_pool_synchronize": ([p], i),
        "cf_pool_get_info": ([p, C.POINTER(CfPoolInfo)], i),
        "cf_pool_status": ([p, C.POINTER(i)], i),
        "cf_pool_step_latencies": ([p, i, vp], i),
        "cf_last_error": ([], cp),
        "cf_version": ([], cp),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc, what=""):
    """Raise CampfireError for a negative status; return rc otherwise."""
    if rc is not None and rc < 0:
        msg = lib().cf_last_error().decode(errors="replace")
        raise CampfireError(rc, (what + ": " if what else "") + msg)
    return rc
