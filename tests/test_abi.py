"""The C-ABI library loads, exports every entry point include/campfire_hip.h
declares, and its host-only paths behave (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from campfire_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "campfire_hip.h")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cf_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_param_table_matches_reference_order():
    # PARAM_FIELDS (js/fluid.js:13-24)
    L = _lib.lib()
    got = [L.cf_param_name(q).decode() for q in range(40)]
    assert got == _lib.PARAM_NAMES
    assert L.cf_param_name(40) is None
    assert C.sizeof(_lib.CfParams) == 160


def test_version_string():
    assert b"gfx950" in _lib.lib().cf_version()


def test_bad_grid_is_coarsest_level_too_large():
    # 34x34 halves to an odd 17x17 = 289 > 256 cells (js/fluid.js:88-91)
    L = _lib.lib()
    h = C.c_void_p()
    rc = L.cf_sim_create(34, 34, 0.01, 0.002, C.byref(h))
    assert rc == _lib.CF_EBADSIZE
    assert b"coarsest level too large" in L.cf_last_error()
    assert not h.value


def test_bad_arguments_are_rejected_without_a_device():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.cf_sim_create(1, 64, 0.01, 0.002, C.byref(h)) == _lib.CF_EINVAL
    assert L.cf_sim_create(64, 64, -1.0, 0.002, C.byref(h)) == _lib.CF_EINVAL
    assert L.cf_sim_create(64, 64, 0.01, 0.002, None) == _lib.CF_EINVAL
    assert L.cf_sim_reset(None) == _lib.CF_EINVAL
    assert L.cf_sim_step(None, 1) == _lib.CF_EINVAL


def test_fluidsim_raises_reference_message():
    from campfire_amd.fluid import FluidSim
    with pytest.raises(_lib.CampfireError, match="coarsest level too large"):
        FluidSim({"nx": 34, "ny": 34, "dt": 0.002})


def test_reference_mode_hierarchies():
    # js/fluid.js:71-91 on the js/config.js:6-10 modes: halve while even and > 256 cells
    from oracle.pyoracle import Oracle
    for (nx, ny), nlev in (((160, 224), 5), ((224, 320), 6), ((352, 512), 6)):
        o = Oracle(nx, ny, 0.7 / nx, 0.003)
        assert o.nlevels == nlev, (nx, ny, o.nlevels)
        o.close()
    # the reference's 6-level cap is lifted: 1024^2 reaches 16x16 in 7 levels
    o = Oracle(1024, 1024, 0.00125, 0.00025)
    assert o.nlevels == 7


def test_flame1d_rejects_bad_sizes_without_device_work():
    import ctypes as C
    from campfire_amd import _lib as L
    c = (L.CfFlameCase * 1)()
    c[0].B, c[0].Ta, c[0].dx, c[0].F, c[0].phi, c[0].L, c[0].tEnd = 2e10, 15000, 1e-5, 1, 1, 0.03, 1
    out = (L.CfFlameResult * 1)()
    assert L.lib().cf_flame1d_run(c, 1, 1, out) == L.CF_EBADSIZE   # 3000 cells > CF_FLAME_MAXN
    c[0].dx = 0.0
    assert L.lib().cf_flame1d_run(c, 1, 1, out) == L.CF_EINVAL
