"""cf_flame1d_run_static, the allocation-free variant of cf_flame1d_run: its
staging lives in a static device array, so two calls in a row (the second with
fewer cases, in the other precision) must each give exactly what
cf_flame1d_run gives for the same cases, and sizes beyond the array fail."""
import ctypes as C

import pytest

from campfire_amd import _lib as L

pytestmark = pytest.mark.gpu


def _cases(spec):
    arr = (L.CfFlameCase * len(spec))()
    for q, (B, dx, F, tEnd, cap) in enumerate(spec):
        arr[q] = L.CfFlameCase(B, 15000.0, dx, F, 1.0, 0.12, tEnd, 293.15, 2000.0, cap, 0)
    return arr


def _run(fn, arr, n, fp64, *extra):
    out = (L.CfFlameResult * n)()
    rc = fn(arr, n, fp64, out, *extra)
    return rc, [(r.SL, r.Tb, r.steps, r.n) for r in out]


def test_static_variant_equals_the_allocating_one_on_two_calls():
    lib = L.lib()
    first = _cases([(2e10 * (1 + 0.1 * q), 2e-4 * (1 + 0.25 * (q % 4)), 1.0 + 0.5 * (q % 3), 0.004 + 0.001 * q, 0)
                    for q in range(12)])
    second = _cases([(1e9, 3e-4, 2.0, 0.02, 3000), (2e10, 1e-3, 1.0, 0.05, 0), (5e9, 5e-4, 1.5, 0.01, 500)])
    print("[no-alloc] call 1: 12 cases FP64; call 2: 3 cases FP32", flush=True)
    for arr, n, fp64 in ((first, 12, 1), (second, 3, 0)):
        rc_s, got = _run(lib.cf_flame1d_run_static, arr, n, fp64, None)
        assert rc_s == 0, lib.cf_last_error()
        rc_a, ref = _run(lib.cf_flame1d_run, arr, n, fp64)
        assert rc_a == 0, lib.cf_last_error()
        assert got == ref
        assert all(r[2] > 0 and r[3] > 0 for r in got)
        print("[no-alloc]   %d cases equal, steps %s" % (n, [r[2] for r in got]), flush=True)


def test_static_variant_rejects_sizes_beyond_its_array():
    lib = L.lib()
    n = 257  # CF_FLAME_STATIC_MAX + 1
    arr = _cases([(2e10, 1e-3, 1.0, 0.001, 10)] * n)
    out = (L.CfFlameResult * n)()
    assert lib.cf_flame1d_run_static(arr, n, 1, out, None) == L.CF_EINVAL
    assert b"CF_FLAME_STATIC_MAX" in lib.cf_last_error()
    assert lib.cf_flame1d_run_static(arr, 0, 1, out, None) == L.CF_EINVAL
    bad = _cases([(2e10, 1e-5, 1.0, 0.001, 10)])  # 12000 cells > CF_FLAME_MAXN
    assert lib.cf_flame1d_run_static(bad, 1, 1, out, None) == L.CF_EBADSIZE
