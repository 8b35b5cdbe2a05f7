"""tests/flame1d.mjs flameSpeed() (:13-71) on the GPU.

    from campfire_amd.flame1d import flame_speed, flame_speeds
    flame_speed(B=2e10, Ta=15000, dx=5e-5, L=0.03, tEnd=3)   # {'SL': ..., 'Tb': ..., 'steps': ..., 'n': ...}

flame_speeds() runs many cases in one launch, one workgroup each.  fp64=True
is the script's double arithmetic; fp64=False the single-precision form.
"""
import ctypes as C

from . import _lib as L

SCRIPT_TAMB = 293.15   # tests/flame1d.mjs:6
SCRIPT_TIGN = 2000.0   # tests/flame1d.mjs:21


def flame_speeds(cases, fp64=True):
    """cases: iterable of dicts with the flameSpeed arguments (B, Ta, dx, and
    optional F = 1, phi = 1, L = 0.12, tEnd = 0.25) plus optional Tamb, Tign
    and max_steps.  Returns one result dict per case."""
    cases = list(cases)
    arr = (L.CfFlameCase * len(cases))()
    for c, d in zip(arr, cases):
        c.B, c.Ta, c.dx = d["B"], d["Ta"], d["dx"]
        c.F, c.phi, c.L, c.tEnd = d.get("F", 1.0), d.get("phi", 1.0), d.get("L", 0.12), d.get("tEnd", 0.25)
        c.Tamb, c.Tign = d.get("Tamb", SCRIPT_TAMB), d.get("Tign", SCRIPT_TIGN)
        c.max_steps = int(d.get("max_steps", 0))
    out = (L.CfFlameResult * len(cases))()
    L.check(L.lib().cf_flame1d_run(arr, len(cases), int(bool(fp64)), out), "flame1d")
    return [{"SL": r.SL, "Tb": r.Tb, "steps": r.steps, "n": r.n} for r in out]


def flame_speed(B, Ta, dx, F=1.0, phi=1.0, L=0.12, tEnd=0.25, fp64=True, **kw):
    return flame_speeds([dict(B=B, Ta=Ta, dx=dx, F=F, phi=phi, L=L, tEnd=tEnd, **kw)], fp64)[0]
