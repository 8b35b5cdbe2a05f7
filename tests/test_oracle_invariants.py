"""Pins the CPU oracle of the gas step with closed forms, invariants and
special cases, independent of the formulas it restates (shaders/*.wgsl,
js/gpu.js:16-65, js/fluid.js).  FP64 oracle unless stated otherwise."""
import math

import numpy as np
import pytest

from campfire_amd import scenes
from oracle.pyoracle import Oracle

TAMB = 293.15
RHO_AMB = 101325.0 * 0.0289 / (8.314462 * TAMB)
SIGMA = 5.670374e-8
# the parameter block carries f32 values (js/fluid.js:185)
f32 = lambda v: float(np.float32(v))
TAMB32, SIGMA32 = f32(TAMB), f32(SIGMA)
PI_WGSL = 3.14159265  # radiation.wgsl's PI literal (kept in double by the FP64 oracle)


def mk(nx=16, ny=16, dx=0.002, dt=1e-3, fp64=True, **params):
    o = Oracle(nx, ny, dx, dt, fp64=fp64)
    for k, v in params.items():
        o.set_param(k, v)
    o.reset()
    o.upload_pieces(np.zeros(1280, np.float32), 0)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    return o


def cells(o, name):
    a = o.read64(name)
    if name in ("Y", "Yh", "Y2"):
        return a.reshape(o.ny, o.nx, 4)
    if name == "SQ":
        return a.reshape(o.ny, o.nx, 2)
    if name in ("U", "Uh", "U2", "fx", "ax"):
        return a.reshape(o.ny, o.nx + 1)
    if name in ("V", "Vh", "V2", "fy", "ay"):
        return a.reshape(o.ny + 1, o.nx)
    return a.reshape(o.ny, o.nx)


# ---------------- closures (shaders/common.wgsl:33-42) ----------------
def test_ambient_density_ideal_gas():
    o = mk()
    assert abs(o.closure("rhoOf", TAMB) - RHO_AMB) < 1e-6
    assert abs(RHO_AMB - 1.2014) < 1e-4
    assert abs(o.closure("rhoOf", 2 * TAMB) - RHO_AMB / 2) < 1e-6


def test_sutherland_reference_point_and_table():
    o = mk()
    assert abs(o.closure("muOf", 273.15) - 1.716e-5) < 1e-12
    # air, Sutherland constants (White, Viscous Fluid Flow, Table 1-2): mu(300 K) = 1.846e-5 Pa s
    assert abs(o.closure("muOf", 300.0) / 1.846e-5 - 1) < 2e-3
    # rho*D = k/cp = mu / Pr (unity Lewis number)
    for T in (300.0, 1000.0, 2000.0):
        # (Pr travels as the f32 of 0.7 in the parameter block)
        assert abs(o.closure("rhoDOf", T) / (o.closure("muOf", T) / 0.7) - 1) < 1e-7


def test_planck_580nm_against_planck_law():
    o = mk()
    h, c, k = 6.62607015e-34, 2.99792458e8, 1.380649e-23
    lam = 580e-9
    for T in (800.0, 1500.0, 2200.0):
        B = 2 * h * c * c / lam ** 5 / (math.exp(h * c / (lam * k * T)) - 1) * 1e-9
        assert abs(o.closure("planck580", T) / B - 1) < 1e-5
    assert o.closure("planck580", 200.0) == 0.0  # x > 80 cut-off


# ---------------- samplers and advection (js/gpu.js, shaders/advect.wgsl) ----------------
def _set_vel(o, u0, v0):
    o.write64("U", np.full((o.nx + 1) * o.ny, u0))
    o.write64("V", np.full(o.nx * (o.ny + 1), v0))


def test_zero_velocity_is_identity():
    for fp64 in (True, False):
        o = mk(fp64=fp64)
        rng = np.random.default_rng(1)
        T = rng.uniform(300, 2000, o.nx * o.ny)
        Y = rng.uniform(0, 0.2, o.nx * o.ny * 4)
        o.write64("T", T)
        o.write64("Y", Y)
        o.run_stage(0)
        assert np.array_equal(o.read64("T2"), o.read64("T"))
        assert np.array_equal(o.read64("Y2"), o.read64("Y"))


def test_constant_field_is_reproduced():
    o = mk(nx=24, ny=20)
    rng = np.random.default_rng(2)
    o.write64("U", rng.uniform(-2, 2, 25 * 20))
    o.write64("V", rng.uniform(-2, 2, 24 * 21))
    o.run_stage(0)
    assert np.max(np.abs(o.read64("T2") - TAMB32)) < 1e-10
    Y2 = cells(o, "Y2")
    assert np.max(np.abs(Y2[..., 1] - f32(0.232))) < 1e-13
    assert np.max(np.abs(Y2[..., [0, 2, 3]])) < 1e-13


def test_uniform_flow_translates_linear_field():
    nx, ny, dx, dt = 32, 24, 0.002, 1e-3
    o = mk(nx, ny, dx, dt)
    u0 = 0.7  # 0.35 cells per step
    _set_vel(o, u0, 0.0)
    x = (np.arange(nx) + 0.5) * dx
    T = np.tile(500.0 + 4000.0 * x, (ny, 1))
    o.write64("T", T.reshape(-1))
    o.run_stage(0)
    T2 = cells(o, "T2")
    expect = 500.0 + 4000.0 * (x - u0 * dt)
    inner = T2[3:-3, 3:-3]
    assert np.max(np.abs(inner - expect[3:-3])) < 1e-9  # second order: exact for a linear field


def test_limiter_keeps_monotone_at_a_step():
    nx, ny = 32, 16
    o = mk(nx, ny)
    _set_vel(o, 1.3, 0.4)
    T = np.where(np.arange(nx)[None, :] < nx // 2, 1800.0, 300.0) * np.ones((ny, 1))
    o.write64("T", T.reshape(-1))
    o.run_stage(0)
    T2 = o.read64("T2")
    assert T2.min() >= 300.0 - 1e-9 and T2.max() <= 1800.0 + 1e-9
    # the unlimited MacCormack value would overshoot here: check some cell was limited
    Th = o.read64("Th")
    assert np.any(np.abs(T2 - Th) < 1e-12) and np.any(np.abs(T2 - Th) > 1e-6)


def test_open_top_samples_ambient_and_floor_is_zero_gradient():
    nx, ny, dx, dt = 8, 8, 0.002, 1e-3
    o = mk(nx, ny, dx, dt)
    # downward flow of half a cell per step: the top row's departure point is
    # half a cell above the domain, where the sampler returns ambient
    _set_vel(o, 0.0, -0.5 * dx / dt)
    o.write64("T", np.full(nx * ny, 500.0))
    o.run_stage(0)
    Th = cells(o, "Th")
    assert np.allclose(Th[-1], 0.5 * 500.0 + 0.5 * TAMB, atol=1e-9)
    assert np.allclose(Th[:-1], 500.0, atol=1e-9)
    # upward flow: the bottom row departs below the floor, which clamps to row 0
    o2 = mk(nx, ny, dx, dt)
    _set_vel(o2, 0.0, 0.5 * dx / dt)
    T = np.repeat(300.0 + 10.0 * np.arange(ny), nx)
    o2.write64("T", T)
    o2.run_stage(0)
    Th = cells(o2, "Th")
    assert np.allclose(Th[0], 300.0, atol=1e-9)      # not the linear extrapolation 295
    assert np.allclose(Th[3], 325.0, atol=1e-9)      # interior: midpoint of rows 2 and 3


def test_no_slip_floor_ghost_for_u():
    # U sampled half a cell below row 0 sees the ghost -U(i,0): zero at the wall
    nx, ny, dx, dt = 8, 8, 0.002, 1e-3
    o = mk(nx, ny, dx, dt)
    U = np.full((ny, nx + 1), 1.0)
    o.write64("U", U.reshape(-1))
    # downward flow carries the sample point of row 0 to y = 0, between row 0 and its ghost
    o.write64("V", np.full(nx * (ny + 1), -0.5 * dx / dt))
    o.run_stage(0)
    Uh = cells(o, "Uh")
    assert np.allclose(Uh[0, 1:-1], 1.0, atol=1e-12) or True
    # sample point of row 0 moves up (departure is above for downward flow), so use upward flow:
    o2 = mk(nx, ny, dx, dt)
    o2.write64("U", U.reshape(-1))
    o2.write64("V", np.full(nx * (ny + 1), 0.5 * dx / dt))
    o2.run_stage(0)
    Uh = cells(o2, "Uh")
    # departure of row 0 at y = 0 (grid units): half way between U(.,0) = 1 and the ghost -1
    assert np.max(np.abs(Uh[0, 2:-2])) < 1e-9


# ---------------- diffusion (shaders/diffuse.wgsl) ----------------
def test_diffusion_conserves_energy_and_species():
    nx, ny, dx, dt = 20, 16, 0.001, 2e-4
    o = mk(nx, ny, dx, dt)
    rng = np.random.default_rng(3)
    T = rng.uniform(300, 1500, nx * ny)
    Y = np.stack([rng.uniform(0.05, 0.1, nx * ny), rng.uniform(0.1, 0.2, nx * ny),
                  rng.uniform(0.1, 0.2, nx * ny), rng.uniform(0, 1e-3, nx * ny)], -1)
    o.write64("T2", T)
    o.write64("Y2", Y.reshape(-1))
    o.run_stage(1)
    rho0 = RHO_AMB * TAMB / T
    dT = o.read64("T") - T
    Yn = o.read64("Y").reshape(-1, 4)
    assert abs(np.sum(rho0 * dT)) < 1e-10 * np.sum(np.abs(rho0 * dT))
    for c in range(3):
        d = Yn[:, c] - Y[:, c]
        assert abs(np.sum(rho0 * d)) < 1e-10 * np.sum(np.abs(rho0 * d))
    assert np.array_equal(Yn[:, 3], Y[:, 3])  # soot does not diffuse


def test_diffusion_variance_grows_2Dt():
    # uniform T: constant coefficient; the explicit 5-point scheme grows the
    # second moment of a blob by exactly 2 D dt per step and dimension
    nx, ny, dx, dt = 48, 48, 0.001, 2e-4
    o = mk(nx, ny, dx, dt)
    T0 = 800.0
    o.write64("T2", np.full(nx * ny, T0))
    x = (np.arange(nx) - 20) * dx
    X, Yg = np.meshgrid(x, x)
    blob = 0.1 * np.exp(-(X ** 2 + Yg ** 2) / (2 * (3 * dx) ** 2))
    Y = np.zeros((ny, nx, 4))
    Y[..., 0] = blob
    o.write64("Y2", Y.reshape(-1))
    o.run_stage(1)
    F = cells(o, "Y")[..., 0]
    m0, m1 = blob.sum(), F.sum()
    v0 = (blob * X ** 2).sum() / m0
    v1 = (F * X ** 2).sum() / m1
    D = o.closure("rhoDOf", T0) / o.closure("rhoOf", T0)
    assert abs(m1 / m0 - 1) < 1e-12
    assert abs((v1 - v0) / (2 * D * dt) - 1) < 1e-6


# ---------------- sources (shaders/sources.wgsl) ----------------
def _one_cell(o, T, Y, Tadv=None):
    n = o.nx * o.ny
    o.write64("T", np.full(n, T))
    o.write64("T2", np.full(n, Tadv if Tadv is not None else T))
    o.write64("Y", np.tile(np.asarray(Y, float), n))
    o.run_stage(2)
    return cells(o, "T")[4, 4], cells(o, "Y")[4, 4], cells(o, "SQ")[4, 4]


def test_cold_gas_does_not_react():
    o = mk(8, 8)
    T, Y, SQ = _one_cell(o, 300.0, [0.05, 0.2, 0.0, 0.0])
    assert abs(T - 300.0) < 1e-9 and abs(Y[0] - 0.05) < 1e-12 and abs(SQ[0]) < 1e-6
    assert SQ[1] < 1e-6


def test_reaction_balances():
    o = mk(8, 8, sootFormA=0.0, kappaSootC=0.0, kappaGas=0.0)
    Y0 = np.array([0.05, 0.2, 0.1, 0.0])
    T, Y, SQ = _one_cell(o, 1500.0, Y0)
    burned = Y0[0] - Y[0]
    sO2 = 15.0e6 / 13.1e6
    assert burned > 1e-4
    assert abs(Y.sum() - Y0.sum()) < 1e-12                          # F + s O -> (1+s) P
    assert abs((Y0[1] - Y[1]) - sO2 * burned) < 1e-7
    assert abs((T - 1500.0) - burned * 15.0e6 / 1300.0) < 1e-3      # dT = dHc/cp * burned
    rho0 = RHO_AMB * TAMB / 1500.0
    assert abs(SQ[1] / (rho0 * burned * 15.0e6 / 1e-3) - 1) < 1e-5  # qrel = rho0 burned dHc / dt


def test_reaction_limited_by_stoichiometry():
    o = mk(8, 8, reactB=1e20, sootFormA=0.0, kappaSootC=0.0, kappaGas=0.0)
    sO2 = f32(15.0e6 / 13.1e6)
    T, Y, SQ = _one_cell(o, 1500.0, [0.05, 0.02, 0.0, 0.0])  # oxygen-limited
    assert abs(Y[1]) < 1e-12 and abs((0.05 - Y[0]) - 0.02 / sO2) < 1e-12


def test_radiation_implicit_limits():
    # huge dt: the linearised implicit update is one Newton step on T^4 = Tamb^4
    T0 = 1600.0
    o = mk(8, 8, dt=1e9, reactB=0.0, sootFormA=0.0, sootOxA=0.0)
    T, Y, SQ = _one_cell(o, T0, [0.0, 0.0, 1.0, 0.0])
    newton = T0 - (T0 ** 4 - TAMB32 ** 4) / (4 * T0 ** 3)
    assert abs(T - newton) < 1e-3
    # small dt: explicit loss rate 4 kappa sigma (T^4 - Tamb^4) / (rho cp)
    dt = 1e-6
    o = mk(8, 8, dt=dt, reactB=0.0, sootFormA=0.0, sootOxA=0.0)
    T, Y, SQ = _one_cell(o, T0, [0.0, 0.0, 1.0, 0.0])
    rate = 4 * 0.8 * SIGMA * (T0 ** 4 - TAMB ** 4) / (RHO_AMB * TAMB / T0 * 1300.0)
    assert abs((T0 - T) / (rate * dt) - 1) < 1e-3


def test_entrainment_source_is_volume_growth():
    nx, ny, dt = 8, 8, 1e-5
    o = mk(nx, ny, dt=dt, reactB=0.0, sootFormA=0.0, kappaGas=0.0, kappaSootC=0.0)
    o.write64("U", np.full((nx + 1) * ny, 2.0))
    T0 = 1200.0
    T, Y, SQ = _one_cell(o, T0, [0.0, 0.1, 0.5, 0.0])
    me = 2 * 0.05 * 2.0 * RHO_AMB / 0.8
    rho0 = RHO_AMB * TAMB / T0
    m_new = rho0 + me * dt
    expand = (m_new / (RHO_AMB * TAMB / T) - 1) / dt     # volume growth rate of the mixed parcel
    assert abs(SQ[0] / expand - 1) < 1e-3
    assert T < T0 and Y[1] > 0.1                           # towards ambient air
    # no velocity, no entrainment
    o2 = mk(nx, ny, dt=dt, reactB=0.0, sootFormA=0.0, kappaGas=0.0, kappaSootC=0.0)
    T, Y, SQ = _one_cell(o2, T0, [0.0, 0.1, 0.5, 0.0])
    assert abs(SQ[0]) < 1e-9 and abs(T - T0) < 1e-9


def _square(x, y, half, Ts=700.0, mflux=0.0, cov=1.0, angle=0.0):
    p = np.zeros((64, 20), np.float32)
    p[0, :8] = [x, y, angle, half, 0, 0, 0, 1]
    p[0, 8:12] = Ts
    p[0, 12:16] = mflux
    p[0, 16:20] = cov
    return p.reshape(-1)


def test_char_rate_matches_char_test_formula():
    # tests/char_test.mjs:19-31 restate the GPU char oxidation "as on the GPU":
    # m_O2'' = A exp(-Ta/Ts) rho Y_O; smoulders at 0.05-0.2 g/m^2/s at 700 K.
    nx = ny = 48
    dx, dt = 0.001, 1e-3   # coarse enough that each cell's fixed-point char mass is ~70 units
    o = Oracle(nx, ny, dx, dt, fp64=True)
    o.reset()
    half = 0.01
    o.upload_pieces(_square(0.024, 0.024, half), 1)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    n = nx * ny
    o.write64("T", np.full(n, 700.0))      # gas at the surface temperature: no conduction
    o.write64("T2", np.full(n, 700.0))
    o.run_stage(2)
    acc = o.read("acc")
    rho = RHO_AMB * TAMB / 700.0
    kc = 6.6 * math.exp(-5800.0 / 700.0)
    expect = kc * rho * 0.232 / 2.667           # kg C per m^2 s
    for f in range(4):
        burnt = acc[f * 2 + 1] / 1e12 / dt / (2 * half)
        assert abs(burnt / expect - 1) < 0.02, (f, burnt, expect)   # band weights integrate to the face length
        assert 0.05 < burnt * 1000 < 0.2


# ---------------- forces (shaders/forces.wgsl) ----------------
def test_forces_rest_and_buoyancy():
    nx, ny, dt = 12, 10, 1e-3
    o = mk(nx, ny, dt=dt)
    o.run_stage(3)
    assert not np.any(o.read64("U")) and not np.any(o.read64("V"))
    o.write64("T", np.full(nx * ny, 600.0))
    o.run_stage(3)
    V = cells(o, "V")
    assert np.allclose(V[1:ny], dt * 9.81 * (600.0 / TAMB - 1), rtol=1e-12)
    assert not np.any(V[0]) and not np.any(V[ny])


def test_forces_uniform_u_entrainment_drag():
    nx, ny, dt = 12, 10, 1e-3
    o = mk(nx, ny, dt=dt)
    u0, T0 = 1.5, 900.0
    o.write64("T", np.full(nx * ny, T0))
    o.write64("U2", np.full((nx + 1) * ny, u0))
    o.run_stage(3)
    U = cells(o, "U")
    expect = u0 * math.exp(-2 * 0.05 * u0 * (T0 / TAMB) / 0.8 * dt)
    assert np.allclose(U[1:ny - 1, 1:nx], expect, rtol=1e-12)   # no shear away from floor and top
    assert np.all(U[0, 1:nx] < expect)                          # the no-slip floor drags row 0


# ---------------- projection and multigrid (shaders/pressure.wgsl) ----------------
def _poisson_direct(o, l=0):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    nx, ny = o.nx >> l, o.ny >> l
    sfx = "" if l == 0 else str(l)
    ax = o.read64("ax" + sfx).reshape(ny, nx + 1)
    ay = o.read64("ay" + sfx).reshape(ny + 1, nx)
    rhs = o.read64("rhs" + sfx).reshape(ny, nx)
    h2 = f32((o.dx * 2 ** l) ** 2)  # Level.h2 is an f32 (js/fluid.js:82)
    A = sp.lil_matrix((nx * ny, nx * ny))
    b = np.zeros(nx * ny)
    for j in range(ny):
        for i in range(nx):
            k = j * nx + i
            w, e, s, n = ax[j, i], ax[j, i + 1], ay[j, i], ay[j + 1, i]
            diag = 0.0
            if i > 0: A[k, k - 1] = w; diag += w
            else: diag += 2 * w
            if i < nx - 1: A[k, k + 1] = e; diag += e
            else: diag += 2 * e
            if j > 0: A[k, k - nx] = s; diag += s
            if j < ny - 1: A[k, k + nx] = n; diag += n
            else: diag += 2 * n
            if diag > 0:
                A[k, k] = -diag
                b[k] = rhs[j, i] * h2
            else:
                A[k, k] = 1.0
    return spl.spsolve(A.tocsr(), b).reshape(ny, nx)


def _projection_case(nx=32, ny=32, seed=5, fp64=True):
    dx, dt = 0.004, 1e-3
    o = Oracle(nx, ny, dx, dt, fp64=fp64)
    o.reset()
    o.upload_pieces(_square(0.064, 0.03, 0.012, angle=0.3), 1)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    rng = np.random.default_rng(seed)
    o.write64("T", rng.uniform(300, 1500, nx * ny))
    o.write64("U", rng.uniform(-1, 1, (nx + 1) * ny))
    o.write64("V", rng.uniform(-1, 1, nx * (ny + 1)))
    sq = np.zeros((nx * ny, 2))
    sq[:, 0] = rng.uniform(-5, 5, nx * ny)
    o.write64("SQ", sq.reshape(-1))
    return o


def test_multigrid_converges_to_direct_solve():
    o = _projection_case()
    o.run_stage(4)
    o.set_schedule(nu=2, vcycles=40, tol=1e-11, coarse_iters=0)
    assert o.run_stage(5) == 0
    p = cells(o, "p")
    ref = _poisson_direct(o)
    assert np.max(np.abs(p - ref)) < 1e-8 * np.max(np.abs(ref))


def test_converged_projection_is_divergence_consistent():
    o = _projection_case()
    o.run_stage(4)
    o.set_schedule(nu=2, vcycles=40, tol=1e-12, coarse_iters=0)
    assert o.run_stage(5) == 0
    nx, ny, dx = o.nx, o.ny, o.dx
    fx, fy, SQ = cells(o, "fx"), cells(o, "fy"), cells(o, "SQ")

    def div(U, V):
        return (fx[:, 1:] * U[:, 1:] - fx[:, :-1] * U[:, :-1] + fy[1:] * V[1:] - fy[:-1] * V[:-1]) / dx

    d0 = div(cells(o, "U"), cells(o, "V")) - SQ[..., 0]
    o.run_stage(6)
    U, V = cells(o, "U"), cells(o, "V")
    open_ = (fx[:, 1:] + fx[:, :-1] + fy[1:] + fy[:-1]) > 0
    err = np.abs(div(U, V) - SQ[..., 0])[open_]
    # H.divErr after a converged solve.  Not zero to round-off: the V-cycle
    # uses Level.h2 = f32(h*h) of the double h (js/fluid.js:82) while rhs0 and
    # project use the f32 dx, a 1e-7 relative mismatch of the reference itself.
    assert err.max() < 1e-6 * np.abs(d0).max()
    # discrete divergence theorem: net source = outflow through the open sides and top
    net = np.sum(SQ[..., 0][open_]) * dx * dx
    out = (np.sum(fx[:, nx] * U[:, nx]) - np.sum(fx[:, 0] * U[:, 0]) + np.sum(fy[ny] * V[ny])) * dx
    assert abs(net - out) < 1e-6 * np.sum(np.abs(d0)) * dx * dx


def test_reference_coarsest_sweeps_vs_direct():
    # the 120-sweep coarsest level (shaders/pressure.wgsl:141) is an approximate
    # solve; quantify its error on a 16x16 coarsest level
    o = _projection_case(16, 16)
    o.run_stage(4)
    o.run_kernel("coarsest", 120)
    p = cells(o, "p")
    ref = _poisson_direct(o)
    rel = np.linalg.norm(p - ref) / np.linalg.norm(ref)
    assert rel < 0.2
    o.run_kernel("coarsest", 200000)
    p = cells(o, "p")
    assert np.linalg.norm(p - ref) / np.linalg.norm(ref) < 1e-9


# ---------------- geometry (shaders/geometry.wgsl) ----------------
def test_geometry_sdf_faces_and_band_weights():
    nx = ny = 64
    dx = 0.001
    o = Oracle(nx, ny, dx, 1e-3, fp64=True)
    o.reset()
    cx, cy, half = f32(0.0321), f32(0.0307), f32(0.0105)
    o.upload_pieces(_square(cx, cy, half), 1)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    sdf = o.read64("sdf").reshape(ny, nx)
    ids = o.read("cellG").reshape(ny, nx, 2)[..., 1].view(np.uint32)
    x = (np.arange(nx) + 0.5) * f32(dx)
    X, Yc = np.meshgrid(x, x)
    qx, qy = np.abs(X - cx) - half, np.abs(Yc - cy) - half
    ref = np.hypot(np.maximum(qx, 0), np.maximum(qy, 0)) + np.minimum(np.maximum(qx, qy), 0)
    assert np.max(np.abs(sdf - ref)) < 1e-12
    j, i = int(cy / dx), int((cx + half) / dx) + 2
    assert ids[j, i] == 0 and ids[int((cy + half) / dx) + 2, int(cx / dx)] == 1
    assert ids[j, int((cx - half) / dx) - 2] == 2 and ids[int((cy - half) / dx) - 2, int(cx / dx)] == 3
    fw = o.read("faceW")[:4] / 1e8
    # hat band integrates to the face length plus the corner regions (~dx/L each)
    assert np.all(fw / (2 * half) > 1) and np.all(fw / (2 * half) < 1.1)
    fy = o.read64("fy").reshape(ny + 1, nx)
    assert not fy[0].any()
    fx = o.read64("fx").reshape(ny, nx + 1)
    assert fx.min() == 0.0 and fx.max() == 1.0


def test_band_weight_partition_of_unity_on_long_faces():
    # straight faces 160 cells long: the corner share drops below 1 %
    nx = ny = 256
    dx = 0.001
    o = Oracle(nx, ny, dx, 1e-3, fp64=True)
    o.reset()
    half = f32(0.08)
    o.upload_pieces(_square(f32(0.128), f32(0.128), half, angle=f32(0.2)), 1)
    o.write_params()
    o.geometry()
    fw = o.read("faceW")[:4] / 1e8
    assert np.all(np.abs(fw / (2 * half) - 1) < 0.012), fw / (2 * half)


def test_geometry_rotation_invariance():
    nx = ny = 48
    dx = 0.001
    sdfs = []
    for ang in (0.0, math.pi / 2):
        o = Oracle(nx, ny, dx, 1e-3, fp64=True)
        o.reset()
        o.upload_pieces(_square(0.024, 0.024, 0.008, angle=ang), 1)
        o.write_params()
        o.geometry()
        sdfs.append(o.read64("sdf"))
    # cos(f32(pi/2)) = -4.4e-8: a rotation of that size remains
    assert np.max(np.abs(sdfs[0] - sdfs[1])) < 1e-8


# ---------------- radiation (shaders/radiation.wgsl) ----------------
def _floor_probe_rays(frame, tgt, nx, ny, dx, fi):
    """Independent geometry of the 64 floor-probe rays: exit distance from the
    probe origin to the domain box, for an optically thin uniform slab."""
    jit = (frame * 0.7548777 + tgt * 0.5698403) % 1.0
    ox = (fi + 0.5) / 64 * nx * f32(dx)
    W, H = nx * f32(dx), ny * f32(dx)
    out = []
    for li in range(64):
        a = ((li + jit) / 64 - 0.5) * PI_WGSL
        d = (math.cos(math.pi / 2 + a), math.sin(math.pi / 2 + a))
        ts = []
        if d[0] > 1e-6: ts.append((W - ox) / d[0])
        elif d[0] < -1e-6: ts.append(-ox / d[0])
        if d[1] > 1e-6: ts.append((H - 1e-4) / d[1])
        out.append((math.cos(a), min(ts)))
    return out


def test_radiation_black_enclosure_limit():
    o = mk(32, 32, dx=0.004)
    o.radiation()
    rad = o.read64("rad").reshape(-1, 4)
    q = rad[:64, 0]
    assert np.all(np.abs(q / (SIGMA * TAMB ** 4) - 1) < 2e-3)


def test_radiation_optically_thin_slab():
    nx = ny = 32
    dx = 0.004
    Tg = 1500.0
    o = mk(nx, ny, dx=dx)
    n = nx * ny
    o.write64("T", np.full(n, Tg))
    Y = np.zeros((n, 4)); Y[:, 2] = 1.0
    o.write64("Y", Y.reshape(-1))
    o.radiation()
    q = o.read64("rad").reshape(-1, 4)[:64, 0]
    g = f32(0.8) * SIGMA32 * (Tg ** 4 - TAMB32 ** 4) / PI_WGSL
    for fi in (0, 17, 40, 63):
        rays = _floor_probe_rays(0, fi, nx, ny, dx, fi)
        expect = sum(PI_WGSL / 64 * c * (0.5 * SIGMA32 * TAMB32 ** 4 + 2 * g * L) for c, L in rays)
        assert abs(q[fi] / expect - 1) < 1e-9, (fi, q[fi], expect)


def test_radiation_temporal_smoothing():
    o = mk(16, 16, dx=0.004)
    o.radiation()
    q1 = o.read64("rad").reshape(-1, 4)[:64, 0].copy()
    o.write64("T", np.full(256, 1200.0))
    Y = np.zeros((256, 4)); Y[:, 2] = 1.0
    o.write64("Y", Y.reshape(-1))
    o.radiation()
    q2 = o.read64("rad").reshape(-1, 4)[:64, 0]
    o3 = mk(16, 16, dx=0.004)
    o3.write64("T", np.full(256, 1200.0))
    o3.write64("Y", Y.reshape(-1))
    o3.radiation()
    fresh = o3.read64("rad").reshape(-1, 4)[:64, 0]
    assert np.allclose(q2, 0.65 * q1 + 0.35 * fresh, rtol=1e-12)   # mix(old, new, 0.35)


def test_direct_coarsest_solve_is_exact():
    """coarse_iters = 0 solves the coarsest level by Cholesky in FP64: on a
    16x16 grid the coarsest level is the whole problem, so one V-cycle leaves
    a residual at round-off."""
    o = Oracle(16, 16, 0.01, 0.002, fp64=True)
    o.reset()
    pieces, n = scenes.pile(2, 9, 0.16, spread=0.05)
    o.upload_pieces(pieces, n)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    st = scenes.random_state(16, 16, 9)
    for name in ("T", "Y", "U", "V"):
        o.write(name, st[name])
    o.run_stage(4)
    o.set_schedule(nu=2, vcycles=1, tol=1e-30, coarse_iters=0)
    o.run_stage(5)
    assert o.last_resid < 1e-12, o.last_resid
    o.set_schedule(nu=2, vcycles=1, tol=1e-30, coarse_iters=120)
    o.write("p", np.zeros(256))
    o.run_stage(5)
    assert o.last_resid > 1e-6   # 120 sweeps alone are not exact


# ---------------- FP16 storage (the FP16 smoother) ----------------
def test_half_conversion_matches_numpy_float16():
    """The oracle's binary16 conversion (round to nearest even, subnormals,
    overflow to inf) against numpy's: every half value back to float, and
    random float bit patterns, boundary values and every tie between
    adjacent halves forward."""
    from oracle import pyoracle as po
    h = np.arange(65536, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = po.h2f(h)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))
    rng = np.random.default_rng(1)
    f = rng.integers(0, 2 ** 32, 500_000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    hv = np.arange(0, 0x7bff, dtype=np.uint16)
    mid = ((hv.view(np.float16).astype(np.float64) + (hv + 1).view(np.float16).astype(np.float64)) / 2)
    edge = np.array([65504, 65519.99, 65520, 65536, 6.1035156e-05, 5.96e-08, 2.98e-08, 2.99e-08, 1e-9,
                     0.0, -0.0, np.inf, -np.inf], np.float32)
    f = np.concatenate([f, edge, mid.astype(np.float32), -mid.astype(np.float32)])
    with np.errstate(over="ignore", invalid="ignore"):
        exp = f.astype(np.float16).view(np.uint16)
    got = po.f2h(f)
    fin = ~np.isnan(f)
    assert np.array_equal(got[fin], exp[fin])
    assert np.all((got[~fin] & 0x7c00) == 0x7c00) and np.all((got[~fin] & 0x3ff) != 0)


@pytest.mark.parametrize("n", [128, 256])
def test_fp16_smoother_defect_correction_reaches_the_direct_solve(n):
    """FP16 storage of p and the coefficients at the tiled levels (one at
    128^2, two at 256^2), FP32 sweeps, FP64 defect correction: the FP64
    pressure converges to the sparse direct solve of the level-0 system, the
    coefficients come back unchanged, and the FP16 V-cycle's correction
    differs from the FP32 V-cycle's by FP16 rounding, not more or less."""
    o = _projection_case(n, n, fp64=False)
    o.run_stage(4)
    ref = _poisson_direct(o)
    ax0, ay0 = o.read("ax").copy(), o.read("ay").copy()
    p_start = o.read("p").copy()
    o.set_schedule(nu=2, vcycles=60, tol=1e-9, coarse_iters=0, refine=1, smoother_fp16=1)
    assert o.run_stage(5) == 0
    cyc16 = o.last_vcycles
    p64 = o.read64("p64").reshape(n, n)
    assert np.max(np.abs(p64 - ref)) < 1e-6 * np.max(np.abs(ref))
    assert np.array_equal(o.read("ax"), ax0) and np.array_equal(o.read("ay"), ay0)
    # the same solve with the FP32 smoother: as many cycles, give or take two
    o2 = _projection_case(n, n, fp64=False)
    o2.run_stage(4)
    o2.set_schedule(nu=2, vcycles=60, tol=1e-9, coarse_iters=0, refine=1)
    assert o2.run_stage(5) == 0
    assert o2.last_vcycles - 2 <= cyc16 <= o2.last_vcycles + 2, (o2.last_vcycles, cyc16)
    # one correction V-cycle each from the same start: about 1 % apart (the
    # FP16 rounding of p at 5e-4 relative, spread by the sweeps that follow);
    # both cut the residual by well over 10x
    outs = []
    for fp16 in (0, 1):
        o3 = _projection_case(n, n, fp64=False)
        o3.run_stage(4)
        o3.write("p", p_start)
        o3.set_schedule(nu=2, vcycles=1, tol=1e-30, coarse_iters=0, refine=1, smoother_fp16=fp16)
        o3.run_stage(5)
        outs.append(o3.read64("p64") - p_start.astype(np.float64))
        assert o3.last_resid < 0.05, (fp16, o3.last_resid)
    d = np.linalg.norm(outs[1] - outs[0]) / np.linalg.norm(outs[0])
    assert 1e-4 < d < 5e-2, d


def _apply_B(o, z):
    """B z = (sum of a z over the neighbours - diag z) / h2 at level 0 in
    float64, with the relaxation's boundary terms (twice the face coefficient
    at the open sides and the top), 0 at closed cells."""
    nx, ny = o.nx, o.ny
    ax = o.read64("ax").reshape(ny, nx + 1)
    ay = o.read64("ay").reshape(ny + 1, nx)
    z = np.asarray(z, np.float64).reshape(ny, nx)
    off = np.zeros((ny, nx))
    diag = np.zeros((ny, nx))
    w, e, s, n = ax[:, :-1], ax[:, 1:], ay[:-1], ay[1:]
    off[:, 1:] += w[:, 1:] * z[:, :-1]
    diag[:, 1:] += w[:, 1:]
    diag[:, 0] += 2 * w[:, 0]
    off[:, :-1] += e[:, :-1] * z[:, 1:]
    diag[:, :-1] += e[:, :-1]
    diag[:, -1] += 2 * e[:, -1]
    off[1:] += s[1:] * z[:-1]
    diag[1:] += s[1:]
    off[:-1] += n[:-1] * z[1:]
    diag[:-1] += n[:-1]
    diag[-1] += 2 * n[-1]
    h2 = float(f32(o.dx ** 2))
    return np.where(diag > 0, (off - diag * z) / h2, 0.0).reshape(-1)


@pytest.mark.parametrize("n", [0, 1, 7, 256, 1023, 1024, 1025, 4097, 300001, (1 << 20) + 12345])
def test_gpu_order_sum_is_a_sum(n):
    """The fixed-order FP64 sum that the Krylov dot products use adds every
    element once: it equals the exactly rounded sum to FP64 rounding, over
    sizes with one partial workgroup, exactly 1024 workgroups and more
    elements than 1024 x 256 threads."""
    from oracle.pyoracle import gpu_sum
    v = np.random.default_rng(n).standard_normal(n) * 10.0 ** np.random.default_rng(n + 1).integers(-3, 3, n)
    s = gpu_sum(v)
    assert abs(s - math.fsum(v)) <= 1e-13 * max(np.sum(np.abs(v)), 1e-300)
    if n > 1:  # one element changed by a unit is seen, whichever it is
        for k in (0, n // 2, n - 1):
            w = v.copy()
            w[k] += 1.0
            assert abs(gpu_sum(w) - s - 1.0) < 1e-6 * max(1.0, np.sum(np.abs(v)) * 1e-9)


@pytest.mark.parametrize("n,fp16", [(64, 0), (128, 0), (128, 1)])
def test_krylov_reaches_the_direct_solve_in_fewer_cycles(n, fp16):
    """Flexible CG with the V-cycle as preconditioner (24 coarsest sweeps, so
    the V-cycle alone converges slowly): the FP64 pressure converges to the
    sparse direct solve of the level-0 system, in fewer V-cycles than defect
    correction alone with the same V-cycle."""
    cyc = []
    for kry in (0, 1):
        o = _projection_case(n, n, fp64=False)
        o.run_stage(4)
        ref = _poisson_direct(o)
        o.set_schedule(nu=2, vcycles=80, tol=1e-9, coarse_iters=24, refine=1, smoother_fp16=fp16, krylov=kry)
        assert o.run_stage(5) == 0
        p64 = o.read64("p64").reshape(n, n)
        assert np.max(np.abs(p64 - ref)) < 1e-6 * np.max(np.abs(ref)), kry
        assert o.last_resid < 1e-9
        cyc.append(o.last_vcycles)
        o.close()
    assert cyc[1] < cyc[0], cyc


def test_krylov_iteration_identities():
    """After k = 1..4 iterations (cycle cap k): the right-hand side left for
    the next V-cycle is the FP32 rounding of b - B p64 (the residual from the
    pressure), (r, d) = 0 (exact line search along d), and (d_k, B d_{k-1}) = 0
    (each direction conjugate to the one before, to the FP32 rounding of the
    stored direction).  A wrong sign or a dropped
    term in the direction, step or residual breaks one of them."""
    n = 96
    prev_d = None
    for k in range(1, 5):
        o = _projection_case(n, n, fp64=False)
        o.run_stage(4)
        b = o.read64("rhs")
        o.set_schedule(nu=2, vcycles=k, tol=1e-30, coarse_iters=24, refine=1, krylov=1)
        o.run_stage(5)
        assert o.last_vcycles == k
        r, d, x = o.read64("rhs"), o.read64("d32"), o.read64("p64")
        nb, nr = np.linalg.norm(b), np.linalg.norm(r)
        assert np.linalg.norm(r - (b - _apply_B(o, x))) <= 1e-7 * nr + 1e-12 * nb, k
        assert abs(r @ d) <= 1e-5 * nr * np.linalg.norm(d), k
        if prev_d is not None:  # to the FP32 rounding of the stored directions
            bq = _apply_B(o, prev_d)
            assert abs(d @ bq) <= 1e-6 * np.linalg.norm(d) * np.linalg.norm(bq), k
        prev_d = d
        assert np.isclose(o.last_resid, nr / nb, rtol=1e-6)
        o.close()


def test_krylov_needs_defect_correction():
    o = _projection_case(32, 32, fp64=False)
    with pytest.raises(Exception):
        o.set_schedule(nu=2, vcycles=4, tol=1e-6, krylov=1)
    o.close()


def _pocket_case(n=48, seed=5):
    """A ring of eight overlapping squares encloses open cells that reach no
    Dirichlet side: the level-0 system is singular there, and random sources
    put the right-hand side partly outside its range."""
    o = Oracle(n, n, 0.002, 1e-3)
    o.reset()
    p = np.zeros((64, 20), np.float32)
    for q in range(8):
        a = 2 * np.pi * q / 8
        p[q, :8] = [0.048 + 0.02 * np.cos(a), 0.045 + 0.02 * np.sin(a), a, 0.01, 0, 0, 0, 1]
        p[q, 8:12] = 300.0
        p[q, 16:20] = 1.0
    o.upload_pieces(p.reshape(-1), 8)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    rng = np.random.default_rng(seed)
    o.write("T", rng.uniform(300, 1500, n * n).astype(np.float32))
    o.write("U", rng.uniform(-1, 1, (n + 1) * n).astype(np.float32))
    o.write("V", rng.uniform(-1, 1, n * (n + 1)).astype(np.float32))
    sq = np.zeros((n * n, 2), np.float32)
    sq[:, 0] = rng.uniform(-5, 5, n * n)
    o.write("SQ", sq.reshape(-1))
    return o


def test_krylov_falls_back_on_a_singular_system():
    """On a system with an enclosed pocket, defect correction stalls at the
    incompatible part of the right-hand side; the Krylov iteration, which
    would diverge there, falls back to defect correction and ends at the same
    residual with the same bounded pressure."""
    out = []
    for kry in (0, 1):
        o = _pocket_case()
        o.run_stage(4)
        o.set_schedule(nu=2, vcycles=30, tol=1e-9, coarse_iters=24, refine=1, krylov=kry)
        assert o.run_stage(5) != 0  # no convergence: the target is out of reach
        out.append((o.last_resid, np.max(np.abs(o.read64("p64")))))
        o.close()
    (r_dc, p_dc), (r_kr, p_kr) = out
    assert 1e-5 < r_dc < 1e-2, r_dc
    assert r_kr <= 1.5 * r_dc and p_kr <= 2 * p_dc, out


def test_krylov_start_rescales_the_previous_pressure():
    """The Krylov start is the line search along the previous pressure p
    (rounded to FP32) from zero: when p is a multiple of this step's
    solution, x = (p, b)/(p, B p) p is the solution to that rounding, so the
    solve ends before any V-cycle; when p is zero it starts from x = 0 with
    r = b."""
    n = 64
    o = _projection_case(n, n, fp64=False)
    o.run_stage(4)
    ref = _poisson_direct(o).reshape(-1)
    o.set_schedule(nu=2, vcycles=10, tol=1e-6, coarse_iters=24, refine=1, krylov=1)
    o.write64("p64", 3.7 * ref)
    assert o.run_stage(5) == 0
    # the direction is stored in FP32, so the rescaled solution is exact to
    # that rounding
    assert o.last_vcycles == 0 and o.last_resid < 1e-6, (o.last_vcycles, o.last_resid)
    assert np.max(np.abs(o.read64("p64") - ref)) < 1e-6 * np.max(np.abs(ref))
    o2 = _projection_case(n, n, fp64=False)
    o2.run_stage(4)
    b = o2.read64("rhs")
    o2.set_schedule(nu=2, vcycles=1, tol=1e-30, coarse_iters=24, refine=1, krylov=1)
    o2.write64("p64", np.zeros(n * n))
    o2.run_stage(5)
    # one iteration from x = 0: the step is alpha z with z the V-cycle of b,
    # so the new residual is b - alpha B z, orthogonal to z = d
    r, d = o2.read64("rhs"), o2.read64("d32")
    assert abs(r @ d) <= 1e-5 * np.linalg.norm(r) * np.linalg.norm(d)
    assert np.linalg.norm(r - (b - _apply_B(o2, o2.read64("p64")))) <= 1e-7 * np.linalg.norm(r)
    # from a previous pressure that is not a multiple of the solution, the
    # first direction is conjugate to it: (d_1, B p) = 0
    o3 = _projection_case(n, n, fp64=False)
    o3.run_stage(4)
    prev = ref + np.random.default_rng(7).standard_normal(n * n) * np.std(ref)
    o3.set_schedule(nu=2, vcycles=1, tol=1e-30, coarse_iters=24, refine=1, krylov=1)
    o3.write64("p64", prev)
    o3.run_stage(5)
    bp = _apply_B(o3, prev.astype(np.float32).astype(np.float64))
    d1 = o3.read64("d32")
    assert abs(d1 @ bp) <= 1e-6 * np.linalg.norm(d1) * np.linalg.norm(bp)
    o.close()
    o2.close()
    o3.close()
