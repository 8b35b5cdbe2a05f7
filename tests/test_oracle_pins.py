"""Pins of the oracle's multigrid, radiation, viscous stencil and face
fractions, each by a property or an independent computation that a plausible
mistake fails -- and the proof that it does: every check is also run against
a build of the oracle with one such mistake planted (CFO_MUTANT in
oracle/cf_oracle.c) and must fail there.

- V-cycle (js/fluid.js:294-312, shaders/pressure.wgsl:77-156): one cycle
  equals an independent dense restatement (sparse operator, colour masks,
  restriction and prolongation matrices) to FP64 round-off; after it the
  colour swept last at level 0 (colour 0) has zero residual and the other
  does not; a single half-sweep zeroes its own colour's residual and leaves
  the other colour alone.
- prolongation (pressure.wgsl:116-129): exact on coarse fields linear in x
  and y away from the edges, the clamped neighbour (constant extrapolation)
  at the edges, no correction in closed cells.
- restriction (pressure.wgsl:104-113): the block mean of the residual, and
  the coarse p zeroed.
- radiation (radiation.wgsl:36-51,64-158): a face under a second piece and a
  downward face whose rays bounce off the floor, against brute-force ray
  marching of the same rays (point-in-box tests every 20 um).
- viscous stencil (forces.wgsl:34-74): a discrete sine mode at uniform T
  grows by exactly its eigenvalue factor; with T varying in y the corner
  viscosities are those of the four-cell mean.
- face fractions (geometry.wgsl:44-50): clamp(1/2 + d/dx, 0, 1) at known
  signed distances from an axis-aligned box; the floor row of fy is 0.
- the oracle's x*sqrt(x) for pow(x, 1.5) is within one ulp of the correctly
  rounded power over the temperatures the closures see.
"""
import concurrent.futures as cf
import math
import os
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
PI_WGSL = 3.14159265
f32 = lambda v: float(np.float32(v))


# ------------------------------------------------------------------ multigrid
def _levels(nx, ny, dx):
    """Level dims and h^2 as js/fluid.js:69-91 builds them (h^2 = f32(h*h))."""
    out, lx, ly, h = [], nx, ny, dx
    while True:
        out.append((lx, ly, float(np.float32(h * h))))
        if lx % 2 or ly % 2 or lx * ly <= 256:
            return out
        lx, ly, h = lx // 2, ly // 2, h * 2


class Lev:
    """Dense/vectorised restatement of one level's operator: the masks of the
    neighbours that exist, the diagonal with the Dirichlet doubling at the
    open sides and the top and nothing at the floor."""

    def __init__(self, ax, ay, h2):
        self.ax, self.ay, self.h2 = ax, ay, h2
        ny, nx1 = ax.shape
        nx = nx1 - 1
        self.nx, self.ny = nx, ny
        i = np.arange(nx)[None, :]
        j = np.arange(ny)[:, None]
        self.hw, self.he = (i > 0) & (j >= 0), (i < nx - 1) & (j >= 0)
        self.hs, self.hn = (j > 0) & (i >= 0), (j < ny - 1) & (i >= 0)
        w, e, s, n = ax[:, :-1], ax[:, 1:], ay[:-1, :], ay[1:, :]
        self.w, self.e, self.s, self.n = w, e, s, n
        self.diag = (np.where(self.hw, w, 2 * w) + np.where(self.he, e, 2 * e) + np.where(self.hs, s, 0 * s)
                     + np.where(self.hn, n, 2 * n))
        self.color = (i + j) & 1

    def off(self, p):
        z = np.zeros_like(p)
        pw = np.zeros_like(p); pw[:, 1:] = p[:, :-1]
        pe = np.zeros_like(p); pe[:, :-1] = p[:, 1:]
        ps = np.zeros_like(p); ps[1:, :] = p[:-1, :]
        pn = np.zeros_like(p); pn[:-1, :] = p[1:, :]
        return (np.where(self.hw, self.w * pw, z) + np.where(self.he, self.e * pe, z)
                + np.where(self.hs, self.s * ps, z) + np.where(self.hn, self.n * pn, z))

    def relax(self, p, rhs, c):
        new = np.where(self.diag > 0, (self.off(p) - rhs * self.h2) / np.where(self.diag > 0, self.diag, 1), 0.0)
        return np.where(self.color == c, new, p)

    def residual(self, p, rhs):
        return np.where(self.diag > 0, rhs - (self.off(p) - self.diag * p) / self.h2, 0.0)


def _restrict_coef(ax, ay):
    nyc, nxc = ax.shape[0] // 2, (ax.shape[1] - 1) // 2
    axc = 0.5 * (ax[0:2 * nyc:2, 0:2 * nxc + 1:2] + ax[1:2 * nyc:2, 0:2 * nxc + 1:2])
    ayc = 0.5 * (ay[0:2 * nyc + 1:2, 0:2 * nxc:2] + ay[0:2 * nyc + 1:2, 1:2 * nxc:2])
    return axc, ayc


def _restrict(r):
    return 0.25 * (r[0::2, 0::2] + r[0::2, 1::2] + r[1::2, 0::2] + r[1::2, 1::2])


def _prolong_matrix(nx, ny):
    """Bilinear interpolation from the coarse cell centres with the
    neighbour index clamped at the edges, as a dense (fine x coarse) matrix."""
    nxc, nyc = nx // 2, ny // 2
    P = np.zeros((ny * nx, nyc * nxc))
    for j in range(ny):
        J = j // 2
        J2 = min(max(J + (1 if j % 2 else -1), 0), nyc - 1)
        for i in range(nx):
            I = i // 2
            I2 = min(max(I + (1 if i % 2 else -1), 0), nxc - 1)
            row = j * nx + i
            for (a, b, wgt) in ((J, I, 9 / 16), (J, I2, 3 / 16), (J2, I, 3 / 16), (J2, I2, 1 / 16)):
                P[row, a * nxc + b] += wgt
    return P


def dense_vcycle(ax0, ay0, rhs0, p0, dx, nu, coarse_iters=120):
    """One V-cycle of js/fluid.js:294-312 from level-0 coefficients."""
    ny, nx = rhs0.shape
    dims = _levels(nx, ny, dx)
    levs, ax, ay = [], ax0, ay0
    for l, (lx, ly, h2) in enumerate(dims):
        levs.append(Lev(ax, ay, h2))
        if l + 1 < len(dims):
            ax, ay = _restrict_coef(ax, ay)
    L = len(levs)
    p = [None] * L
    rhs = [None] * L
    p[0], rhs[0] = p0.copy(), rhs0.copy()
    for l in range(L - 1):
        for _ in range(nu):
            p[l] = levs[l].relax(p[l], rhs[l], 0)
            p[l] = levs[l].relax(p[l], rhs[l], 1)
        rhs[l + 1] = _restrict(levs[l].residual(p[l], rhs[l]))
        p[l + 1] = np.zeros_like(rhs[l + 1])
    pl = np.zeros_like(rhs[L - 1])
    for _ in range(coarse_iters):
        pl = levs[L - 1].relax(pl, rhs[L - 1], 0)
        pl = levs[L - 1].relax(pl, rhs[L - 1], 1)
    p[L - 1] = pl
    for l in range(L - 2, -1, -1):
        lv = levs[l]
        c = (_prolong_matrix(lv.nx, lv.ny) @ p[l + 1].reshape(-1)).reshape(lv.ny, lv.nx)
        p[l] = p[l] + np.where(lv.diag > 0, c, 0.0)
        for _ in range(nu):
            p[l] = lv.relax(p[l], rhs[l], 1)
            p[l] = lv.relax(p[l], rhs[l], 0)
    return p, rhs, levs


def _mg_problem(nx, ny, seed, lib_path=None, dx=0.01):
    """Level-0 coefficients with closed and fractional faces (the floor row
    closed), a random right-hand side and a random warm start."""
    rng = np.random.default_rng(seed)
    ax = rng.uniform(0.3, 3.0, (ny, nx + 1))
    ay = rng.uniform(0.3, 3.0, (ny + 1, nx))
    ax[rng.random(ax.shape) < 0.08] = 0.0
    ay[rng.random(ay.shape) < 0.08] = 0.0
    ax[rng.random(ax.shape) < 0.05] *= 1e-3
    ay[0, :] = 0.0
    # a closed cell (all four faces shut)
    ax[ny // 2, nx // 3] = ax[ny // 2, nx // 3 + 1] = 0.0
    ay[ny // 2, nx // 3] = ay[ny // 2 + 1, nx // 3] = 0.0
    rhs = rng.normal(size=(ny, nx))
    p = 0.1 * rng.normal(size=(ny, nx))
    o = Oracle(nx, ny, dx, 1e-3, fp64=True, lib_path=lib_path)
    o.reset()
    o.write64("ax", ax)
    o.write64("ay", ay)
    o.write64("rhs", rhs)
    o.write64("p", p)
    for l in range(o.nlevels - 1):
        o.run_kernel("coefRestrict", l)
    return o, ax, ay, rhs, p


def check_vcycle_matches_dense(lib_path=None):
    for (nx, ny, nu, seed) in ((64, 48, 2, 1), (48, 32, 1, 2), (32, 32, 3, 3)):
        o, ax, ay, rhs, p0 = _mg_problem(nx, ny, seed, lib_path)
        o.run_kernel("vcycle", nu)
        got = o.read64("p").reshape(ny, nx)
        ps, rs, _ = dense_vcycle(ax, ay, rhs, p0, 0.01, nu)
        err = np.max(np.abs(got - ps[0])) / np.max(np.abs(ps[0]))
        assert err < 1e-12, (nx, ny, nu, err)
        # the coarse right-hand sides stay as the down leg left them
        for l in range(1, o.nlevels):
            r = o.read64("rhs%d" % l).reshape(rs[l].shape)
            assert np.max(np.abs(r - rs[l])) <= 1e-12 * np.max(np.abs(rs[l])), (nx, ny, l)
        o.close()


def check_vcycle_last_colour(lib_path=None):
    nx, ny = 64, 48
    o, ax, ay, rhs, p0 = _mg_problem(nx, ny, 4, lib_path)
    o.run_kernel("vcycle", 2)
    p = o.read64("p").reshape(ny, nx)
    lv = Lev(ax, ay, _levels(nx, ny, 0.01)[0][2])
    r = lv.residual(p, rhs)
    scale = np.sqrt(np.mean(rhs ** 2))
    open_ = lv.diag > 0
    r0 = np.abs(r[(lv.color == 0) & open_])
    r1 = np.abs(r[(lv.color == 1) & open_])
    assert r0.max() < 1e-12 * scale * 1e3, r0.max()     # swept last: solved to round-off
    assert np.sqrt(np.mean(r1 ** 2)) > 1e-3 * scale      # the other colour is not
    o.close()


def check_half_sweep(lib_path=None):
    nx, ny = 48, 32
    for c in (0, 1):
        o, ax, ay, rhs, p0 = _mg_problem(nx, ny, 5 + c, lib_path)
        o.run_kernel("relax%d" % c, 0)
        p = o.read64("p").reshape(ny, nx)
        lv = Lev(ax, ay, _levels(nx, ny, 0.01)[0][2])
        other = lv.color != c
        assert np.array_equal(p[other], p0[other])
        r = lv.residual(p, rhs)
        mine = (lv.color == c) & (lv.diag > 0)
        assert np.max(np.abs(r[mine])) < 1e-9 * np.sqrt(np.mean(rhs ** 2)), c
        assert np.all(p[(lv.color == c) & (lv.diag <= 0)] == 0)
        o.close()


def check_prolongation(lib_path=None):
    nx, ny = 32, 24
    o, ax, ay, rhs, p0 = _mg_problem(nx, ny, 7, lib_path)
    nxc, nyc = nx // 2, ny // 2
    a, bx, by = 0.7, 0.31, -0.23
    # coarse centres at fine coordinates (2I + 1, 2J + 1)
    X = 2 * np.arange(nxc)[None, :] + 1.0
    Y = 2 * np.arange(nyc)[:, None] + 1.0
    pc = a + bx * X + by * Y
    o.write64("p1", pc)
    o.write64("p", np.zeros((ny, nx)))
    o.run_kernel("prolong", 0)
    p = o.read64("p").reshape(ny, nx)
    lv = Lev(ax, ay, 1.0)
    xf = np.arange(nx)[None, :] + 0.5
    yf = np.arange(ny)[:, None] + 0.5
    lin = a + bx * xf + by * yf
    open_ = lv.diag > 0
    inner = np.zeros((ny, nx), bool)
    inner[1:-1, 1:-1] = True
    assert np.max(np.abs(np.where(inner & open_, p - lin, 0))) < 1e-12
    # edges: the clamped neighbour is the cell itself, so the value is that
    # of the nearest coarse centre along the clamped axis
    xe = np.clip(xf, 1.0, 2 * nxc - 1.0)
    ye = np.clip(yf, 1.0, 2 * nyc - 1.0)
    edge = a + bx * xe + by * ye
    assert np.max(np.abs(np.where(open_, p - edge, 0))) < 1e-12
    assert np.all(p[~open_] == 0)
    o.close()


def check_restriction(lib_path=None):
    nx, ny = 32, 24
    o, ax, ay, rhs, p0 = _mg_problem(nx, ny, 8, lib_path)
    lv = Lev(ax, ay, 1.0)
    rng = np.random.default_rng(9)
    blocks = rng.normal(size=(ny // 2, nx // 2))
    rhs = np.kron(blocks, np.ones((2, 2)))
    o.write64("rhs", rhs)
    o.write64("p", np.zeros((ny, nx)))          # residual = rhs in every open cell
    o.write64("p1", np.full((ny // 2, nx // 2), 7.0))
    o.run_kernel("restrictRes", 0)
    rc = o.read64("rhs1").reshape(ny // 2, nx // 2)
    open_ = lv.diag > 0
    nopen = open_.reshape(ny // 2, 2, nx // 2, 2).sum(axis=(1, 3))
    expect = blocks * nopen / 4.0                 # a closed cell contributes 0
    assert np.max(np.abs(rc - expect)) < 1e-13
    assert np.all(o.read64("p1") == 0)
    o.close()


# ------------------------------------------------------------------ radiation
def _piece(x, y, half, angle=0.0, Ts=(293.15,) * 4):
    p = np.zeros(20, np.float32)
    p[0], p[1], p[2], p[3] = x, y, angle, half
    p[7] = 1.0
    p[8:12] = Ts
    return p


def _ray_setup(pc, f, tgt, frame):
    """The rays of target `tgt` (face f of piece pc) as radiation.wgsl:75-99
    samples them: origins along the face, stratified jittered angles."""
    c, s = math.cos(float(pc[2])), math.sin(float(pc[2]))
    ln = [(1.0, 0.0), (0.0, 1.0), (-1.0, 0.0), (0.0, -1.0)][f]
    nrm = (c * ln[0] - s * ln[1], s * ln[0] + c * ln[1])
    tang = (-nrm[1], nrm[0])
    half = float(pc[3])
    jit = math.modf(frame * 0.7548777 + tgt * 0.5698403)[0]
    base = math.atan2(nrm[1], nrm[0])
    rays = []
    for li in range(64):
        u = (math.modf(li * 0.618034 + frame * 0.1234)[0] - 0.5) * 1.6 * half
        o = (float(pc[0]) + nrm[0] * (half + 1e-4) + tang[0] * u, float(pc[1]) + nrm[1] * (half + 1e-4) + tang[1] * u)
        a = ((li + jit) / 64 - 0.5) * PI_WGSL
        rays.append((o, (math.cos(base + a), math.sin(base + a)), PI_WGSL / 64 * math.cos(a)))
    return rays


def _march(o, d, pieces, tmax=4.0, ds=2e-5):
    """First piece hit along o + t d by point-in-box tests every ds, skipping
    the first 1e-5 (as the ray-box test does); (t, piece, face) or None."""
    t = np.arange(1e-5 + ds / 2, tmax, ds)
    x = o[0] + d[0] * t
    y = o[1] + d[1] * t
    best = None
    for n, pc in enumerate(pieces):
        c, s = math.cos(float(pc[2])), math.sin(float(pc[2]))
        qx, qy = x - float(pc[0]), y - float(pc[1])
        lx, ly = c * qx + s * qy, -s * qx + c * qy
        h = float(pc[3])
        inside = np.nonzero((np.abs(lx) <= h) & (np.abs(ly) <= h))[0]
        if inside.size == 0:
            continue
        k = inside[0]
        if best is not None and t[k] >= best[0]:
            continue
        # the face crossed: the side whose plane the entry point is nearest
        gx, gy = h - abs(lx[k]), h - abs(ly[k])
        face = (0 if lx[k] > 0 else 2) if gx < gy else (1 if ly[k] > 0 else 3)
        assert abs(gx - gy) > 1e-3, "ray passes too close to a corner for a brute-force test"
        best = (t[k], n, face)
    return best


def _trace_T(o, d, pieces, Tamb):
    """Wall temperature a ray sees: a piece face, or one specular floor
    bounce and then a piece face, or the ambient."""
    tfloor = -o[1] / d[1] if d[1] < -1e-6 else 1e9
    hit = _march(o, d, pieces, tmax=min(4.0, tfloor))
    if hit is not None:
        return float(pieces[hit[1]][8 + hit[2]]), "hit"
    if tfloor < 1e8:
        o2 = (o[0] + d[0] * tfloor, 1e-5)
        d2 = (d[0], -d[1])
        hit = _march(o2, d2, pieces)
        if hit is not None:
            return float(pieces[hit[1]][8 + hit[2]]), "bounce-hit"
        return Tamb, "bounce"
    return Tamb, "escape"


def _radiation_case(pieces, tgt, frame, lib_path):
    nx = ny = 128
    dx = 0.005
    o = Oracle(nx, ny, dx, 1e-3, fp64=True, lib_path=lib_path)
    o.reset()
    buf = np.zeros(1280, np.float32)
    for n, p in enumerate(pieces):
        buf[20 * n:20 * n + 20] = p
    o.upload_pieces(buf, len(pieces))
    o.set_param("frame", frame)
    o.write_params()
    o.radiation()   # T = Tamb, Y = (0, YO, 0, 0): no absorber, J = 0
    q = float(o.read64("rad").reshape(-1, 4)[tgt, 0])
    o.close()
    pc = pieces[tgt // 4]
    sig, Tamb = f32(5.670374e-8), f32(293.15)
    expect, kinds = 0.0, {}
    for (orig, d, w) in _ray_setup(pc, tgt % 4, tgt, frame):
        T, kind = _trace_T(orig, d, pieces, Tamb)
        kinds[kind] = kinds.get(kind, 0) + 1
        expect += w * (0.5 * sig * T ** 4)
    return q, expect, kinds


def check_radiation_piece_hits(lib_path=None):
    A = _piece(0.32, 0.10, 0.01, 0.0, (600, 610, 620, 630))
    B = _piece(0.35, 0.30, 0.06, 0.3, (900, 1000, 1100, 1200))
    q, expect, kinds = _radiation_case([A, B], 1, 7, lib_path)   # A's top face
    assert kinds.get("hit", 0) >= 8 and kinds.get("escape", 0) >= 8, kinds
    assert abs(q / expect - 1) < 1e-6, (q, expect, kinds)


def check_radiation_floor_bounce(lib_path=None):
    A = _piece(0.20, 0.35, 0.01, 0.0, (700, 710, 720, 730))
    B = _piece(0.55, 0.35, 0.05, 0.2, (900, 1000, 1100, 1200))
    q, expect, kinds = _radiation_case([A, B], 3, 5, lib_path)   # A's bottom face
    assert kinds.get("bounce-hit", 0) >= 4 and kinds.get("bounce", 0) >= 8, kinds
    assert abs(q / expect - 1) < 1e-6, (q, expect, kinds)


# ------------------------------------------------------------------ forces
def _mu(T, viscMult=1.0):
    return 1.716e-5 * (T / 273.15) ** 1.5 * (273.15 + 110.4) / (T + 110.4) * viscMult


def _forces_oracle(nx, ny, dx, dt, T, U2, lib_path):
    o = Oracle(nx, ny, dx, dt, fp64=True, lib_path=lib_path)
    o.set_param("exchAlpha", 0.0)   # no entrainment drag: exchange() = 1
    o.reset()
    o.upload_pieces(np.zeros(1280, np.float32), 0)
    o.write_params()
    o.clear_geometry()
    o.geometry()                     # no pieces: every face open, the floor row of fy closed
    o.write64("T", T)
    o.write64("U2", U2)
    o.write64("V2", np.zeros((ny + 1) * nx))
    o.run_kernel("forces")
    U = o.read64("U").reshape(ny, nx + 1)
    o.close()
    return U


def check_viscous_sine_mode(lib_path=None):
    nx, ny, dx, dt = 64, 48, 0.002, 0.004
    Tamb, rhoAmb = f32(293.15), f32(101325.0 * 0.0289 / (8.314462 * 293.15))
    mx, my = 29, 17
    theta = (my + 0.5) * math.pi / ny
    i = np.arange(nx + 1)[None, :]
    j = np.arange(ny)[:, None]
    U2 = np.sin(math.pi * mx * i / nx) * np.sin(theta * (j + 0.5))
    U = _forces_oracle(nx, ny, dx, dt, np.full(nx * ny, Tamb), U2, lib_path)
    dx, dt = f32(dx), f32(dt)  # the parameter block's values
    lam = (2 * math.cos(math.pi * mx / nx) - 2) + (2 * math.cos(theta) - 2)
    nu = _mu(Tamb) / (rhoAmb * Tamb / Tamb)
    factor = 1 + dt * nu * lam / (dx * dx)
    assert factor < 0.99   # the mode visibly decays
    inner = (slice(None), slice(1, nx))
    err = np.max(np.abs(U[inner] - factor * U2[inner]))
    assert err < 1e-12, err


def check_viscous_corner_viscosity(lib_path=None):
    nx, ny, dx, dt = 32, 40, 0.002, 0.004
    Tamb, rhoAmb = f32(293.15), f32(101325.0 * 0.0289 / (8.314462 * 293.15))
    Tj = 300.0 + 1200.0 * np.arange(ny) / (ny - 1)
    T = np.repeat(Tj[:, None], nx, axis=1)
    f = np.sin(math.pi * (np.arange(ny) + 0.5) / ny) + 0.3 * np.cos(0.37 * np.arange(ny))
    U2 = np.repeat(f[:, None], nx + 1, axis=1)
    U = _forces_oracle(nx, ny, dx, dt, T.reshape(-1), U2, lib_path)
    dx, dt = f32(dx), f32(dt)  # the parameter block's values
    for jj in range(1, ny - 1):
        rho = rhoAmb * Tamb / (0.5 * (Tj[jj] + Tj[jj]))
        muN = _mu(0.25 * (Tj[jj] + Tj[jj] + Tj[jj + 1] + Tj[jj + 1]))
        muS = _mu(0.25 * (Tj[jj] + Tj[jj] + Tj[jj - 1] + Tj[jj - 1]))
        expect = f[jj] + dt * (muN * (f[jj + 1] - f[jj]) + muS * (f[jj - 1] - f[jj])) / (dx * dx) / rho
        got = U[jj, 1:nx]
        assert np.max(np.abs(got - expect)) < 1e-11 * max(1.0, abs(expect)), (jj, got[0], expect)


# ------------------------------------------------------------------ face fractions
def check_face_fractions(lib_path=None):
    nx = ny = 128
    dx = 0.005
    xc, yc, h = f32(0.3203), f32(0.2107), f32(0.05)
    o = Oracle(nx, ny, dx, 1e-3, fp64=True, lib_path=lib_path)
    o.reset()
    buf = np.zeros(1280, np.float32)
    buf[:20] = _piece(xc, yc, h)
    o.upload_pieces(buf, 1)
    o.write_params()
    o.clear_geometry()
    o.geometry()
    fx = o.read64("fx").reshape(ny, nx + 1)
    fy = o.read64("fy").reshape(ny + 1, nx)
    o.close()
    dx = f32(dx)  # the parameter block's dx
    n = 0
    for j in range(ny):
        y = (j + 0.5) * dx
        if abs(y - yc) > h - 2 * dx:
            continue
        for i in range(nx + 1):
            x = i * dx
            d = abs(x - xc) - h     # the nearest feature is a vertical face
            if abs(d) > 3 * dx:
                continue
            expect = min(max(0.5 + d / dx, 0.0), 1.0)
            assert abs(fx[j, i] - expect) < 1e-12, (i, j, fx[j, i], expect)
            n += 1
    for j in range(1, ny + 1):
        y = j * dx
        for i in range(nx):
            x = (i + 0.5) * dx
            if abs(x - xc) > h - 2 * dx:
                continue
            d = abs(y - yc) - h
            if abs(d) > 3 * dx:
                continue
            expect = min(max(0.5 + d / dx, 0.0), 1.0)
            assert abs(fy[j, i] - expect) < 1e-12, (i, j, fy[j, i], expect)
            n += 1
    assert n > 100
    assert np.all(fy[0] == 0)
    # partially open faces exist (the box is off the grid)
    assert np.any((fx > 0.05) & (fx < 0.95))


# ------------------------------------------------------------------ tests
CHECKS = {
    "vcycle_matches_dense": check_vcycle_matches_dense,
    "vcycle_last_colour": check_vcycle_last_colour,
    "half_sweep": check_half_sweep,
    "prolongation": check_prolongation,
    "restriction": check_restriction,
    "radiation_piece_hits": check_radiation_piece_hits,
    "radiation_floor_bounce": check_radiation_floor_bounce,
    "viscous_sine_mode": check_viscous_sine_mode,
    "viscous_corner_viscosity": check_viscous_corner_viscosity,
    "face_fractions": check_face_fractions,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_pin(name):
    CHECKS[name]()


def test_pow15_within_one_ulp_of_the_power():
    """The oracle (and the kernels) evaluate pow(x, 1.5) of the Sutherland
    closures as x*sqrt(x) in FP32; over T/273.15 for T in 137-4370 K this is
    within one ulp of the correctly rounded power (WGSL allows pow a larger
    error), so it is the reference's power to within the reference's own
    precision."""
    lo, hi = np.float32(0.5), np.float32(16.0)
    bits = np.arange(lo.view(np.int32), hi.view(np.int32), 3, dtype=np.int32)
    x = bits.view(np.float32)
    mine = x * np.sqrt(x)
    exact = (x.astype(np.float64) ** 1.5).astype(np.float32)
    ulps = np.abs(mine.view(np.int32).astype(np.int64) - exact.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1


# ------------------------------------------------------------------ the mutants
# mutant -> the checks that must catch it (oracle/cf_oracle.c, MUTANT(k))
MUTANTS = {
    1: ("V-cycle up leg sweeps colour 0 first", ["vcycle_matches_dense", "vcycle_last_colour"]),
    2: ("V-cycle down leg sweeps colour 1 first", ["vcycle_matches_dense"]),
    3: ("prolongation weights 3/16 and 1/16 swapped", ["prolongation", "vcycle_matches_dense"]),
    4: ("prolongation mirrors instead of clamping at the edges", ["prolongation"]),
    5: ("restriction weight 1/2 instead of 1/4", ["restriction", "vcycle_matches_dense"]),
    6: ("coarse p not zeroed by the restriction", ["restriction"]),
    7: ("viscous term with 2/dx^2", ["viscous_sine_mode"]),
    8: ("north corner viscosity from the face's own two cells", ["viscous_corner_viscosity"]),
    9: ("no-slip floor ghost with the wrong sign", ["viscous_sine_mode"]),
    10: ("open fraction without the 1/2", ["face_fractions"]),
    11: ("no floor bounce", ["radiation_floor_bounce"]),
    12: ("ray hits take the opposite face's temperature", ["radiation_piece_hits"]),
}


def _build_mutant(k, outdir):
    cflags = ["-O1", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-std=c11", "-w",
              "-DCFO_MUTANT=%d" % k]
    objs = []
    for tag, extra, src in (("32", [], "cf_oracle.c"), ("64", ["-DCFO_DOUBLE"], "cf_oracle.c"),
                            ("f", [], "flame1d.c"), ("f32", ["-DFL_FLOAT"], "flame1d.c")):
        obj = os.path.join(outdir, "m%d_%s.o" % (k, tag))
        subprocess.run(["gcc"] + cflags + extra + ["-c", os.path.join(ORACLE_DIR, src), "-o", obj], check=True)
        objs.append(obj)
    so = os.path.join(outdir, "liboracle_mutant%d.so" % k)
    subprocess.run(["gcc", "-shared", "-fopenmp", "-o", so] + objs + ["-lm"], check=True)
    return so


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("oracle_mutants"))
    with cf.ThreadPoolExecutor(max_workers=6) as ex:
        futs = {k: ex.submit(_build_mutant, k, out) for k in MUTANTS}
        return {k: f.result() for k, f in futs.items()}


@pytest.mark.parametrize("k", sorted(MUTANTS))
def test_mutant_is_caught(k, mutant_libs):
    what, checks = MUTANTS[k]
    for name in checks:
        with pytest.raises(AssertionError):
            CHECKS[name](mutant_libs[k])
