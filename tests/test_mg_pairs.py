"""CPU check of the halo arithmetic behind the two-level multigrid legs
(MgDown2Op, MgUp2Op in campfire_amd/csrc/cf_mg.hip): one workgroup runs the
legs of a level and of the next coarser one for a tile of the coarser level.

Down pair: the coarse leg needs its tile plus 2nu+1 coarse cells; the coarse
rhs there is the restriction of the fine residual under those cells.  The
fine leg's region is that residual region plus 2nu-1 fine cells, two less
than a lone down leg needs: the residual of the residual region's two outer
fine rings is wrong, but it only reaches the coarse region's outer ring,
whose cells the coarse leg never updates (a coarse region edge never lies on
the domain edge: the tile origin is even, the halo odd, the level even).
Up pair: the fine leg needs its tile plus 2nu fine cells, whose prolongation
reads coarse cells up to nu+1 beyond the coarse tile, so the coarse leg's
region is the tile plus 3nu+1 coarse cells.

numpy emulation of the tile algorithm (colour-ordered updates; region-edge
cells with an in-domain neighbour outside are never updated) against global
red-black sweeps, restriction and prolongation (shaders/pressure.wgsl:77-129,
js/fluid.js:294-312): equal to the last bit, and one cell less of either
halo is not."""
import numpy as np
import pytest


def _level(nx, ny, rng):
    ax = rng.uniform(0, 1, (ny, nx + 1))
    ay = rng.uniform(0, 1, (ny + 1, nx))
    ay[0] = 0
    # a closed block (diag = 0)
    bx, by = nx // 3, ny // 3
    ax[by:by + 3, bx:bx + 4] = 0
    ay[by:by + 4, bx:bx + 3] = 0
    return ax, ay


def _terms(ax, ay, x, y, nx, ny):
    w, e, s, n = ax[y, x], ax[y, x + 1], ay[y, x], ay[y + 1, x]
    diag = 0.0
    diag += w if x > 0 else 2 * w
    diag += e if x < nx - 1 else 2 * e
    diag += s if y > 0 else 0.0
    diag += n if y < ny - 1 else 2 * n
    return w, e, s, n, diag


def _off(ax, ay, get, x, y, nx, ny):
    w, e, s, n, diag = _terms(ax, ay, x, y, nx, ny)
    off = 0.0
    if x > 0: off += w * get(x - 1, y)
    if x < nx - 1: off += e * get(x + 1, y)
    if y > 0: off += s * get(x, y - 1)
    if y < ny - 1: off += n * get(x, y + 1)
    return off, diag


def _update(ax, ay, rhs, h2, get, x, y, nx, ny):
    off, diag = _off(ax, ay, get, x, y, nx, ny)
    return (off - rhs[y, x] * h2) / diag if diag > 0 else 0.0


def _resid(ax, ay, rhs, h2, get, x, y, nx, ny):
    off, diag = _off(ax, ay, get, x, y, nx, ny)
    return 0.0 if diag <= 0 else rhs[y, x] - (off - diag * get(x, y)) / h2


def _prolong(pc, x, y):
    nyc, nxc = pc.shape
    I, J = x // 2, y // 2
    I2 = min(max(I + (1 if x & 1 else -1), 0), nxc - 1)
    J2 = min(max(J + (1 if y & 1 else -1), 0), nyc - 1)
    return 0.5625 * pc[J, I] + 0.1875 * (pc[J, I2] + pc[J2, I]) + 0.0625 * pc[J2, I2]


# ---------------------------------------------------------------- global legs
def _sweeps(ax, ay, rhs, h2, p, cols):
    ny, nx = p.shape
    p = p.copy()
    for col in cols:
        for y in range(ny):
            for x in range(nx):
                if (x + y) % 2 == col:
                    p[y, x] = _update(ax, ay, rhs, h2, lambda a, b: p[b, a], x, y, nx, ny)
    return p


def _restrict(r):
    return 0.25 * (r[0::2, 0::2] + r[0::2, 1::2] + r[1::2, 0::2] + r[1::2, 1::2])


def _global_down(F, C, nu):
    pf = _sweeps(F["ax"], F["ay"], F["rhs"], F["h2"], F["p"], [0, 1] * nu)
    ny, nx = pf.shape
    rf = np.array([[_resid(F["ax"], F["ay"], F["rhs"], F["h2"], lambda a, b: pf[b, a], x, y, nx, ny)
                    for x in range(nx)] for y in range(ny)])
    rc = _restrict(rf)
    pc = _sweeps(C["ax"], C["ay"], rc, C["h2"], np.zeros_like(rc), [0, 1] * nu)
    nyc, nxc = pc.shape
    rcr = np.array([[_resid(C["ax"], C["ay"], rc, C["h2"], lambda a, b: pc[b, a], x, y, nxc, nyc)
                     for x in range(nxc)] for y in range(nyc)])
    return pf, pc, rc, _restrict(rcr)


def _global_up(F, C, qf, qc, pn, nu):
    nyc, nxc = qc.shape
    vc = np.array([[qc[y, x] + (_prolong(pn, x, y) if _terms(C["ax"], C["ay"], x, y, nxc, nyc)[4] > 0 else 0.0)
                    for x in range(nxc)] for y in range(nyc)])
    pc = _sweeps(C["ax"], C["ay"], C["rhs"], C["h2"], vc, [1, 0] * nu)
    ny, nx = qf.shape
    vf = np.array([[qf[y, x] + (_prolong(pc, x, y) if _terms(F["ax"], F["ay"], x, y, nx, ny)[4] > 0 else 0.0)
                    for x in range(nx)] for y in range(ny)])
    return _sweeps(F["ax"], F["ay"], F["rhs"], F["h2"], vf, [1, 0] * nu), pc


# ---------------------------------------------------------------- tiled pairs
def _region_sweeps(L, rhs_get, sp, x0, y0, R, cols):
    """Colour-ordered updates of the region [x0, x0+R)^2 held in sp (dict of
    in-domain cells); cells on the region edge with an in-domain neighbour
    outside the region are not updated."""
    ny, nx = L["nx_ny"][1], L["nx_ny"][0]
    for col in cols:
        new = {}
        for (x, y) in sp:
            rx, ry = x - x0, y - y0
            edge = (rx == 0 and x > 0) or (rx == R - 1 and x < nx - 1) or \
                   (ry == 0 and y > 0) or (ry == R - 1 and y < ny - 1)
            if (x + y) % 2 == col and not edge:
                off, diag = _off(L["ax"], L["ay"], lambda a, b: sp[(a, b)], x, y, nx, ny)
                new[(x, y)] = (off - rhs_get(x, y) * L["h2"]) / diag if diag > 0 else 0.0
        sp.update(new)


def _cells(x0, y0, R, nx, ny):
    return [(x, y) for y in range(max(0, y0), min(ny, y0 + R)) for x in range(max(0, x0), min(nx, x0 + R))]


def _tiled_down(F, C, nu, TC, HF=None, HC=None):
    HF = 2 * nu - 1 if HF is None else HF
    HC = 2 * nu + 1 if HC is None else HC
    ny, nx = F["p"].shape
    nyc, nxc = ny // 2, nx // 2
    qf, qc, rc = np.full((ny, nx), np.nan), np.full((nyc, nxc), np.nan), np.full((nyc, nxc), np.nan)
    rn = np.full((nyc // 2, nxc // 2), np.nan)
    F["nx_ny"], C["nx_ny"] = (nx, ny), (nxc, nyc)
    for Y0 in range(0, nyc, TC):
        for X0 in range(0, nxc, TC):
            RC = TC + 2 * HC
            xc0, yc0 = X0 - HC, Y0 - HC
            xr0, yr0 = 2 * xc0, 2 * yc0
            xf0, yf0, RF = xr0 - HF, yr0 - HF, 2 * RC + 2 * HF
            sp = {c: F["p"][c[1], c[0]] for c in _cells(xf0, yf0, RF, nx, ny)}
            _region_sweeps(F, lambda x, y: F["rhs"][y, x], sp, xf0, yf0, RF, [0, 1] * nu)
            res = {c: _resid(F["ax"], F["ay"], F["rhs"], F["h2"], lambda a, b: sp[(a, b)], c[0], c[1], nx, ny)
                   for c in _cells(xr0, yr0, 2 * RC, nx, ny)}
            for (x, y) in _cells(2 * X0, 2 * Y0, 2 * TC, nx, ny):
                qf[y, x] = sp[(x, y)]
            crhs = {(X, Y): 0.25 * (res[(2 * X, 2 * Y)] + res[(2 * X + 1, 2 * Y)] + res[(2 * X, 2 * Y + 1)] +
                                    res[(2 * X + 1, 2 * Y + 1)]) for (X, Y) in _cells(xc0, yc0, RC, nxc, nyc)}
            spc = {c: 0.0 for c in crhs}
            _region_sweeps(C, lambda x, y: crhs[(x, y)], spc, xc0, yc0, RC, [0, 1] * nu)
            cres = {}
            for (X, Y) in _cells(X0, Y0, TC, nxc, nyc):
                qc[Y, X] = spc[(X, Y)]
                rc[Y, X] = crhs[(X, Y)]
                off, diag = _off(C["ax"], C["ay"], lambda a, b: spc[(a, b)], X, Y, nxc, nyc)
                cres[(X, Y)] = 0.0 if diag <= 0 else crhs[(X, Y)] - (off - diag * spc[(X, Y)]) / C["h2"]
            for (I, J) in _cells(X0 // 2, Y0 // 2, TC // 2, nxc // 2, nyc // 2):
                rn[J, I] = 0.25 * (cres[(2 * I, 2 * J)] + cres[(2 * I + 1, 2 * J)] + cres[(2 * I, 2 * J + 1)] +
                                   cres[(2 * I + 1, 2 * J + 1)])
    return qf, qc, rc, rn


def _tiled_up(F, C, qf, qc, pn, nu, TC, HF=None, HC=None):
    HF = 2 * nu if HF is None else HF
    HC = 3 * nu + 1 if HC is None else HC
    ny, nx = qf.shape
    nyc, nxc = qc.shape
    F["nx_ny"], C["nx_ny"] = (nx, ny), (nxc, nyc)
    pf, pc = np.full((ny, nx), np.nan), np.full((nyc, nxc), np.nan)
    for Y0 in range(0, nyc, TC):
        for X0 in range(0, nxc, TC):
            RC = TC + 2 * HC
            xc0, yc0 = X0 - HC, Y0 - HC
            spc = {}
            for (X, Y) in _cells(xc0, yc0, RC, nxc, nyc):
                d = _terms(C["ax"], C["ay"], X, Y, nxc, nyc)[4]
                spc[(X, Y)] = qc[Y, X] + (_prolong(pn, X, Y) if d > 0 else 0.0)
            _region_sweeps(C, lambda x, y: C["rhs"][y, x], spc, xc0, yc0, RC, [1, 0] * nu)
            for (X, Y) in _cells(X0, Y0, TC, nxc, nyc):
                pc[Y, X] = spc[(X, Y)]
            # the coarse values the prolongation reads, from the region (NaN outside it)
            lc = np.full((nyc, nxc), np.nan)
            for (X, Y), v in spc.items():
                lc[Y, X] = v
            xf0, yf0, RF = 2 * X0 - HF, 2 * Y0 - HF, 2 * TC + 2 * HF
            sp = {}
            for (x, y) in _cells(xf0, yf0, RF, nx, ny):
                d = _terms(F["ax"], F["ay"], x, y, nx, ny)[4]
                sp[(x, y)] = qf[y, x] + (_prolong(lc, x, y) if d > 0 else 0.0)
            _region_sweeps(F, lambda x, y: F["rhs"][y, x], sp, xf0, yf0, RF, [1, 0] * nu)
            for (x, y) in _cells(2 * X0, 2 * Y0, 2 * TC, nx, ny):
                pf[y, x] = sp[(x, y)]
    return pf, pc


def _problem(nx, ny, seed):
    rng = np.random.default_rng(seed)
    fax, fay = _level(nx, ny, rng)
    cax, cay = _level(nx // 2, ny // 2, rng)
    F = {"ax": fax, "ay": fay, "rhs": rng.uniform(-1, 1, (ny, nx)), "p": rng.uniform(-1, 1, (ny, nx)), "h2": 0.37}
    C = {"ax": cax, "ay": cay, "rhs": rng.uniform(-1, 1, (ny // 2, nx // 2)), "h2": 1.48}
    return F, C, rng


# 44 x 36 fine (22 x 18 coarse, 11 x 9 next): partial tiles of 4 and 8 coarse cells
@pytest.mark.parametrize("nu,TC", [(1, 4), (2, 4), (2, 8)])
def test_down_pair_halos(nu, TC):
    F, C, _ = _problem(44, 36, seed=nu)
    g = _global_down(F, C, nu)
    t = _tiled_down(F, C, nu, TC)
    for a, b in zip(g, t):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("which", ["fine", "coarse"])
def test_down_pair_one_cell_less_differs(which):
    nu, TC = 2, 4
    F, C, _ = _problem(44, 36, seed=5)
    g = _global_down(F, C, nu)
    kw = {"HF": 2 * nu - 2} if which == "fine" else {"HC": 2 * nu}
    t = _tiled_down(F, C, nu, TC, **kw)
    assert not all(np.array_equal(a, b) for a, b in zip(g, t))


@pytest.mark.parametrize("nu,TC", [(1, 4), (2, 4), (2, 8)])
def test_up_pair_halos(nu, TC):
    F, C, rng = _problem(44, 36, seed=10 + nu)
    qf = rng.uniform(-1, 1, (36, 44))
    qc = rng.uniform(-1, 1, (18, 22))
    pn = rng.uniform(-1, 1, (9, 11))
    g = _global_up(F, C, qf, qc, pn, nu)
    t = _tiled_up(F, C, qf, qc, pn, nu, TC)
    for a, b in zip(g, t):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("which", ["fine", "coarse"])
def test_up_pair_one_cell_less_differs(which):
    nu, TC = 2, 4
    F, C, rng = _problem(44, 36, seed=20)
    qf = rng.uniform(-1, 1, (36, 44))
    qc = rng.uniform(-1, 1, (18, 22))
    pn = rng.uniform(-1, 1, (9, 11))
    g = _global_up(F, C, qf, qc, pn, nu)
    kw = {"HF": 2 * nu - 1} if which == "fine" else {"HC": 3 * nu}
    t = _tiled_up(F, C, qf, qc, pn, nu, TC, **kw)
    assert not all(np.array_equal(a, b) for a, b in zip(g, t))
