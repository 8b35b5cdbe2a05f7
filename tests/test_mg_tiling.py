"""CPU check of the halo arithmetic behind the fused multigrid kernels
(campfire_amd/csrc/cf_mg.hip): a tile plus a halo of k cells, updating only
cells whose in-domain neighbours are inside the region, reproduces k global
red-black half-sweeps exactly (red-black Gauss-Seidel = colour-ordered
Jacobi); one cell less of halo does not.  numpy emulation of the tile
algorithm, run against global sweeps of shaders/pressure.wgsl:87-94."""
import numpy as np
import pytest


def _problem(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    ax = rng.uniform(0, 1, (ny, nx + 1))
    ay = rng.uniform(0, 1, (ny + 1, nx))
    ay[0] = 0
    ax[5:9, 10:14] = 0
    ay[5:9, 10:14] = 0
    return ax, ay, rng.uniform(-1, 1, (ny, nx)), rng.uniform(-1, 1, (ny, nx))


def _update(ax, ay, rhs, h2, get, x, y, nx, ny):
    w, e, s, n = ax[y, x], ax[y, x + 1], ay[y, x], ay[y + 1, x]
    off = diag = 0.0
    if x > 0: off += w * get(x - 1, y); diag += w
    else: diag += 2 * w
    if x < nx - 1: off += e * get(x + 1, y); diag += e
    else: diag += 2 * e
    if y > 0: off += s * get(x, y - 1); diag += s
    if y < ny - 1: off += n * get(x, y + 1); diag += n
    else: diag += 2 * n
    return (off - rhs[y, x] * h2) / diag if diag > 0 else 0.0


def _global(ax, ay, rhs, p, cols, h2=0.37):
    ny, nx = p.shape
    p = p.copy()
    for col in cols:
        for y in range(ny):
            for x in range(nx):
                if (x + y) % 2 == col:
                    p[y, x] = _update(ax, ay, rhs, h2, lambda a, b: p[b, a], x, y, nx, ny)
    return p


def _tiled(ax, ay, rhs, p, cols, H, TX=16, TY=16, h2=0.37):
    ny, nx = p.shape
    out = p.copy()
    for y0 in range(0, ny, TY):
        for x0 in range(0, nx, TX):
            RX, RY = TX + 2 * H, TY + 2 * H
            sp = {(x, y): p[y, x] for y in range(max(0, y0 - H), min(ny, y0 + TY + H))
                  for x in range(max(0, x0 - H), min(nx, x0 + TX + H))}
            for col in cols:
                new = {}
                for (x, y) in sp:
                    rx, ry = x - (x0 - H), y - (y0 - H)
                    edge = (rx == 0 and x > 0) or (rx == RX - 1 and x < nx - 1) or \
                           (ry == 0 and y > 0) or (ry == RY - 1 and y < ny - 1)
                    if (x + y) % 2 == col and not edge:
                        new[(x, y)] = _update(ax, ay, rhs, h2, lambda a, b: sp[(a, b)], x, y, nx, ny)
                sp.update(new)
            for y in range(y0, min(y0 + TY, ny)):
                for x in range(x0, min(x0 + TX, nx)):
                    out[y, x] = sp[(x, y)]
    return out


@pytest.mark.parametrize("nu", [1, 2])
def test_down_leg_halo(nu):
    ax, ay, rhs, p = _problem(38, 26)
    g = _global(ax, ay, rhs, p, [0, 1] * nu)
    assert np.array_equal(_tiled(ax, ay, rhs, p, [0, 1] * nu, 2 * nu + 1), g)


@pytest.mark.parametrize("nu", [1, 2])
def test_up_leg_halo_and_too_small(nu):
    ax, ay, rhs, p = _problem(38, 26, seed=1)
    g = _global(ax, ay, rhs, p, [1, 0] * nu)
    assert np.array_equal(_tiled(ax, ay, rhs, p, [1, 0] * nu, 2 * nu), g)
    assert not np.array_equal(_tiled(ax, ay, rhs, p, [1, 0] * nu, 2 * nu - 1), g)
