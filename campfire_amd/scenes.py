"""Seeded scene builder: piece piles, ignition discs and smooth random fields.

Every random choice comes from PCG32 (O'Neill, pcg-random.org, XSH-RR
variant) seeded explicitly, so a scene is reproducible from its seed alone.

Piece records follow the layout of js/main.js:195-199 / shaders/common.wgsl:17-23:
[x, y, angle, half, vx, vy, omega, alive, Ts[4], mflux[4], charCov[4]].
"""
import math

import numpy as np

MAX_PIECES, PIECE_FLOATS = 64, 20
STICK, LOG = 0.025, 0.05  # index.html data-size: sticks 2.5 cm, logs 5 cm


class PCG32:
    MULT = 6364136223846793005
    MASK = (1 << 64) - 1

    def __init__(self, seed, seq=54):
        self.state = 0
        self.inc = ((seq << 1) | 1) & self.MASK
        self.next_u32()
        self.state = (self.state + (seed & self.MASK)) & self.MASK
        self.next_u32()

    def next_u32(self):
        old = self.state
        self.state = (old * self.MULT + self.inc) & self.MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * (self.next_u32() / 4294967296.0)


def lit_piece(x, y, side, angle=0.0, Ts=880.0, mflux=0.008, char=1.0):
    """A burning piece with prescribed face state (the pre-lit pieces of
    js/main.js:144-147 hold Ts = 880 K, a char layer and pyrolysing faces)."""
    r = np.zeros(PIECE_FLOATS, np.float32)
    r[0:8] = [x, y, angle, side / 2, 0, 0, 0, 1]
    r[8:12] = Ts
    r[12:16] = mflux
    r[16:20] = char
    return r


def pile(n, seed, width, log_fraction=0.4, margin=0.05, spread=None):
    """n sticks and logs dropped at seeded x onto a height map, so pieces rest
    on the floor or on earlier pieces.  Returns (pieces[64*20], n)."""
    g = PCG32(seed)
    spread = spread if spread is not None else min(0.25, width / 3)
    cx = width / 2
    res = 0.0025
    nb = int(math.ceil(width / res)) + 1
    hmap = np.zeros(nb)
    out = np.zeros((MAX_PIECES, PIECE_FLOATS), np.float32)
    for q in range(n):
        side = LOG if g.uniform() < log_fraction else STICK
        x = g.uniform(cx - spread / 2, cx + spread / 2)
        x = min(max(x, margin + side / 2), width - margin - side / 2)
        a = int((x - side / 2) / res)
        b = int(math.ceil((x + side / 2) / res))
        base = float(hmap[max(a, 0):min(b, nb)].max()) if b > a else 0.0
        y = base + side / 2 + 0.0005
        hmap[max(a, 0):min(b, nb)] = y + side / 2
        ang = g.uniform(-0.15, 0.15)
        Ts = g.uniform(820, 950)
        out[q] = lit_piece(x, y, side, ang, Ts=Ts, mflux=g.uniform(0.004, 0.012))
    return out.reshape(-1), n


def stacked_logs(seed, width, count=3):
    """`count` 5 cm logs stacked on the floor at a seeded x (two on the floor,
    one on top for count = 3)."""
    g = PCG32(seed)
    cx = g.uniform(0.4 * width, 0.6 * width)
    out = np.zeros((MAX_PIECES, PIECE_FLOATS), np.float32)
    s = LOG
    pos = [(cx - 0.55 * s, s / 2), (cx + 0.55 * s, s / 2), (cx, 1.5 * s + 0.001)]
    for q in range(count):
        x, y = pos[q % 3]
        out[q] = lit_piece(x, y + (q // 3) * s, s, 0.0, Ts=g.uniform(850, 920), mflux=0.008)
    return out.reshape(-1), count


def disc(nx, ny, dx, cx, cy, r, T_in, T_out):
    """Cell temperatures with a hot disc (H.blob, dev/gas_test.html:112-120)."""
    x = (np.arange(nx) + 0.5) * dx - cx
    y = (np.arange(ny) + 0.5) * dx - cy
    X, Y = np.meshgrid(x, y)
    T = np.full((ny, nx), T_out, np.float32)
    T[X * X + Y * Y < r * r] = T_in
    return T.reshape(-1)


def blobs(nx, ny, seed, lo, hi, n=32, shape=None):
    """Smooth seeded field in [lo, hi]: a sum of n Gaussian blobs, rescaled."""
    g = PCG32(seed)
    shape = shape or (ny, nx)
    H, W = shape
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    f = np.zeros(shape)
    for _ in range(n):
        cx, cy = g.uniform(0, W), g.uniform(0, H)
        s = g.uniform(0.05, 0.25) * max(W, H)
        amp = g.uniform(-1, 1)
        f += amp * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    f -= f.min()
    m = f.max()
    if m > 0:
        f /= m
    return (lo + (hi - lo) * f).astype(np.float32)


def random_state(nx, ny, seed, vmax=3.0, hot=True):
    """Smooth fields in the ranges of a burning scene (T 293-2200 K,
    Y_F <= 0.3, Y_O <= 0.233, Y_S <= 1e-3, |u| <= vmax).  Returns a dict of
    arrays in the C-ABI layouts (Y as vec4 AoS)."""
    T = blobs(nx, ny, seed, 293.15, 2200.0 if hot else 400.0)
    F = blobs(nx, ny, seed + 1, 0.0, 0.3)
    O = blobs(nx, ny, seed + 2, 0.0, 0.233)
    S = blobs(nx, ny, seed + 3, 0.0, 1e-3)
    Pr = np.clip(1.0 - F - O - S, 0, 1) * 0.5
    Y = np.stack([F, O, Pr, S], axis=-1).astype(np.float32)
    U = blobs(nx + 1, ny, seed + 4, -vmax, vmax, shape=(ny, nx + 1))
    V = blobs(nx, ny + 1, seed + 5, -vmax, vmax, shape=(ny + 1, nx))
    return {"T": T.reshape(-1), "Y": Y.reshape(-1), "U": U.reshape(-1), "V": V.reshape(-1)}


def air_params(Tamb=300.0, YO=0.233):
    """Ambient-air parameters for a given temperature: rho from the ideal gas
    law with the reference's p0, W and R (js/config.js:15-18,46)."""
    return {"Tamb": Tamb, "rhoAmb": 101325.0 * 0.0289 / (8.314462 * Tamb), "YOamb": YO}


# ---- the benchmark workloads ------------------------------------------------------
def config_plain_256(seed=1):
    """256x256, h = 5 mm, 300 K air with Y_O = 0.233, three 5 cm logs stacked on
    the floor at a seeded x, a 1200 K ignition disc of r = 3 cm at the pile.
    Each step takes dt = 0.5 h / max(max|u|, 1 m/s) and runs V-cycles until
    ||r|| / ||rhs|| < 1e-6 (at most 60), with the level-0 pressure and
    residual in FP64 (defect correction): an FP32 pressure stalls near 1e-5."""
    nx = ny = 256
    h = 0.005
    pieces, n = stacked_logs(seed, nx * h)
    cx = float(pieces[0] + pieces[20]) / 2
    return {"nx": nx, "ny": ny, "dx": h, "dt": 0.5 * h / 1.0, "pieces": pieces, "n": n,
            "params": air_params(), "disc": (cx, 0.05, 0.03, 1200.0),
            "schedule": {"nu": 2, "vcycles": 60, "tol": 1e-6, "coarse_iters": 120, "cfl": 0.5, "umin": 1.0,
                         "refine": 1}}


def config_pile_1024(seed=1):
    """1024x1024, h = 1.25 mm, 300 K air, 12 sticks and logs in a seeded random
    pile, a 1200 K ignition disc of r = 3 cm at the pile."""
    nx = ny = 1024
    h = 0.00125
    pieces, n = pile(12, seed, nx * h)
    cx = float(np.mean(pieces.reshape(64, 20)[:n, 0]))
    return {"nx": nx, "ny": ny, "dx": h, "dt": 0.5 * h / 2.5, "pieces": pieces, "n": n,
            "params": air_params(), "disc": (cx, 0.06, 0.03, 1200.0)}


def apply(sim, cfg):
    """Load a workload into a FluidSim (or an oracle with the same methods)."""
    for k, v in cfg["params"].items():
        sim.set_param(k, v)
    if "schedule" in cfg:
        sim.set_schedule(**cfg["schedule"])
    sim.reset()
    Tamb = cfg["params"]["Tamb"]
    cx, cy, r, Tin = cfg["disc"]
    T = disc(cfg["nx"], cfg["ny"], cfg["dx"], cx, cy, r, Tin, Tamb)
    for name in ("T", "T2", "Th"):
        if hasattr(sim, "write_field"):
            sim.write_field(name, T)
        else:
            sim.write(name, T)
    sim.upload_pieces(cfg["pieces"], cfg["n"])
    sim.write_params()
    return sim


def config_batch(seed=1, n=64, width=1.28):
    """n scenes with sizes drawn by seed from {256^2, 512^2, 1024^2} over the
    same width (h = 5, 2.5, 1.25 mm), 2-12 seeded pieces each, a 1200 K
    ignition disc at each pile; dt = 0.5 h / 2.5 m/s."""
    g = PCG32(seed)
    out = []
    for q in range(n):
        nx = (256, 512, 1024)[min(int(g.uniform(0, 3)), 2)]
        h = width / nx
        npc = 2 + min(int(g.uniform(0, 11)), 10)
        pieces, k = pile(npc, seed * 1000 + q, width)
        cx = float(np.mean(pieces.reshape(64, 20)[:k, 0]))
        out.append({"nx": nx, "ny": nx, "dx": h, "dt": 0.5 * h / 2.5, "pieces": pieces, "n": k,
                    "params": air_params(), "disc": (cx, 0.06, 0.03, 1200.0)})
    return out

