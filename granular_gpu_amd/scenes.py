"""Synthetic scenes: seeded jittered lattices and the reference's wall recipes.

All generators are deterministic (SplitMix64, Steele/Lea/Flood 2014) and
produce float32 arrays bit-identically on any machine: values are formed in
float64 and rounded once, or formed in float32 exactly as the reference does.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

DENSITY_R001 = np.float32(238732.4146)  # m = 1 <=> r = 0.01 (main.cu:47)
CELL_LENGTH = np.float32(0.0404)        # scenes/*.json "cell_length"

_M64 = (1 << 64) - 1


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of SplitMix64 starting from state `seed` (uint64)."""
    out = np.empty(count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & _M64) + np.arange(1, count + 1, dtype=np.uint64) * np.uint64(
            0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        out[:] = z ^ (z >> np.uint64(31))
    return out


def uniform01(seed: int, count: int) -> np.ndarray:
    """Floats in [0,1) from the top 24 bits of SplitMix64 (float64, exact)."""
    return (splitmix64(seed, count) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))


def cell_counts(space, cell_length=CELL_LENGTH):
    """cell_size = ceil(space / L) per axis (main.cu:158-160), in float32."""
    L = np.float32(cell_length)
    return tuple(int(math.ceil(float(np.float32(s) / L))) for s in space)


def box_walls(space, cell_length=CELL_LENGTH, scene: str = "box", translation=(0, 0, 0)):
    """Wall particles of the reference BOX / PILING / CORNER / EXCAVATOR /
    FUNNEL / BOX_OPENING recipes (main.cu:822-1167), in the reference's
    insertion order, computed in float32 as the reference does."""
    f = np.float32
    S = np.asarray(space, dtype=np.float32)
    cc = np.array(cell_counts(space, cell_length), dtype=np.int64) * 2
    den = (cc - 1).astype(np.float32)
    out = []

    def put(i, j, k):
        x = (np.array([i, j, k], dtype=np.float32) / den) * S
        out.append(f(0.99) * x + f(0.005) * S)

    nx, ny, nz = (int(v) for v in cc)
    fb = scene in ("box", "excavator", "funnel", "box_opening", "corner")
    if fb:  # front and back
        for i in range(nx):
            for j in range(ny):
                put(i, j, 0)
                if scene != "corner":
                    put(i, j, nz - 1)
    # top and bottom
    for i in range(nx):
        for j in range(nz - 2):
            put(i, 0, j + 1)
            if scene in ("box", "piling", "box_opening"):
                put(i, ny - 1, j + 1)
    if fb:  # left and right
        for i in range(ny - 2):
            for j in range(nz - 2):
                put(0, i + 1, j + 1)
                if scene != "corner":
                    put(nx - 1, i + 1, j + 1)
    arr = np.asarray(out, dtype=np.float32).reshape(-1, 3)
    if scene not in ("excavator", "funnel"):
        arr = arr + np.asarray(translation, dtype=np.float32)
    return arr


def jittered_lattice(nx, ny, nz, spacing, origin, jitter, seed=1):
    """Lattice with uniform jitter U(-jitter, +jitter) per coordinate.

    Loop order y (outer), x, z (inner), as main.cu:751-760; one SplitMix64
    draw per coordinate in x, y, z order.  Positions are formed in float64
    and rounded once to float32."""
    iy, ix, iz = np.meshgrid(np.arange(ny), np.arange(nx), np.arange(nz), indexing="ij")
    ix, iy, iz = ix.ravel(), iy.ravel(), iz.ravel()
    n = ix.size
    u = uniform01(seed, 3 * n).reshape(n, 3)
    j = (2.0 * u - 1.0) * float(jitter)
    o = [float(v) for v in origin]
    s = float(spacing)
    pos = np.empty((n, 3), dtype=np.float64)
    pos[:, 0] = (o[0] + s * ix) + j[:, 0]
    pos[:, 1] = (o[1] + s * iy) + j[:, 1]
    pos[:, 2] = (o[2] + s * iz) + j[:, 2]
    return pos.astype(np.float32)


@dataclass
class Scene:
    name: str
    pos: np.ndarray          # (n,3) float32
    vel: np.ndarray          # (n,3) float32
    mass: np.ndarray         # (n,) float32
    boundary: np.ndarray     # (nb,3) float32
    space: tuple
    cell_length: np.float32
    cells: tuple
    dt: np.float32
    gravity: tuple
    density: np.float32
    radius: float

    @property
    def n(self) -> int:
        return self.pos.shape[0]


def radius_mass(r: float) -> np.float32:
    """Mass of a particle of radius r at the scenes' density (m = 1 at
    r = 0.01): m = (r / 0.01)^3, formed in float64 and rounded once."""
    return np.float32((float(r) / 0.01) ** 3)


def plain_box(n_side: int = 32, box: float = 1.0, seed: int = 1, r: float = 0.01,
              dt: float = 1.0 / 240.0, g: float = -9.81) -> Scene:
    """Plain box scene: n_side^3 particles of radius r on a 2r lattice
    (jitter U(+-0.1r), seeded), centred in x/z, bottoms 0.1 above the floor
    surface (floor particles at 0.005*box with radius 0.01), BOX walls.

    Masses follow the radius (radius_mass).
    n_side=32, box=1   -> 32,768 particles, 25^3 cells, 14,408 walls.
    n_side=100, box=2.4 -> 1,000,000 particles, 60^3 cells, 84,968 walls."""
    space = (box, box, box)
    spacing = 2.0 * r
    span = spacing * (n_side - 1)
    x0 = (box - span) / 2.0
    floor_top = 0.005 * box + 0.01
    y0 = floor_top + 0.1 + r
    pos = jittered_lattice(n_side, n_side, n_side, spacing, (x0, y0, x0), 0.1 * r, seed)
    n = pos.shape[0]
    walls = box_walls(space, CELL_LENGTH, "box")
    return Scene(
        name=f"plain_box_{n_side}^3_in_{box}m", pos=pos, vel=np.zeros((n, 3), np.float32),
        mass=np.full(n, radius_mass(r), np.float32), boundary=walls, space=space,
        cell_length=CELL_LENGTH,
        cells=cell_counts(space), dt=np.float32(dt), gravity=(0.0, float(np.float32(g)), 0.0),
        density=DENSITY_R001, radius=r)


def fine_cloud(coarse: np.ndarray, per: int, radius: float, seed: int = 7) -> np.ndarray:
    """`per` fine particles around each coarse centre, offsets uniform in
    [-radius, radius) per axis (the reference draws 25 per coarse particle
    the same way, main.cu:804-816, with its own generator).  Fine particles
    of one coarse particle are contiguous, in coarse order."""
    c = np.asarray(coarse, np.float32)
    n = c.shape[0]
    u = uniform01(seed, n * per * 3).reshape(n, per, 3)
    off = ((2.0 * u - 1.0) * float(radius)).astype(np.float32)
    return (c[:, None, :] + off).reshape(n * per, 3).astype(np.float32)


def batch_box_sizes(count: int = 64, nmin: int = 4096, nmax: int = 262144, seed: int = 4):
    """Per-scene particle counts log-uniform in [nmin, nmax] (SplitMix64,
    seeded), rounded to lattice sides, and per-scene friction (mu_s uniform
    in [0.3, 0.7], mu_k = 0.8 mu_s).  Returns [(n_side, mu_s, mu_k)]."""
    u = uniform01(seed, 2 * count).reshape(count, 2)
    lo, hi = math.log(nmin), math.log(nmax)
    out = []
    for a, b in u:
        n = math.exp(lo + (hi - lo) * float(a))
        side = max(2, int(round(n ** (1.0 / 3.0))))
        mu_s = 0.3 + 0.4 * float(b)
        out.append((side, mu_s, 0.8 * mu_s))
    return out


def batch_box(n_side: int, seed: int = 1) -> Scene:
    """A plain box scene sized for n_side^3 particles (r = 0.01): box edge
    max(0.024 n_side, 0.02 (n_side - 1) + 0.25), dt = 0.002, g = -9.8."""
    box = max(0.024 * n_side, 0.02 * (n_side - 1) + 0.25)
    return plain_box(n_side, float(np.float32(box)), seed=seed, dt=0.002, g=-9.8)
