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


def batch_box(n_side: int, seed: int = 1, dt: float = 0.002, g: float = -9.8) -> Scene:
    """A plain box scene sized for n_side^3 particles (r = 0.01): box edge
    max(0.024 n_side, 0.02 (n_side - 1) + 0.25), dt = 0.002 and g = -9.8
    unless given."""
    box = max(0.024 * n_side, 0.02 * (n_side - 1) + 0.25)
    return plain_box(n_side, float(np.float32(box)), seed=seed, dt=dt, g=g)


def _largest_remainder(weights, total: int) -> np.ndarray:
    """Integer shares of `total` proportional to `weights` (floors, then one
    more for the largest fractional parts; ties go to the lower index)."""
    w = np.asarray(weights, np.float64)
    share = total * w / w.sum()
    m = np.floor(share).astype(np.int64)
    rest = total - int(m.sum())
    order = np.lexsort((np.arange(w.size), -(share - m)))
    m[order[:rest]] += 1
    return m


def cone_funnel(count: int, half_angle_deg: float = 30.0, outlet: float = 0.1,
                height: float = 0.3, centre=(0.5, 0.3, 0.5)) -> np.ndarray:
    """`count` boundary particles on the lateral surface of a conical funnel
    with a vertical (y) axis: the outlet circle of diameter `outlet` centred
    at `centre`, the wall opening upward at half_angle_deg from the axis to
    `height` above the outlet.

    Rings at equal slant steps (their number chosen so that the step along
    the slant equals the spacing along a ring); ring k gets a share of
    `count` proportional to its radius (largest remainder), and odd rings
    are turned by half a spacing.  Formed in float64, rounded once."""
    th = math.radians(half_angle_deg)
    slant = height / math.cos(th)
    r0 = 0.5 * outlet
    r1 = r0 + slant * math.sin(th)
    nr = max(2, int(round(math.sqrt(count * slant / (math.pi * (r0 + r1))))))
    s = slant * np.arange(nr) / (nr - 1)
    rk = r0 + s * math.sin(th)
    m = _largest_remainder(rk, count)
    cx, cy, cz = (float(v) for v in centre)
    out = np.empty((count, 3), np.float64)
    k0 = 0
    for k in range(nr):
        j = np.arange(m[k])
        phi = 2.0 * math.pi * (j + 0.5 * (k % 2)) / max(int(m[k]), 1)
        sl = slice(k0, k0 + m[k])
        out[sl, 0] = cx + rk[k] * np.cos(phi)
        out[sl, 1] = cy + s[k] * math.cos(th)
        out[sl, 2] = cz + rk[k] * np.sin(phi)
        k0 += m[k]
    return out.astype(np.float32)


def funnel_pour(n: int = 8000, cone_count: int = 12000, r: float = 0.01, box: float = 2.0,
                seed: int = 2, dt: float = 0.002, g: float = -9.8,
                half_angle_deg: float = 30.0, outlet: float = 0.1,
                cell_length: float | None = None) -> Scene:
    """Funnel pour: a cone_funnel of `cone_count` particles (axis at the
    box centre in x and z, outlet at 0.3 box, height 0.3 box) inside BOX
    walls, and a column of n particles at seeded uniform-random positions
    in a cylinder above it: radius 0.8 of the cone's rim radius, from 0.02
    above the rim, tall enough for (2.4 r)^3 of volume per particle.
    cell_length defaults to the scenes' 0.0404."""
    L = CELL_LENGTH if cell_length is None else np.float32(cell_length)
    space = (box, box, box)
    h = 0.3 * box
    centre = (0.5 * box, 0.3 * box, 0.5 * box)
    cone = cone_funnel(cone_count, half_angle_deg, outlet, h, centre)
    rim = 0.5 * outlet + h * math.tan(math.radians(half_angle_deg))
    R = 0.8 * rim
    y0 = centre[1] + h + 0.02
    tall = n * (2.4 * r) ** 3 / (math.pi * R * R)
    if y0 + tall > 0.99 * box:
        raise ValueError("funnel_pour: the column does not fit in the box")
    u = uniform01(seed, 3 * n).reshape(n, 3)
    rad = R * np.sqrt(u[:, 0])
    phi = 2.0 * math.pi * u[:, 1]
    pos = np.stack([centre[0] + rad * np.cos(phi), y0 + tall * u[:, 2],
                    centre[2] + rad * np.sin(phi)], 1).astype(np.float32)
    walls = box_walls(space, L, "box")
    return Scene(
        name=f"funnel_pour_{n}_cone_{cone_count}", pos=pos, vel=np.zeros((n, 3), np.float32),
        mass=np.full(n, radius_mass(r), np.float32),
        boundary=np.concatenate([walls, cone]).astype(np.float32), space=space,
        cell_length=L, cells=cell_counts(space, L), dt=np.float32(dt),
        gravity=(0.0, float(np.float32(g)), 0.0), density=DENSITY_R001, radius=r)


def funnel_column(n: int = 1_000_000, cone_count: int = 400_000, r: float = 0.01,
                  cone_height: float = 2.5, outlet: float = 0.1, half_angle_deg: float = 30.0,
                  gap: float = 0.6, seed: int = 2, dt: float = 0.002,
                  g: float = -9.8) -> Scene:
    """Funnel pour with the space sized to the column: a cone_funnel of
    `cone_count` particles (outlet `gap` above the floor, `cone_height`
    tall) in BOX walls 0.1 wider than the rim on each side, and n particles
    of radius r at seeded uniform-random positions in a cylinder of 0.8 of
    the rim radius from 0.02 above the rim, (2.4 r)^3 of volume each; the
    space ends 0.1 above the column.

    n = 1M, 400k cone (defaults): space 3.187 x 6.303 x 3.187, 79 x 157 x 79
    cells of 0.0404 (979,843), a column 3.08 tall."""
    L = CELL_LENGTH
    rim = 0.5 * outlet + cone_height * math.tan(math.radians(half_angle_deg))
    R = 0.8 * rim
    tall = n * (2.4 * r) ** 3 / (math.pi * R * R)
    w = 2.0 * rim + 0.2
    y0 = gap + cone_height + 0.02
    space = (w, y0 + tall + 0.1, w)
    centre = (0.5 * w, gap, 0.5 * w)
    cone = cone_funnel(cone_count, half_angle_deg, outlet, cone_height, centre)
    u = uniform01(seed, 3 * n).reshape(n, 3)
    rad = R * np.sqrt(u[:, 0])
    phi = 2.0 * math.pi * u[:, 1]
    pos = np.stack([centre[0] + rad * np.cos(phi), y0 + tall * u[:, 2],
                    centre[2] + rad * np.sin(phi)], 1).astype(np.float32)
    walls = box_walls(space, L, "box")
    return Scene(
        name=f"funnel_column_{n}_cone_{cone_count}", pos=pos,
        vel=np.zeros((n, 3), np.float32), mass=np.full(n, radius_mass(r), np.float32),
        boundary=np.concatenate([walls, cone]).astype(np.float32), space=space,
        cell_length=L, cells=cell_counts(space, L), dt=np.float32(dt),
        gravity=(0.0, float(np.float32(g)), 0.0), density=DENSITY_R001, radius=r)


TEMPLATE8 = np.array([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)],
                     np.float64) - 0.5


def paired_boxes(n_side: int, box: float, r: float = 0.02, seed: int = 1,
                 dt: float = 0.002, g: float = -9.8) -> tuple[Scene, Scene]:
    """A coarse plain box (plain_box(n_side, box, seed, r)) and its fine
    twin seeded identically: every coarse particle i is replaced by the 8
    particles of radius r/2 at x_i + r T_k (T_k the corners of the unit cube
    centred at 0, k = 4 sx + 2 sy + sz), fine index 8 i + k, which tile the
    coarse particle's cube exactly.  Same walls, cells, dt and gravity; the
    two are stepped side by side to give the learned upsampler's training
    pairs."""
    coarse = plain_box(n_side, box, seed=seed, r=r, dt=dt, g=g)
    fpos = (coarse.pos.astype(np.float64)[:, None, :] + r * TEMPLATE8[None]).reshape(-1, 3)
    nf = fpos.shape[0]
    fine = Scene(name=f"{coarse.name}_fine", pos=fpos.astype(np.float32),
                 vel=np.zeros((nf, 3), np.float32),
                 mass=np.full(nf, radius_mass(0.5 * r), np.float32),
                 boundary=coarse.boundary, space=coarse.space, cell_length=coarse.cell_length,
                 cells=coarse.cells, dt=coarse.dt, gravity=coarse.gravity,
                 density=coarse.density, radius=0.5 * r)
    return coarse, fine


def disk_heap(nx: int = 16, ny: int = 40, r: float = 0.005, disk_radius: float = 0.12,
              disk_y: float = 0.1, box: float = 0.6, seed: int = 3, dt: float = 0.002,
              g: float = -9.8) -> Scene:
    """A fixed-base heap: an nx x ny x nx jittered lattice (spacing 2r,
    jitter U(+-0.1r)) of radius-r particles standing on a horizontal disk of
    boundary particles (radius disk_radius at height disk_y, a square grid
    of spacing r clipped to the disk), centred in a box of BOX walls with
    cells of 4.04 r.  The column collapses, the excess falls off the disk's
    rim and the heap left on the disk stands at the angle of repose."""
    L = np.float32(4.04 * r)
    space = (box, box, box)
    c = 0.5 * box
    k = int(math.floor(disk_radius / r))
    gx, gz = np.meshgrid(np.arange(-k, k + 1), np.arange(-k, k + 1), indexing="ij")
    keep = (gx * gx + gz * gz) * r * r <= disk_radius * disk_radius
    disk = np.stack([c + r * gx[keep], np.full(int(keep.sum()), disk_y), c + r * gz[keep]], 1)
    sp = 2.0 * r
    x0 = c - 0.5 * sp * (nx - 1)
    y0 = disk_y + 0.01 + r + 0.1 * r
    pos = jittered_lattice(nx, ny, nx, sp, (x0, y0, x0), 0.1 * r, seed)
    walls = box_walls(space, L, "box")
    n = pos.shape[0]
    return Scene(
        name=f"disk_heap_{n}", pos=pos, vel=np.zeros((n, 3), np.float32),
        mass=np.full(n, radius_mass(r), np.float32),
        boundary=np.concatenate([walls, disk]).astype(np.float32), space=space, cell_length=L,
        cells=cell_counts(space, L), dt=np.float32(dt),
        gravity=(0.0, float(np.float32(g)), 0.0), density=DENSITY_R001, radius=r)


def heap_angle(pos: np.ndarray, centre, base_y: float, base_radius: float,
               bins: int = 12) -> float:
    """Angle of repose (degrees) of the heap on a disk of base_radius whose
    top surface is at base_y, centred at centre = (x, z): the particles above
    the disk and inside its rim, binned by distance from the axis; the
    heap's surface height in a bin is the 90th percentile of the particle
    heights there, and the angle is that of the least-squares line through
    the bins between 0.2 and 0.9 of the radius."""
    p = np.asarray(pos, np.float64)
    rr = np.hypot(p[:, 0] - centre[0], p[:, 2] - centre[1])
    h = p[:, 1] - base_y
    on = (rr < base_radius) & (h > 0.0)
    rr, h = rr[on], h[on]
    edges = np.linspace(0.0, base_radius, bins + 1)
    rb, hb = [], []
    for a, b in zip(edges[:-1], edges[1:]):
        m = (rr >= a) & (rr < b)
        if m.sum() >= 5 and 0.2 * base_radius <= 0.5 * (a + b) <= 0.9 * base_radius:
            rb.append(0.5 * (a + b))
            hb.append(np.quantile(h[m], 0.9))
    if len(rb) < 3:
        raise ValueError("heap_angle: too few populated bins")
    slope = np.polyfit(np.array(rb), np.array(hb), 1)[0]
    return math.degrees(math.atan(-slope))


def bucket_scoop(count: int, depth: float = 0.2, width: float = 0.3, height: float = 0.15,
                 corner=(0.0, 0.0, 0.0)) -> np.ndarray:
    """About `count` boundary particles on an open scoop: a bottom plate
    (x in [0, depth], z in [0, width]), a back plate at x = depth and two
    side plates at z = 0 and z = width (y in [0, height]); the open mouth
    faces -x.  One square grid of spacing sqrt(area / count); each shared
    edge belongs to one plate only.  `corner` is the (x, y, z) of the
    scoop's origin."""
    area = depth * width + height * width + 2.0 * depth * height
    s = math.sqrt(area / count)
    nx = max(2, int(round(depth / s)) + 1)
    ny = max(2, int(round(height / s)) + 1)
    nz = max(2, int(round(width / s)) + 1)
    xs = np.linspace(0.0, depth, nx)
    ys = np.linspace(0.0, height, ny)
    zs = np.linspace(0.0, width, nz)
    pts = []
    X, Z = np.meshgrid(xs, zs, indexing="ij")                       # bottom
    pts.append(np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], 1))
    Y, Z = np.meshgrid(ys[1:], zs, indexing="ij")                   # back
    pts.append(np.stack([np.full(Y.size, depth), Y.ravel(), Z.ravel()], 1))
    X, Y = np.meshgrid(xs[:-1], ys[1:], indexing="ij")              # sides
    for z in (0.0, width):
        pts.append(np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], 1))
    out = np.concatenate(pts) + np.asarray(corner, np.float64)
    return out.astype(np.float32)


@dataclass
class DigScene:
    scene: Scene
    animated: np.ndarray   # (nb,) int32: 1 for the bucket
    script: dict           # animation-state fields of the scripted arc


def excavator_dig(bed=(0.6, 0.3), fill: float = 0.12, r: float = 0.01,
                  bucket_count: int = 3000, seed: int = 5, dt: float = 0.002,
                  g: float = -9.8) -> DigScene:
    """Excavator dig: a bed of particles on a 2r lattice (jitter 0.1 r,
    seeded) covering a bed[0] x bed[1] floor up to `fill`, in BOX walls of
    (bed[0], 4 fill, bed[1]), and a bucket_scoop (depth 0.3 bed[0], width
    0.6 bed[1], height fill) hanging 0.02 above the bed near the +x wall,
    mouth towards -x.  The scripted arc (animation fields in `script`):
    0.25 bed[0] (0.5 m for a 2 m bed) in 1 s, 15 degrees below -x into the
    bed, then 90 degrees about z at 45 degrees/s about the bucket's centroid,
    turning the mouth up.  The cell length and the wall spacing scale with
    r (4.04 r); masses follow radius_mass(r)."""
    bx, bz = bed
    space = (bx, 4.0 * fill, bz)
    spacing = 2.0 * r
    lo = 0.005 * np.asarray(space) + 0.01 + r
    nxl = int((bx - 2 * lo[0]) / spacing) + 1
    nzl = int((bz - 2 * lo[2]) / spacing) + 1
    nyl = max(1, int((fill - lo[1]) / spacing) + 1)
    pos = jittered_lattice(nxl, nyl, nzl, spacing, (lo[0], lo[1], lo[2]), 0.1 * r, seed)
    n = pos.shape[0]
    d, w, h = 0.3 * bx, 0.6 * bz, fill
    top = float(pos[:, 1].max()) + r
    bucket = bucket_scoop(bucket_count, d, w, h, (bx - d - 0.05 * bx, top + 0.02, 0.5 * (bz - w)))
    L = np.float32(float(CELL_LENGTH) * r / 0.01)
    walls = box_walls(space, L, "box")
    boundary = np.concatenate([walls, bucket]).astype(np.float32)
    animated = np.zeros(boundary.shape[0], np.int32)
    animated[walls.shape[0]:] = 1
    a = math.radians(15.0)
    script = dict(is_animating=1, start_index=walls.shape[0], num_particles=bucket.shape[0],
                  translation_direction=(-math.cos(a), -math.sin(a), 0.0),
                  translation_speed=0.25 * bx,
                  translation_time=1.0, rotation_speed=45.0, rotation_limit=90.0,
                  center_from_marked=1)
    sc = Scene(name=f"excavator_dig_{n}_bucket_{bucket.shape[0]}", pos=pos,
               vel=np.zeros((n, 3), np.float32), mass=np.full(n, radius_mass(r), np.float32),
               boundary=boundary, space=space, cell_length=L,
               cells=cell_counts(space, L), dt=np.float32(dt),
               gravity=(0.0, float(np.float32(g)), 0.0), density=DENSITY_R001, radius=r)
    return DigScene(sc, animated, script)
