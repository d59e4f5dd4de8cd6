"""Generated boundaries and scenes for the funnel pour and the excavator dig:
the cone sampler against the cone's equation and exact count, the bucket
sampler against its planes and spacing, and the scenes' placement rules.
CPU only."""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from granular_gpu_amd import scenes


@pytest.mark.parametrize("count", [500, 12000, 400000])
def test_cone_on_surface_with_exact_count(count):
    c0, h, th = (1.0, 0.6, 1.0), 0.6, math.radians(30.0)
    pts = scenes.cone_funnel(count, 30.0, 0.1, h, c0)
    assert pts.shape == (count, 3) and pts.dtype == np.float32
    p = pts.astype(np.float64)
    rho = np.hypot(p[:, 0] - c0[0], p[:, 2] - c0[2])
    y = p[:, 1] - c0[1]
    # radius = outlet/2 + (height above the outlet) tan(half angle)
    assert np.abs(rho - (0.05 + y * math.tan(th))).max() < 2e-6
    assert y.min() >= -1e-6 and y.max() <= h + 1e-6
    assert np.isclose(rho.min(), 0.05, atol=1e-6)
    assert np.isclose(rho.max(), 0.05 + h * math.tan(th), atol=2e-6)
    assert len(np.unique(pts, axis=0)) == count


def test_cone_spacing_is_even():
    count = 20000
    pts = scenes.cone_funnel(count, 30.0, 0.1, 0.6, (1.0, 0.6, 1.0)).astype(np.float64)
    th = math.radians(30.0)
    slant = 0.6 / math.cos(th)
    area = math.pi * (0.05 + 0.05 + slant * math.sin(th)) * slant
    s = math.sqrt(area / count)
    d, _ = cKDTree(pts).query(pts, k=2)
    nn = d[:, 1]
    # every particle has a neighbour about one spacing away: no holes, no
    # clumps (the boundary must not leak particles of radius ~ s / 2)
    assert nn.max() < 1.3 * s and np.median(nn) > 0.6 * s


def test_funnel_pour_column_above_cone():
    s = scenes.funnel_pour(8000, 12000, seed=2)
    assert s.n == 8000 and s.boundary.shape[0] == 12000 + scenes.box_walls(
        s.space, s.cell_length, "box").shape[0]
    box = s.space[0]
    rim_y = 0.6 * box
    rim = 0.05 + 0.3 * box * math.tan(math.radians(30.0))
    p = s.pos.astype(np.float64)
    rho = np.hypot(p[:, 0] - 0.5 * box, p[:, 2] - 0.5 * box)
    assert rho.max() <= 0.8 * rim + 1e-6
    assert p[:, 1].min() >= rim_y + 0.02 - 1e-6 and p[:, 1].max() < 0.99 * box
    # seeded: the same seed gives the same column, another seed another
    assert np.array_equal(s.pos, scenes.funnel_pour(8000, 12000, seed=2).pos)
    assert not np.array_equal(s.pos, scenes.funnel_pour(8000, 12000, seed=3).pos)
    with pytest.raises(ValueError):
        scenes.funnel_pour(10 ** 6, 12000, box=2.0)


def test_bucket_planes_and_spacing():
    d, w, h = 0.2, 0.3, 0.15
    b = scenes.bucket_scoop(3000, d, w, h, (1.0, 2.0, 3.0)).astype(np.float64)
    assert abs(b.shape[0] - 3000) < 0.1 * 3000
    assert len(np.unique(b, axis=0)) == b.shape[0]
    q = b - [1.0, 2.0, 3.0]
    on = (np.isclose(q[:, 1], 0.0, atol=1e-6) | np.isclose(q[:, 0], d, atol=1e-6) |
          np.isclose(q[:, 2], 0.0, atol=1e-6) | np.isclose(q[:, 2], w, atol=1e-6))
    assert on.all()
    assert (q[:, 0] >= -1e-6).all() and (q[:, 0] <= d + 1e-6).all()
    assert (q[:, 1] >= -1e-6).all() and (q[:, 1] <= h + 1e-6).all()
    # the mouth (x = 0, above the bottom) is open
    assert not (np.isclose(q[:, 0], 0.0, atol=1e-6) & (q[:, 1] > 1e-6) &
                (q[:, 2] > 1e-6) & (q[:, 2] < w - 1e-6)).any()
    s = math.sqrt((d * w + h * w + 2 * d * h) / 3000)
    nn = cKDTree(b).query(b, k=2)[0][:, 1]
    assert nn.max() < 1.2 * s and nn.min() > 0.8 * s


def test_excavator_dig_layout():
    dig = scenes.excavator_dig()
    s = dig.scene
    nb_walls = s.boundary.shape[0] - int(dig.animated.sum())
    assert (dig.animated[:nb_walls] == 0).all() and (dig.animated[nb_walls:] == 1).all()
    bucket = s.boundary[dig.animated == 1]
    # the bucket hangs above the bed, inside the box, and leaves by -x, down
    assert bucket[:, 1].min() > s.pos[:, 1].max()
    assert (bucket >= 0).all() and (bucket <= np.asarray(s.space)).all()
    assert (s.pos > 0).all() and (s.pos < np.asarray(s.space)).all()
    sc = dig.script
    assert sc["start_index"] == nb_walls and sc["num_particles"] == bucket.shape[0]
    t = np.asarray(sc["translation_direction"])
    assert np.isclose(np.linalg.norm(t), 1.0) and t[0] < 0 and t[1] < 0
    assert sc["translation_speed"] * sc["translation_time"] == pytest.approx(0.25 * s.space[0])
    assert sc["rotation_limit"] / sc["rotation_speed"] == pytest.approx(2.0)
    # at r = 0.004 the grid scales with the radius
    big = scenes.excavator_dig(bed=(0.4, 0.2), fill=0.05, r=0.004, bucket_count=2000).scene
    assert big.cell_length == np.float32(0.0404 * 0.4)
