"""Pins the plain-upsampler oracle (update_upsampled_cuda, Solver.cu:1244-1387)
with kinematic closed forms, exact special cases of the blend, the kernel's
support, the stencil truncation, the boundary repulsion's sign and size, the
clamps and an exact scaling invariant.  No GPU needed."""
import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

H = 0.045            # kernel support: r9 = 1/H^2
G_DT = np.float32(-9.8) * np.float32(0.002)


def _scene():
    return scenes.plain_box(2, 1.0)


def _run(fine, fvel, coarse=(), cvel=(), boundary=(), steps=1, u=None):
    s = _scene()
    p = po.params_from(s)
    fine = np.asarray(fine, np.float32).reshape(-1, 3)
    fvel = np.asarray(fvel, np.float32).reshape(-1, 3)
    coarse = np.asarray(coarse, np.float32).reshape(-1, 3)
    cvel = np.asarray(cvel, np.float32).reshape(-1, 3)
    if coarse.shape[0]:
        _, perm, cs = po.grid_build(coarse, s.cell_length, s.cells)
        coarse, cvel = coarse[perm], cvel[perm]
    else:
        cs = np.zeros(s.cells[0] * s.cells[1] * s.cells[2] + 2, np.int32)
    b = np.asarray(boundary, np.float32).reshape(-1, 3)
    if b.shape[0]:
        bpos, bcs = po.sorted_boundary(b, s.cell_length, s.cells)
    else:
        bpos, bcs = b, np.zeros_like(cs)
    return po.upsample_plain(fine, fvel, coarse, cvel, cs, bpos, bcs, p, u=u, steps=steps)


def _nog():
    u = po.upsample_params()
    u.gy = 0.0
    return u


def test_defaults_are_reference_constants():
    u = po.upsample_params()
    assert np.float32(u.r9) == np.float32(493.8271605)
    assert abs(1.0 / np.sqrt(u.r9) - H) < 1e-7
    assert (u.dt, u.gy, u.max_velocity, u.boundary_radius) == (
        np.float32(0.002), np.float32(-9.8), 7.0, np.float32(0.015))


def test_free_fall_matches_symplectic_euler():
    # no coarse, no boundary: v_k = k g dt, y_k = y0 + g dt^2 k(k+1)/2
    y0, k = 0.8, 40
    pos, vel = _run([[0.5, y0, 0.5]], [[0, 0, 0]], steps=k)
    g, dt = -9.8, 0.002
    assert abs(vel[0, 1] - k * g * dt) < 1e-5
    assert abs(pos[0, 1] - (y0 + g * dt * dt * k * (k + 1) / 2)) < 1e-5
    assert pos[0, 0] == np.float32(0.5) and pos[0, 2] == np.float32(0.5)
    # one step is exactly v = g*dt, x = x + v*dt
    pos, vel = _run([[0.5, y0, 0.5]], [[0, 0, 0]])
    assert vel[0, 1] == G_DT
    assert pos[0, 1] == np.float32(y0) + G_DT * np.float32(0.002)


def test_coincident_coarse_takes_its_velocity():
    # d = 0 -> w = 1 = max w -> alpha = 0 -> v = sum(w v)/sum(w) = v_c exactly
    vc = [0.25, -0.5, 1.5]
    pos, vel = _run([[0.5, 0.5, 0.5]], [[3, 3, 3]], [[0.5, 0.5, 0.5]], [vc])
    assert np.array_equal(vel[0], np.float32(vc))
    assert np.array_equal(pos[0], np.float32(0.5) + np.float32(vc) * np.float32(0.002))


@pytest.mark.parametrize("d", [0.0, 0.01, 0.02, 0.03, 0.04, 0.0449])
def test_single_coarse_kernel_shape(d):
    # g = 0, v0 = 0: v = (1 - alpha) v_c = max w * v_c with w = (1 - d^2/H^2)^3
    vc = np.float32([0.0, 0.0, 2.0])
    pos, vel = _run([[0.5, 0.5, 0.5]], [[0, 0, 0]], [[0.5 + d, 0.5, 0.5]], [vc], u=_nog())
    w = max(0.0, 1.0 - d * d / (H * H)) ** 3
    if w <= 1e-6:  # total weight below min_weight: the ballistic branch
        w = 0.0
    assert vel[0, 2] == pytest.approx(w * 2.0, rel=2e-5, abs=1e-7)
    assert vel[0, 0] == 0.0 and vel[0, 1] == 0.0


def test_kernel_support_ends_at_h():
    # just outside the support: no weight at all -> ballistic branch
    pos, vel = _run([[0.5, 0.5, 0.5]], [[0, 1, 0]], [[0.5 + 0.0451, 0.5, 0.5]], [[5, 5, 5]])
    assert vel[0, 1] == np.float32(1.0) + G_DT and vel[0, 0] == 0.0


def test_two_coarse_weighted_average():
    # two coarse at equal distance: sum(w v)/sum(w) is the plain mean
    d = 0.02
    c = [[0.5 + d, 0.5, 0.5], [0.5 - d, 0.5, 0.5]]
    v = [[1.0, 0.0, 0.0], [3.0, 0.0, 0.0]]
    pos, vel = _run([[0.5, 0.5, 0.5]], [[0, 0, 0]], c, v, u=_nog())
    w = (1.0 - d * d / (H * H)) ** 3
    assert vel[0, 0] == pytest.approx(w * 2.0, rel=2e-5)


def test_stencil_truncates_support():
    # H > L: a coarse particle within H but two cells away is not seen
    L = float(np.float32(0.0404))
    x = 12 * L + 0.039            # near the right edge of cell 12
    cx = x + 0.044                # cell 14, distance 0.044 < H
    assert int(np.floor(np.float32(cx) / np.float32(L))) == 14
    pos, vel = _run([[x, 0.5, 0.5]], [[0, 0, 0]], [[cx, 0.5, 0.5]], [[5, 5, 5]])
    assert vel[0, 0] == 0.0 and vel[0, 1] == G_DT
    # the same pair one cell closer is seen
    x2 = 12 * L + 0.041
    pos, vel = _run([[x2, 0.5, 0.5]], [[0, 0, 0]], [[x2 + 0.044, 0.5, 0.5]], [[5, 5, 5]])
    assert vel[0, 0] > 0.0


@pytest.mark.parametrize("y,ground", [(0.5, False), (0.08, True)])
def test_boundary_repulsion(y, ground):
    # one wall particle 0.01 below-left along -x: pushed towards +x with
    # 3 (0.015 - d)/0.015 * d/(d + 1e-6); no coarse -> near-wall damping 0.8
    d = 0.01
    pos, vel = _run([[0.5, y, 0.5]], [[0, 1.0, 0]], boundary=[[0.5 - d, y, 0.5]])
    force = 3.0 * (0.015 - d) / 0.015 * d / (d + 1e-6)
    assert vel[0, 0] == pytest.approx(force, rel=1e-5)
    vy = (1.0 - 9.8 * 0.002) * 0.8 * (0.5 if ground else 1.0)
    assert vel[0, 1] == pytest.approx(vy, rel=1e-6)
    # beyond the repulsion radius: no force, no damping
    pos, vel = _run([[0.5, y, 0.5]], [[0, 1.0, 0]], boundary=[[0.5 - 0.016, y, 0.5]])
    assert vel[0, 0] == 0.0 and vel[0, 1] == np.float32(1.0) + G_DT


def test_velocity_clamp_and_nan():
    pos, vel = _run([[0.5, 0.5, 0.5]], [[0, 0, 0]], [[0.5, 0.5, 0.5]], [[100, -100, 3]])
    assert np.array_equal(vel[0], np.float32([7, -7, 3]))
    # a NaN component meets the clamp first: fmin(NaN, 7) = 7 (IEEE minNum),
    # so it leaves as +7 and the reference's later NaN test never fires
    pos, vel = _run([[0.5, 0.5, 0.5]], [[0, 0, 0]], [[0.5, 0.5, 0.5]], [[np.nan, 0, 0]])
    assert np.array_equal(vel[0], np.float32([7, 0, 0]))
    assert pos[0, 0] == np.float32(0.5) + np.float32(7) * np.float32(0.002)


def test_floor_and_side_clamps():
    pos, vel = _run([[0.5, 0.041, 0.5]], [[1.0, -1.0, 2.0]])
    assert pos[0, 1] == np.float32(0.04) and vel[0, 1] == 0.0
    assert vel[0, 0] == np.float32(1.0) * np.float32(0.95)
    assert vel[0, 2] == np.float32(2.0) * np.float32(0.95)
    pos, vel = _run([[0.001, 0.5, 0.001]], [[-1.0, 0.0, -1.0]])
    assert pos[0, 0] == 0.0 and pos[0, 2] == 0.0 and vel[0, 0] == 0.0 and vel[0, 2] == 0.0


def test_velocity_scaling_is_exact():
    # with g = 0 and v0 = 0 the result is linear in the coarse velocities and
    # doubling them doubles it bit for bit (scaling by 2 is exact in float)
    rng = np.random.default_rng(5)
    coarse = (0.5 + rng.uniform(-0.06, 0.06, (40, 3))).astype(np.float32)
    cvel = rng.uniform(-1, 1, (40, 3)).astype(np.float32)
    fine = scenes.fine_cloud(coarse[:10], 4, 0.02)
    f0 = np.zeros_like(fine)
    _, v1 = _run(fine, f0, coarse, cvel, u=_nog())
    _, v2 = _run(fine, f0, coarse, 2 * cvel, u=_nog())
    assert np.abs(v1).max() > 0.05
    assert np.array_equal(v2, 2 * v1)


def test_fine_cloud_offsets_bounded():
    c = np.float32([[0.5, 0.5, 0.5], [0.2, 0.3, 0.4]])
    f = scenes.fine_cloud(c, 8, 0.01)
    assert f.shape == (16, 3)
    assert np.all(np.abs(f.reshape(2, 8, 3) - c[:, None]) <= 0.01 + 1e-7)
    assert np.array_equal(f, scenes.fine_cloud(c, 8, 0.01))
