"""The oracle's kinematic boundary motion (animateParticles and
updateAnimation, main.cu:1175-1277), pinned by closed forms: exact float
translation, rigid rotation against float64 geometry with a known sign,
untouched unmarked particles, and the updateAnimation timeline (translation
while time <= 2.5 s, centre recomputed once, rotation until 45 degrees with
the last increment not applied).  CPU only."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


def _points(n=300, seed=11):
    rng = np.random.default_rng(seed)
    pos = rng.random((n, 3)).astype(np.float32)
    animated = (rng.random(n) < 0.4).astype(np.int32)
    animated[:3] = [1, 0, 2]  # only the value 1 marks a particle
    return pos, animated


def test_rotation_cs_known_values():
    assert po.rotation_cs(0.0) == (f32(1.0), f32(0.0))
    c, s = po.rotation_cs(90.0)
    assert s == f32(1.0) and c == f32(np.cos(np.float64(f32(np.pi / 2))))
    for deg in (-0.18, -0.9, 30.0, -45.0):
        rad = np.float64(f32(np.float64(deg) * np.pi / 180.0))
        c, s = po.rotation_cs(deg)
        assert abs(float(c) - math.cos(rad)) <= 6e-8 and abs(float(s) - math.sin(rad)) <= 6e-8


def test_translation_is_exact_float_addition():
    pos, animated = _points()
    t = (f32(0.013), f32(-0.25), f32(1e-3))
    got = po.animate(pos, animated, t)
    m = animated == 1
    assert np.array_equal(got[m], pos[m] + np.float32(t))
    assert np.array_equal(got[~m], pos[~m])


def test_rotation_is_rigid_about_z():
    pos, animated = _points()
    cx, cy, deg = 0.4, 0.55, -27.5
    got = po.animate(pos, animated, (0.0, 0.0, 0.0), deg, (cx, cy))
    m = animated == 1
    a = math.radians(deg)
    u, v = pos[m, 0].astype(np.float64) - cx, pos[m, 1].astype(np.float64) - cy
    want_x = u * math.cos(a) - v * math.sin(a) + cx
    want_y = u * math.sin(a) + v * math.cos(a) + cy
    assert np.abs(got[m, 0] - want_x).max() < 3e-7
    assert np.abs(got[m, 1] - want_y).max() < 3e-7
    assert np.array_equal(got[m, 2], pos[m, 2])
    assert np.array_equal(got[~m], pos[~m])


def test_positive_angle_turns_counter_clockwise():
    pos = np.float32([[1.0, 0.0, 0.5], [0.0, 2.0, 0.0]])
    got = po.animate(pos, np.int32([1, 1]), (0.0, 0.0, 0.0), 90.0, (0.0, 0.0))
    assert np.allclose(got, [[0.0, 1.0, 0.5], [-2.0, 0.0, 0.0]], atol=2e-7)


def test_translate_then_rotate_order():
    # the translation is applied before the rotation about the fixed centre
    pos = np.float32([[1.0, 0.0, 0.0]])
    got = po.animate(pos, np.int32([1]), (1.0, 0.0, 0.0), 90.0, (0.0, 0.0))
    assert np.allclose(got, [[0.0, 2.0, 0.0]], atol=3e-7)


def test_update_animation_timeline():
    pos, _ = _points(200, seed=2)
    animated = np.zeros(200, np.int32)
    animated[150:] = 1
    st = po.animation_state(is_animating=1, start_index=150, num_particles=50)
    start = pos.copy()
    dt = f32(0.01)
    n_translate, n_rotate = 0, 0
    for _ in range(400):
        moved = po.update_animation(st, pos, animated, dt)
        if not st.translation_complete:
            n_translate += moved
        elif moved:
            n_rotate += 1
    # translation steps: animation_time <= 2.5 s, about 250 at dt = 0.01
    assert abs(n_translate - 250) <= 1
    # rotation: 0.9 degrees per call until the angle reaches 45; the call that
    # reaches it moves nothing
    assert st.is_animating == 0 and st.translation_complete == 1
    assert st.rotation_angle >= 45.0 and st.rotation_angle < 45.0 + 0.9 + 1e-4
    applied = st.rotation_angle - 0.9
    assert n_rotate == round(applied / 0.9)
    # geometry: rigid rotation of the translated cloud by -applied degrees
    # about the centroid recomputed when the translation ended
    g = np.array(st.rotation_center, np.float64)
    tr = start[150:].astype(np.float64)
    tr[:, 0] += -0.5 * 0.01 * n_translate
    assert np.allclose(g, tr.mean(0), atol=2e-4)
    a = math.radians(-applied)
    u, v = tr[:, 0] - g[0], tr[:, 1] - g[1]
    want = np.stack([u * math.cos(a) - v * math.sin(a) + g[0],
                     u * math.sin(a) + v * math.cos(a) + g[1], tr[:, 2]], 1)
    assert np.abs(pos[150:] - want).max() < 5e-5
    assert np.array_equal(pos[:150], start[:150])
    # finished: further calls move nothing
    frozen = pos.copy()
    assert not po.update_animation(st, pos, animated, dt)
    assert np.array_equal(pos, frozen)


def test_explosion_closed_forms():
    # two equal masses at +-a on x: centre 0, each pushed out by
    # F / (1 + a^2) / m; a particle on the centre is left alone
    a, F, m = 0.3, 50.0, 2.0
    pos = np.float32([[-a, 0.0, 0.0], [a, 0.0, 0.0], [0.0, 0.0, 0.0]])
    mass = np.float32([m, m, 1.0])
    com = po.center_of_mass(pos, mass)
    assert np.array_equal(com, np.zeros(3, np.float32))
    vel = po.explode(pos, np.zeros_like(pos), mass, com, F)
    dv = F / (1 + a * a) / m
    assert np.allclose(vel[0], [-dv, 0, 0], rtol=1e-6) and np.allclose(vel[1], [dv, 0, 0], rtol=1e-6)
    assert np.array_equal(vel[2], [0, 0, 0])


def test_explosion_radial_and_momentum():
    rng = np.random.default_rng(7)
    pos = rng.random((500, 3)).astype(np.float32)
    mass = (1.0 + 3.0 * rng.random(500)).astype(np.float32)
    com = po.center_of_mass(pos, mass)
    want = (pos.astype(np.float64) * mass[:, None]).sum(0) / mass.astype(np.float64).sum()
    assert np.abs(com - want).max() < 1e-7
    v0 = rng.standard_normal((500, 3)).astype(np.float32)
    vel = po.explode(pos, v0, mass, com, 50.0)
    dv = (vel - v0).astype(np.float64)
    d = pos.astype(np.float64) - com
    r = np.linalg.norm(d, axis=1)
    # along the radius, magnitude F / (1 + r^2) / m
    assert np.abs(np.cross(dv, d)).max() / np.abs(dv).max() < 1e-5
    assert np.allclose(np.linalg.norm(dv, axis=1), 50.0 / (1 + r * r) / mass, rtol=1e-5)
    # mirror-symmetric pairs about the centre: total momentum change is zero
    sym = np.concatenate([d[:50], -d[:50]]) + com
    ms = np.concatenate([mass[:50], mass[:50]])
    dvs = po.explode(sym.astype(np.float32), np.zeros((100, 3), np.float32), ms, com, 50.0)
    p = (dvs.astype(np.float64) * ms[:, None]).sum(0)
    assert np.abs(p).max() < 1e-3 * np.abs(dvs * ms[:, None]).sum(0).max()


def test_pick_sphere_brute_force():
    rng = np.random.default_rng(3)
    pos = rng.random((3000, 3)).astype(np.float32)
    c, r = np.float32([0.4, 0.5, 0.6]), 0.2
    got = po.pick_sphere(pos, c, r)
    d = np.linalg.norm(pos.astype(np.float64) - c.astype(np.float64), axis=1)
    sure_in = np.nonzero(d < r - 1e-6)[0]
    sure_out = np.nonzero(d > r + 1e-6)[0]
    assert np.all(np.diff(got) > 0)
    assert set(sure_in) <= set(got) and not (set(sure_out) & set(got))
    assert po.pick_sphere(pos, c, 0.0).size == 0


def test_move_picked_clamps():
    pos = np.float32([[0.5, 0.5, 0.5], [0.98, 0.02, 0.5], [0.3, 0.3, 0.3]])
    got = po.move_picked(pos, [0, 1], (0.05, -0.05, 0.0))
    assert np.allclose(got[0], [0.55, 0.45, 0.5])
    assert np.array_equal(got[1], np.float32([0.99, 0.01, 0.5]))
    assert np.array_equal(got[2], pos[2])


def test_update_animation_scripted_arc():
    # a scripted move: 0.5 m along x in 1 s, then 90 degrees at 45 deg/s
    pos, _ = _points(64, seed=4)
    animated = np.ones(64, np.int32)
    st = po.animation_state(is_animating=1, start_index=0, num_particles=64,
                            translation_direction=(1.0, 0.0, 0.0), translation_speed=0.5,
                            translation_time=1.0, rotation_speed=45.0, rotation_limit=90.0)
    dt = f32(1.0 / 64)
    calls = 0
    while st.is_animating and calls < 1000:
        po.update_animation(st, pos, animated, dt)
        calls += 1
    assert st.is_animating == 0
    assert st.rotation_angle == pytest.approx(90.0, abs=45.0 * float(dt) + 1e-4)
    assert calls == pytest.approx(64 + 1 + 128, abs=2)
