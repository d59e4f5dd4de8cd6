"""GPU tests of the host C++ API (DArray, GranularParticles, Solver,
GranularSystem) through its C-ABI: container semantics, the reference's
system semantics (one step in the constructor, step() returns 1, frame
times, N changes through merge/split) and bit-exact parity of whole system
steps (granular, fine and adaptive) with the oracle."""
import sys

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [0, 1, 1023, 1024, 1025, 70001])
def test_darray_selftest(gpu, n):
    from granular_gpu_amd import host
    _log(f"DArray self-test n={n}")
    assert host.darray_selftest(n) == "ok"


@pytest.mark.parametrize("n", [0, 1, 257, 5000])
def test_solver_stage_methods(gpu, n):
    # add_external_force, update_particle_positions, apply_mass_scaling,
    # update_neighborhood and trigger_explosion called as Solver methods
    from granular_gpu_amd import host
    _log(f"Solver stage self-test n={n}")
    assert host.solver_stages_selftest(n) == "ok"


def _unsorted_keys(keys_sorted, perm):
    k = np.empty_like(keys_sorted)
    k[perm] = keys_sorted
    return k


def test_system_plain_with_fine_bitexact(gpu, tmp_path):
    from granular_gpu_amd import host
    s = scenes.plain_box(10, 1.0, dt=0.002, g=-9.8)
    fine = scenes.fine_cloud(s.pos, 6, 0.01, seed=5)
    _log(f"host system plain n={s.n} fine={fine.shape[0]}")
    sysm = host.HostSystem.from_scene(s, fine=fine)
    assert sysm.size("granular") == s.n and sysm.size("fine") == fine.shape[0]
    assert sysm.size("total") == s.n + s.boundary.shape[0] + fine.shape[0]
    k = 7
    assert sysm.step(k) == 1
    assert sysm.frame_times().shape == (k + 1,)     # the constructor ran one step
    assert (sysm.frame_times() > 0).all()
    gp, gv = sysm.read("granular", "pos"), sysm.read("granular", "vel")
    fp, fv = sysm.read("fine", "pos"), sysm.read("fine", "vel")
    # oracle: the same pipeline, step by step
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    cp, cv, cm = s.pos, s.vel, np.ones(s.n, np.float32)
    op, ov = fine, np.zeros_like(fine)
    for _ in range(k + 1):
        _, _, cs = po.grid_build(cp, s.cell_length, s.cells)
        cp, cv, cm = po.plain_steps_f32(cp, cv, cm, bpos, bcs, p, 1)[:3]
        keys, perm, _ = po.grid_build(op, s.cell_length, s.cells)
        keys = _unsorted_keys(keys, perm)
        op, ov = op[perm], ov[perm]
        op, ov = po.upsample_plain(op, ov, cp, cv, cs, bpos, bcs, p)
    assert np.array_equal(gp, cp) and np.array_equal(gv, cv)
    assert np.array_equal(fp, op) and np.array_equal(fv, ov)
    assert np.array_equal(sysm.read("fine", "particle_2_cell"), keys)
    assert np.array_equal(sysm.read("boundary", "pos"), bpos)
    assert np.array_equal(sysm.read("granular", "mass"), np.ones(s.n, np.float32))
    sysm.save_stats(tmp_path / "stats.csv")
    assert (tmp_path / "stats.csv").read_text().splitlines() == [
        "Frame,MergedParticles,SplitParticles,TotalParticles"]
    sysm.close()


def test_system_solver_options(gpu):
    # iteration counts and friction set on the system's solver apply from the
    # next step on (the constructor's step used the reference's 3 + 5, no
    # friction)
    from granular_gpu_amd import host
    s = scenes.plain_box(8, 1.0, dt=0.002, g=-9.8)
    a = host.HostSystem.from_scene(s)
    a.set_solver(3, 10, 0.5, 0.4)
    a.step(3)
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    cp, cv, cm = po.plain_steps_f32(s.pos, s.vel, s.mass, bpos, bcs, p, 1)[:3]
    p2 = po.params_from(s, 3, 10, 0.5, 0.4)
    cp, cv, cm = po.plain_steps_f32(cp, cv, cm, bpos, bcs, p2, 3)[:3]
    assert np.array_equal(a.read("granular", "pos"), cp)
    assert np.array_equal(a.read("granular", "vel"), cv)
    a.close()


def test_system_adaptive_bitexact(gpu, tmp_path):
    from granular_gpu_amd import host
    s = scenes.plain_box(16, 1.0, dt=0.002, g=-9.8)
    _log(f"host system adaptive n={s.n}")
    sysm = host.HostSystem.from_scene(s, is_adaptive=1)
    k = 79
    sysm.step(k)
    n = sysm.size("granular")
    st = po.AdaptiveState(s.pos, s.vel, s.mass, cap=2 * s.n)
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    stats = po.adaptive_steps(st, bpos, bcs, p, k + 1)
    _log(f"merged {stats[:, 0].sum()} split {stats[:, 1].sum()} final n {st.n} (gpu {n})")
    assert stats[:, 0].sum() > 0, "no merge happened: the test scene is too small"
    assert n == st.n < s.n
    assert np.array_equal(sysm.particle_stats(), stats)
    assert np.array_equal(sysm.read("granular", "pos"), st.pos[:n])
    assert np.array_equal(sysm.read("granular", "vel"), st.vel[:n])
    assert np.array_equal(sysm.read("granular", "mass"), st.mass[:n])
    assert np.array_equal(sysm.read("granular", "adaptive_last_step"), st.adl[:n])
    sysm.save_stats(tmp_path / "stats.csv")
    rows = (tmp_path / "stats.csv").read_text().splitlines()
    assert rows[0] == "Frame,MergedParticles,SplitParticles,TotalParticles"
    assert rows[1:] == [f"{i},{a},{b},{c}" for i, (a, b, c) in enumerate(stats)]
    sysm.close()


def test_system_moving_boundary(gpu):
    # a kinematic boundary: the animated wall particles move between steps
    # and the next step re-sorts the boundary (is_move_boundary)
    from granular_gpu_amd import host
    s = scenes.plain_box(8, 1.0, dt=0.002, g=-9.8)
    anim = np.zeros(s.boundary.shape[0], np.int32)
    anim[s.boundary[:, 1] < 0.02] = 1            # the floor
    sysm = host.HostSystem.from_scene(s, is_move_boundary=True, boundary_animated=anim)
    sysm.step(2)
    gp0, gv0 = sysm.read("granular", "pos"), sysm.read("granular", "vel")
    b0 = sysm.read("boundary", "pos")
    a0 = sysm.read("boundary", "is_animated")
    assert sysm.translate_animated((0.0, 0.03, 0.0)) == int(anim.sum())
    sysm.step(1)
    moved = b0.copy()
    moved[a0 == 1] += np.float32([0.0, 0.03, 0.0])
    bpos, bcs = po.sorted_boundary(moved, s.cell_length, s.cells)
    assert np.array_equal(sysm.read("boundary", "pos"), bpos)
    p = po.params_from(s, 3, 5)
    cp, cv, _ = po.plain_steps_f32(gp0, gv0, np.ones(s.n, np.float32), bpos, bcs, p, 1)[:3]
    assert np.array_equal(sysm.read("granular", "pos"), cp)
    assert np.array_equal(sysm.read("granular", "vel"), cv)
    sysm.close()


def test_boundary_animation_matches_oracle(gpu, tmp_path):
    # animateParticles and the updateAnimation timeline on the device against
    # the oracle on the same arrays, bit for bit; then the moving-boundary
    # re-sort of the next step
    from granular_gpu_amd import host
    _log("boundary animation: gg_boundary_animate + update_animation")
    s = scenes.plain_box(8, 1.0, dt=0.002, g=-9.8)
    anim = np.zeros(s.boundary.shape[0], np.int32)
    anim[s.boundary[:, 2] < 0.02] = 1            # the back wall
    sysm = host.HostSystem.from_scene(s, is_move_boundary=True, boundary_animated=anim)
    b = sysm.read("boundary", "pos")
    a = sysm.read("boundary", "is_animated")
    assert int(a.sum()) == int(anim.sum())
    sysm.animate((0.01, 0.0, -0.02), -33.0, (0.5, 0.4, 0.0))
    b = po.animate(b, a, (0.01, 0.0, -0.02), -33.0, (0.5, 0.4))
    assert np.array_equal(sysm.read("boundary", "pos"), b)
    st = host.animation_default()
    st.is_animating, st.start_index, st.num_particles = 1, 40, 300
    ost = po.animation_state(is_animating=1, start_index=40, num_particles=300)
    dt = 0.05
    calls = 0
    while ost.is_animating:
        assert sysm.update_animation(st, dt) == po.update_animation(ost, b, a, dt)
        calls += 1
        assert calls < 100
    assert st.is_animating == 0 and st.translation_complete == 1
    assert st.rotation_angle == ost.rotation_angle
    assert list(st.rotation_center) == list(ost.rotation_center)
    assert np.array_equal(sysm.read("boundary", "pos"), b)
    sysm.step(1)
    bpos, _ = po.sorted_boundary(b, s.cell_length, s.cells)
    assert np.array_equal(sysm.read("boundary", "pos"), bpos)
    assert np.isfinite(sysm.read("granular", "pos")).all()
    # the animated particles as a VTK file, in set order, y and z swapped
    marks = sysm.read("boundary", "is_animated")
    n = sysm.save_boundary_vtk(tmp_path / "bnd", 3)
    assert n == int((marks == 1).sum())
    lines = (tmp_path / "bnd" / "3.vtk").read_text().splitlines()
    assert lines[1] == "Animated Boundary Particles" and lines[4] == f"POINTS {n} float"
    xzy = np.array([l.split() for l in lines[5:5 + n]], np.float64)
    assert np.allclose(xzy, bpos[marks == 1][:, [0, 2, 1]], rtol=1e-5, atol=1e-6)
    sysm.close()


def test_interactive_extras_match_oracle(gpu):
    # explosion impulse, sphere picking and the clamped move, against the
    # oracle on the system's own arrays
    from granular_gpu_amd import host
    _log("interactive extras: gg_explosion, gg_pick_sphere, gg_move_picked")
    s = scenes.plain_box(12, 1.0, dt=0.002, g=-9.8)
    sysm = host.HostSystem.from_scene(s)
    sysm.step(3)
    pos, vel = sysm.read("granular", "pos"), sysm.read("granular", "vel")
    mass = sysm.read("granular", "mass")
    com = sysm.explode(50.0)
    ocom = po.center_of_mass(pos, mass)
    assert (np.abs(com - ocom) <= np.spacing(np.abs(ocom))).all()
    assert np.array_equal(sysm.read("granular", "vel"), po.explode(pos, vel, mass, com, 50.0))
    c, r = pos[100] + np.float32([0.01, 0.0, -0.01]), 0.05
    idx = sysm.pick(c, r)
    assert idx.size > 10 and np.array_equal(idx, po.pick_sphere(pos, c, r))
    assert sysm.move_picked(idx, (0.2, 0.9, -0.05)) == idx.size
    assert np.array_equal(sysm.read("granular", "pos"),
                          po.move_picked(pos, idx, (0.2, 0.9, -0.05)))
    with pytest.raises(host.HostError, match="out of range"):
        sysm.move_picked([0, s.n], (0.0, 0.0, 0.0))
    assert sysm.pick(c, -1.0).size == 0
    sysm.step(2)
    assert np.isfinite(sysm.read("granular", "pos")).all()
    sysm.close()


def test_motion_and_explosion_at_full_size(gpu):
    # a 400k-particle cone boundary (the funnel's size) rotated and shifted,
    # and the explosion impulse on 1M particles, bit for bit with the oracle
    from granular_gpu_amd import host
    s = scenes.plain_box(100, 2.4, dt=0.002, g=-9.8)
    cone = scenes.cone_funnel(400000, 30.0, 0.1, 0.6, (1.2, 0.9, 1.2))
    boundary = np.concatenate([s.boundary, cone]).astype(np.float32)
    anim = np.zeros(boundary.shape[0], np.int32)
    anim[s.boundary.shape[0]:] = 1
    _log(f"full size: n={s.n} boundary={boundary.shape[0]}")
    sysm = host.HostSystem(s.pos, boundary, np.zeros((0, 3), np.float32), s.space,
                           s.cell_length, s.dt, s.gravity, s.cells, s.density,
                           is_move_boundary=True, boundary_animated=anim)
    b, a = sysm.read("boundary", "pos"), sysm.read("boundary", "is_animated")
    assert int(a.sum()) == 400000
    sysm.animate((0.001, -0.002, 0.0), -1.5, (1.2, 0.9, 0.0))
    assert np.array_equal(sysm.read("boundary", "pos"),
                          po.animate(b, a, (0.001, -0.002, 0.0), -1.5, (1.2, 0.9)))
    pos, vel = sysm.read("granular", "pos"), sysm.read("granular", "vel")
    mass = sysm.read("granular", "mass")
    com = sysm.explode(50.0)
    assert (np.abs(com - po.center_of_mass(pos, mass)) <= np.spacing(np.abs(com))).all()
    assert np.array_equal(sysm.read("granular", "vel"), po.explode(pos, vel, mass, com, 50.0))
    sysm.close()


def test_system_outputs(gpu, tmp_path):
    # VTK files and the checkpoint carry the system's own arrays
    from granular_gpu_amd import host
    _log("outputs: VTK and checkpoint")
    s = scenes.plain_box(12, 1.0)
    fine = scenes.fine_cloud(s.pos, 4, 0.005, seed=3)
    sysm = host.HostSystem.from_scene(s, fine=fine, is_adaptive=1)
    sysm.step(5)
    fp = sysm.read("fine", "pos")
    n = sysm.save_upsampled_vtk(tmp_path / "out", 7, is_adaptive=True)
    assert n == int(np.isfinite(fp).all(axis=1).sum())
    lines = (tmp_path / "out_adaptive" / "7.vtk").read_text().splitlines()
    assert lines[4] == f"POINTS {n} float"
    xzy = np.array([l.split() for l in lines[5:5 + n]], np.float64)
    ok = fp[np.isfinite(fp).all(axis=1)].astype(np.float64)
    assert np.allclose(xzy, ok[:, [0, 2, 1]], rtol=1e-5, atol=1e-6)
    mass = sysm.read("granular", "mass")
    counts = sysm.save_mass_vtk(tmp_path / "out", 7)
    assert counts.tolist() == [int((mass < 2).sum()), int(((mass >= 2) & (mass <= 3)).sum()),
                               int((mass > 3).sum())]
    sysm.save_checkpoint(tmp_path / "s.ggck")
    ck = host.read_checkpoint(tmp_path / "s.ggck")
    for which, field in ck:
        assert ck[(which, field)].tobytes() == sysm.read(which, field).tobytes(), (which, field)
