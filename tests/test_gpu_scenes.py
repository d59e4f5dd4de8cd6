"""The funnel pour and the excavator dig at test size, through the host
system's C-ABI, against the oracle bit for bit: a conical boundary with
adaptive merge/split, and a scripted bucket moved by the animation kernel
and re-sorted before every step."""
import sys

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_funnel_pour_adaptive_bitexact(gpu):
    from granular_gpu_amd import host
    s = scenes.funnel_pour(6000, 9000, seed=2)
    _log(f"funnel pour n={s.n} boundary={s.boundary.shape[0]}")
    sysm = host.HostSystem.from_scene(s, is_adaptive=1)
    k = 59
    sysm.step(k)
    n = sysm.size("granular")
    st = po.AdaptiveState(s.pos, s.vel, np.ones(s.n, np.float32), cap=2 * s.n)
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    stats = po.adaptive_steps(st, bpos, bcs, p, k + 1)
    _log(f"merged {stats[:, 0].sum()} split {stats[:, 1].sum()} final n {st.n} (gpu {n})")
    assert n == st.n
    assert np.array_equal(sysm.particle_stats(), stats)
    assert np.array_equal(sysm.read("granular", "pos"), st.pos[:n])
    assert np.array_equal(sysm.read("granular", "vel"), st.vel[:n])
    assert np.array_equal(sysm.read("granular", "mass"), st.mass[:n])
    # the column has fallen onto the cone: some particles are inside it
    pos = st.pos[:n]
    assert pos[:, 1].min() < 0.6 * s.space[1]
    sysm.close()


def test_demo_runs_the_excavator_scene_animated(gpu, tmp_path):
    # the reference-style main over the host API: scene JSON with an
    # obj_file point cloud, the moving boundary animated every frame
    import json
    import subprocess
    from pathlib import Path
    demo = Path(__file__).resolve().parents[1] / "granular_gpu_amd" / "granular_demo"
    if not demo.exists():
        pytest.fail(f"{demo} is not built")
    cloud = scenes.bucket_scoop(600, 0.2, 0.3, 0.15, (-0.1, 0.0, -0.15))
    (tmp_path / "bucket.txt").write_text(
        "".join(f"{x!r} {y!r} {z!r}\n" for x, y, z in cloud.tolist()))
    scene = json.loads((Path(__file__).parent / "golden" / "scenes" / "excavator.json")
                       .read_text())
    scene["box_size"] = {"x": 8, "y": 8, "z": 8}
    scene["obj_file"] = str(tmp_path / "bucket.txt")
    scene["save_file"] = str(tmp_path / "out")
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    _log("demo: excavator scene, 12 frames, animated bucket")
    r = subprocess.run([str(demo), str(tmp_path / "scene.json"), "12", str(tmp_path), "animate"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"(animated {cloud.shape[0]})" in r.stdout, r.stdout
    times = (tmp_path / "frame_times.txt").read_text().split()
    assert len(times) == 12 and all(float(t) > 0 for t in times)
    rows = (tmp_path / "stats.csv").read_text().splitlines()
    assert rows[0] == "Frame,MergedParticles,SplitParticles,TotalParticles"


def test_excavator_dig_bitexact(gpu):
    from granular_gpu_amd import host
    dig = scenes.excavator_dig(bed=(0.5, 0.25), fill=0.1, bucket_count=2500)
    s = dig.scene
    _log(f"excavator dig n={s.n} boundary={s.boundary.shape[0]} "
         f"bucket={int(dig.animated.sum())}")
    # the scripted arc, compressed: about 15 steps of translation, then 3
    # degrees per step up to 30 degrees, which swings the bucket's back into
    # the bed
    script = dict(dig.script, translation_time=0.03, rotation_speed=1500.0, rotation_limit=30.0)
    st = host.animation_default()
    for key, v in script.items():
        if key in ("translation_direction", "rotation_center"):
            getattr(st, key)[:] = tuple(float(x) for x in v)
        else:
            setattr(st, key, v)
    ost = po.animation_state(**script)
    sysm = host.HostSystem.from_scene(s, is_move_boundary=True, boundary_animated=dig.animated)
    gp, gv = sysm.read("granular", "pos"), sysm.read("granular", "vel")
    gm = sysm.read("granular", "mass")
    b = sysm.read("boundary", "pos")
    m = sysm.read("boundary", "is_animated")
    p = po.params_from(s, 3, 5)
    dt = float(s.dt)
    moved = calls = 0
    while ost.is_animating:
        mg = sysm.update_animation(st, dt)
        assert mg == po.update_animation(ost, b, m, dt)
        moved += mg
        calls += 1
        assert calls < 100
        sysm.step(1)
        _, perm, bcs = po.grid_build(b, s.cell_length, s.cells)
        b, m = np.ascontiguousarray(b[perm]), np.ascontiguousarray(m[perm])
        gp, gv, gm = po.plain_steps_f32(gp, gv, gm, b, bcs, p, 1)[:3]
    _log(f"{calls} animation calls, {moved} moves")
    assert st.is_animating == 0 and moved == calls - 1 >= 20
    assert list(st.rotation_center) == list(ost.rotation_center)
    assert np.array_equal(sysm.read("boundary", "pos"), b)
    assert np.array_equal(sysm.read("boundary", "is_animated"), m)
    assert np.array_equal(sysm.read("granular", "pos"), gp)
    assert np.array_equal(sysm.read("granular", "vel"), gv)
    # the bucket reached the bed: bed particles were pushed
    assert np.abs(gv).max() > 0.5
    sysm.close()
