"""The reference's own box scene (scenes/box.json: 100,000 granular
particles, 51,088 walls, 2.5M fine particles, adaptive) through the host
C++ API on the GPU: bit-exact against the oracle over the first merge wave,
and the particle-count trajectory against the statistics the reference
recorded for this scene (tests/golden/box_stats.csv)."""
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
sys.path.insert(0, str(Path(__file__).parent))
from test_oracle_box_scene import box_scene, n_after, reference_stats  # noqa: E402


def _log(msg):
    print(msg, flush=True)


def _system(sc, fine=True):
    from granular_gpu_amd import host
    f = host.upsampled_cloud(sc.pos, 0.01, 25) if fine else None
    return host.HostSystem.from_scene(sc, fine=f, is_adaptive=1)


def test_box_scene_first_merge_wave_bitexact(gpu):
    sc = box_scene()
    frames = 45
    _log(f"box scene: {frames} frames on the GPU and in the oracle")
    sysm = _system(sc, fine=False)
    sysm.step(frames - 1)
    p = po.params_from(sc, 3, 5)
    bpos, bcs = po.sorted_boundary(sc.boundary, sc.cell_length, sc.cells)
    st = po.AdaptiveState(sc.pos, sc.vel, sc.mass, cap=2 * sc.n)
    stats = po.adaptive_steps(st, bpos, bcs, p, frames)
    assert stats[40, 0] > 10000
    assert np.array_equal(sysm.particle_stats(), stats)
    n = st.n
    assert sysm.size("granular") == n
    assert np.array_equal(sysm.read("granular", "pos"), st.pos[:n])
    assert np.array_equal(sysm.read("granular", "vel"), st.vel[:n])
    assert np.array_equal(sysm.read("granular", "mass"), st.mass[:n])
    sysm.close()


def test_box_scene_count_trajectory(gpu):
    sc = box_scene()
    frames = 400
    _log(f"box scene with 2.5M fine particles: {frames} frames")
    sysm = _system(sc, fine=True)
    assert sysm.size("fine") == 2500000
    sysm.step(frames - 1)
    stats = sysm.particle_stats()
    ref = reference_stats()[:frames]
    ours, theirs = n_after(stats), n_after(ref)
    rel = np.abs(ours - theirs) / theirs
    _log(f"N at frame {frames - 1}: {ours[-1]} (reference {theirs[-1]}); "
         f"max rel 100-399 {rel[100:].max():.4f}, 300-399 {rel[300:].max():.4f}")
    assert stats[:40, 0].sum() == 0 and stats[40, 0] > 0
    assert rel[100:].max() < 0.07
    assert rel[300:].max() < 0.03
    assert abs(stats[:, 0].sum() - ref[:, 1].sum()) < 0.05 * ref[:, 1].sum()
    ft = sysm.frame_times()
    assert ft.shape == (frames,)
    _log(f"median frame time {np.median(ft):.3f} ms")
    sysm.close()
