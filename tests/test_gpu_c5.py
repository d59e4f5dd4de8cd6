"""The excavator dig at full size through the host system: 4M particles of
r = 0.004 on a 2 x 1 m bed (a 2r lattice, jitter 0.1 r, seeded), a bucket of
200k boundary particles on the scripted arc (0.5 m in 1 s, then 90 degrees
at 45 degrees/s; the first 2 s, 1000 steps, here), adaptive merge/split with
the radius-level rules, and the learned upsampler writing 8 fine particles
per granular particle (32M at the start) from GranularSystem::step.  Too large for the oracle: checked through properties
(mass, counts, the merge/split bookkeeping, finite state inside the space,
the bucket's motion and the bed's response, the fine set's size and
placement)."""
import math
import sys

import numpy as np
import pytest

from granular_gpu_amd import scenes

pytestmark = pytest.mark.gpu
R = 0.004


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _trained_net():
    """A network trained on a coarse/fine twin pair (r = 0.02 / 0.01); its
    features are scaled by the coarse radius, so it is applied at r = 0.004."""
    from granular_gpu_amd.device import PlainSim
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet, Trainer, collect_pairs
    coarse, fine = scenes.paired_boxes(16, 1.0, r=0.02, seed=1)
    net = MlpNet(seed=6)
    up = LearnedUpsampler(net)
    x, t = collect_pairs(up, PlainSim(coarse), PlainSim(fine), snapshots=8, every=5, warmup=40)
    Trainer(net, batch=512).fit(x, t, steps=1000, lr=2e-3, lr_final=1e-4, seed=0)
    w, b = net.host_bf16()
    return (w.view(np.uint32) >> 16).astype(np.uint16), b


def test_excavator_4m_bucket_200k_adaptive_learned_32m(gpu):
    from granular_gpu_amd import host
    w, b = _trained_net()
    dig = scenes.excavator_dig(bed=(2.0, 1.0), fill=1.11, r=R, bucket_count=200_000)
    s = dig.scene
    # the host system starts every granular mass at 1 (the reference's
    # constructor): the density makes mass 1 a particle of radius R
    s.density = np.float32(3.0 / (4.0 * math.pi * R ** 3))
    nb_bucket = int(dig.animated.sum())
    _log(f"excavator n={s.n} boundary={s.boundary.shape[0]} bucket={nb_bucket} "
         f"cells={s.cells}")
    assert 3.95e6 < s.n < 4.1e6 and nb_bucket > 195_000
    sysm = host.HostSystem.from_scene(s, is_move_boundary=True, is_adaptive=1,
                                      boundary_animated=dig.animated)
    sysm.set_merge_rules()  # radius levels r/2, r, 2r: masses 0.125, 1, 8
    # the bucket's 200k particles are 3.4 mm apart: a particle beside it has
    # about 80 boundary candidates within contact distance plus the skin;
    # merged (2r) particles among split (r/2) ones have 100+ granular ones
    sysm.set_lists(gmax=160, bmax=96, skin=0.75 * R)
    # lists lapse inside the step where splits overlap and the bucket pushes:
    # rebuilt mid-step when the sweeps' stencil scans say so
    from granular_gpu_amd import _lib
    sysm.set_solver(3, 5, 0.0, 0.0, _lib.GG_FLAG_LIST_REFRESH)
    # the network's guards in units of the coarse radius as at r = 0.02
    # (boundary 2.5 r, floor 4 r, sides 2.5 r)
    sysm.set_learned(w, b, radius=R, guard_boundary=2.5 * R, guard_floor=4 * R,
                     guard_side=2.5 * R)
    st = host.animation_default()
    for key, v in dig.script.items():
        if key in ("translation_direction", "rotation_center"):
            getattr(st, key)[:] = tuple(float(x) for x in v)
        else:
            setattr(st, key, v)
    dt = float(s.dt)
    b0 = sysm.read("boundary", "pos")
    anim0 = sysm.read("boundary", "is_animated")
    bucket0 = b0[anim0 == 1].astype(np.float64).mean(0)
    n0 = sysm.size("granular")
    hi = np.array(s.space)
    done, vmax = 0, 0.0
    for chunk in (250, 250, 500):
        for _ in range(chunk):
            sysm.update_animation(st, dt)
            sysm.step(1)
        done += chunk
        n = sysm.size("granular")
        mass = sysm.read("granular", "mass")
        vel = sysm.read("granular", "vel")
        pos = sysm.read("granular", "pos")
        vmax = max(vmax, float(np.abs(vel).max()))
        stats = sysm.particle_stats()
        _log(f"step {done}: n {n} merged {stats[:, 0].sum()} split {stats[:, 1].sum()} "
             f"mass {mass.sum(dtype=np.float64):.2f} |v|max {np.abs(vel).max():.3f} "
             f"fine {sysm.size('fine')} frame {sysm.frame_times()[-1]:.2f} ms")
        assert np.isfinite(pos).all() and np.isfinite(vel).all()
        assert abs(float(mass.sum(dtype=np.float64)) - n0) < 1e-5 * n0
        outside = int(((pos < 0) | (pos > hi)).any(1).sum())
        assert outside <= n // 10000, outside
        # the fine set is written from the step's state before merge/split
        assert sysm.size("fine") == 8 * int(stats[-1, 2])
    fine = sysm.read("fine", "pos").reshape(-1, 8, 3).astype(np.float64)
    assert np.isfinite(fine).all()
    # the 8 fine particles of a parent sit around it on the template
    # (corners at 0.87 R from the centre) moved by dt times their velocity
    d = np.linalg.norm(fine - fine.mean(1, keepdims=True), axis=2).max()
    _log(f"largest fine distance from its group's centre {d / R:.2f} R")
    assert d < 2 * R
    # the merge/split bookkeeping: each row's N follows from the one before
    stats = sysm.particle_stats().astype(np.int64)
    assert np.array_equal(stats[1:, 2], stats[:-1, 2] - stats[:-1, 0] + stats[:-1, 1])
    assert stats[:, 0].sum() > 0 and stats[:, 1].sum() > 0
    # the bucket followed the arc (0.5 m along the translation in 1 s)
    b1 = sysm.read("boundary", "pos")
    anim1 = sysm.read("boundary", "is_animated")
    bucket1 = b1[anim1 == 1].astype(np.float64).mean(0)
    _log(f"bucket centroid {bucket0.round(3)} -> {bucket1.round(3)}")
    assert np.linalg.norm(bucket1 - bucket0) > 0.3
    # and the bed answered: particles were pushed
    assert vmax > 0.3
    sysm.close()
