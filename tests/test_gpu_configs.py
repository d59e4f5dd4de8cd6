"""The stated configurations against the oracle on the GPU:
- the plain box (32,768 particles, dt 1/240, 3 + 10 sweeps, friction
  0.5/0.4) through all of its 500 steps, bit for bit;
- batched scenes: eight scenes drawn from the 64-scene draw (counts
  log-uniform in [4k, 256k], per-scene friction), dt 1/240, 3 + 10 sweeps,
  advanced by graph launches, each scene bit for bit against the oracle at
  its own friction;
- the learned upsampler's inputs on the 250k-coarse piling state: kNN
  features, flags and the teacher bit for bit, and the fused inference
  against the training forward."""
import sys

import numpy as np
import pytest
import torch

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_plain_box_500_steps_bitexact(gpu):
    from granular_gpu_amd.device import PlainSim
    s = scenes.plain_box(32, 1.0, dt=1.0 / 240.0, g=-9.81)
    _log(f"plain box n={s.n}, 500 steps")
    sim = PlainSim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    sim.step(500)
    gp, gv, gm = sim.host()
    p = po.params_from(s, 3, 10, 0.5, 0.4)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    op, ov, om = po.plain_steps_f32(s.pos, s.vel, s.mass, bpos, bcs, p, 500)[:3]
    assert np.array_equal(gp, op)
    assert np.array_equal(gv, ov)
    assert np.array_equal(gm, om)
    # (at dt 1/240 with the reference's constraint averaging, some of the
    # bottom layers are pushed through the floor lattice: the same ones in
    # both runs)
    _log(f"below the floor after 500 steps: {int((gp[:, 1] < 0.0).sum())} particles")
    assert np.isfinite(gp).all() and np.isfinite(gv).all()


def _spread(sizes, k):
    """k scenes of the draw spanning its range of sizes (sorted by side,
    evenly spaced ranks)."""
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i][0], i))
    pick = [order[round(j * (len(order) - 1) / (k - 1))] for j in range(k)]
    return sorted(set(pick))


def test_batched_scenes_each_match_oracle(gpu):
    from granular_gpu_amd.device import BatchSim, PlainSim
    sizes = scenes.batch_box_sizes(64)
    pick = _spread(sizes, 8)
    dt, launches = 1.0 / 240.0, 10
    sims, defs = [], []
    for k in pick:
        side, mu_s, mu_k = sizes[k]
        sc = scenes.batch_box(side, seed=1 + k, dt=dt, g=-9.81)
        sims.append(PlainSim(sc, stab=3, solver=10, mu_s=mu_s, mu_k=mu_k))
        defs.append((sc, mu_s, mu_k))
    ns = [s.n for s in sims]
    _log(f"batch of {len(sims)} scenes from the 64-scene draw, n = {ns}")
    assert min(ns) < 8000 and max(ns) > 150000
    b = BatchSim(sims, branches=4)
    b.step(launches)
    for sim, (sc, mu_s, mu_k) in zip(sims, defs):
        gp, gv, _ = sim.host()
        p = po.params_from(sc, 3, 10, mu_s, mu_k)
        bpos, bcs = po.sorted_boundary(sc.boundary, sc.cell_length, sc.cells)
        op, ov = po.plain_steps_f32(sc.pos, sc.vel, sc.mass, bpos, bcs, p, launches)[:2]
        _log(f"scene n={sc.n} mu_s={mu_s:.3f}: {int((gp != op).any(axis=1).sum())} differ")
        assert np.array_equal(gp, op)
        assert np.array_equal(gv, ov)
    b.close()


@pytest.mark.parametrize("captured", [False, True])
def test_scenes_in_one_set_each_match_oracle(gpu, captured):
    # the same eight scenes stored back to back in one set: one launch per
    # stage for all of them (eagerly, or replayed from a captured graph with
    # eager work and synchronisations between the replays), each scene bit
    # for bit against the oracle at its own friction
    from granular_gpu_amd.device import SceneBatch
    sizes = scenes.batch_box_sizes(64)
    pick = _spread(sizes, 8)
    dt, steps = 1.0 / 240.0, 10
    scs = [scenes.batch_box(sizes[k][0], seed=1 + k, dt=dt, g=-9.81) for k in pick]
    fr = [(sizes[k][1], sizes[k][2]) for k in pick]
    b = SceneBatch(scs, fr, stab=3, solver=10)
    _log(f"{len(scs)} scenes in one set: n = {b.n}, cells = {b.total_cells}, "
         f"captured = {captured}")
    if captured:
        b.capture(steps_per_launch=1)
        for _ in range(steps):
            b.launch(1)
            torch.cuda.synchronize()
            torch.zeros(1024, device="cuda").sum().item()  # eager work between replays
    else:
        b.step(steps)
    for k, (sc, (mu_s, mu_k)) in enumerate(zip(scs, fr)):
        gp, gv, gm = b.scene(k)
        p = po.params_from(sc, 3, 10, mu_s, mu_k)
        bpos, bcs = po.sorted_boundary(sc.boundary, sc.cell_length, sc.cells)
        op, ov, om = po.plain_steps_f32(sc.pos, sc.vel, sc.mass, bpos, bcs, p, steps)[:3]
        _log(f"scene {k}: n={sc.n} mu_s={mu_s:.3f}: {int((gp != op).any(axis=1).sum())} differ")
        assert np.array_equal(gp, op)
        assert np.array_equal(gv, ov)
        assert np.array_equal(gm, om)
    b.close_graph()


def _bf16_to_f32(t: torch.Tensor) -> np.ndarray:
    u = t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32)


def test_learned_inputs_at_250k_coarse(gpu):
    from granular_gpu_amd.device import PlainSim
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet
    s = scenes.plain_box(63, 3.0, dt=0.002, g=-9.8, r=0.02)
    _log(f"piling scene: coarse n={s.n}")
    sim = PlainSim(s, stab=3, solver=5)
    sim.step(40)
    sim.grid_build()
    torch.cuda.synchronize()
    up = LearnedUpsampler(MlpNet(seed=2))
    feat, flags, knn = up.features(sim)
    y, fp, _ = up.teacher(sim)
    n = sim.n
    pos = sim.pos[:n].cpu().numpy().copy()
    vel = sim.vel[:n].cpu().numpy().copy()
    cs, bcs = sim.cs.cpu().numpy().copy(), sim.bcs.cpu().numpy().copy()
    bpos = sim.bpos.cpu().numpy().copy()
    p = po.params_from(s)
    cfg = po.mlp_cfg(0.02)
    of, ofl, oknn = po.mlp_features(pos, vel, cs, bpos, bcs, p, cfg)
    _log(f"features: flagged {int((ofl != 0).sum())} of {n}")
    assert np.array_equal(knn.cpu().numpy(), oknn)
    assert np.array_equal(flags.cpu().numpy(), ofl)
    assert np.array_equal(_bf16_to_f32(feat).reshape(-1, 96), of)
    oy, ofp = po.mlp_teacher(pos, vel, cs, bpos, bcs, p, cfg)
    assert np.array_equal(y.cpu().numpy(), oy)
    assert np.array_equal(fp.cpu().numpy(), ofp)
    # fused inference against the training forward on the same features
    from granular_gpu_amd.learned import Trainer
    fine = up.infer(sim)
    yf = Trainer(up.net, batch=n).forward(feat).cpu().numpy()
    want = po.mlp_positions(pos, vel, yf, cfg)
    fl = ofl
    got = fine.cpu().numpy()
    assert np.array_equal(got[fl != 0], ofp[fl != 0])
    d = np.abs(got[fl == 0] - want[fl == 0]).max()
    _log(f"inference vs training forward: max |dx| {d:.2e} m")
    assert d < 1e-5
