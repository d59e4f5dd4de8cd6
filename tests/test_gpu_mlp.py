"""GPU tests of the learned upsampler: the MFMA lane layout, features and
teacher bit for bit against the oracle, the MFMA training forward and
backward against the oracle's (gradient check), Adam bit for bit, the fused
inference kernel against the training forward and the teacher fallback,
and what training gains over the method without the network."""
import sys

import numpy as np
import pytest
import torch

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
R = 0.02


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bf16_to_f32(t: torch.Tensor) -> np.ndarray:
    u = t.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32)


def _pile(side=12, steps=150):
    from granular_gpu_amd.device import PlainSim
    s = scenes.plain_box(side, 1.0, r=R, dt=0.002, g=-9.8)
    sim = PlainSim(s)
    sim.step(steps)
    sim.grid_build()
    torch.cuda.synchronize()
    return s, sim


def _host_state(s, sim):
    n = sim.n
    pos = sim.pos[:n].cpu().numpy().copy()
    vel = sim.vel[:n].cpu().numpy().copy()
    cs = sim.cs.cpu().numpy().copy()
    bpos = sim.bpos.cpu().numpy().copy()
    bcs = sim.bcs.cpu().numpy().copy()
    return pos, vel, cs, bpos, bcs, po.params_from(s)


def test_mfma_lane_layout(gpu):
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import _ptr, _stream
    rng = np.random.default_rng(0)
    A = (rng.integers(-8, 9, (16, 16)) / 4.0).astype(np.float32)
    B = (rng.integers(-8, 9, (16, 16)) / 4.0).astype(np.float32)
    to16 = lambda a: torch.from_numpy((a.view(np.uint32) >> 16).astype(np.int16)).cuda()
    a, b = to16(A), to16(B)
    d = torch.zeros((16, 16), dtype=torch.float32, device="cuda")
    _lib.check("gg_mfma_probe", _lib.load().gg_mfma_probe(_ptr(a), _ptr(b), _ptr(d), _stream()))
    assert np.array_equal(d.cpu().numpy(), A @ B)


def test_bf16_hardware_rounding_matches_definition(gpu):
    # the inference kernel rounds hidden activations with v_cvt_pk_bf16_f32;
    # it must equal the defined round-to-nearest-even on every non-NaN float
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import _ptr, _stream
    bad = torch.zeros(1, dtype=torch.int64, device="cuda")
    _lib.check("gg_bf16_round_check",
               _lib.load().gg_bf16_round_check(0, 1 << 32, _ptr(bad), _stream()))
    torch.cuda.synchronize()
    assert int(bad.item()) == 0


def test_features_and_teacher_bitexact(gpu):
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet
    s, sim = _pile()
    up = LearnedUpsampler(MlpNet(seed=1))
    feat, flags, knn = up.features(sim)
    y, fp, fv = up.teacher(sim)
    pos, vel, cs, bpos, bcs, p = _host_state(s, sim)
    cfg = po.mlp_cfg(R)
    of, ofl, oknn = po.mlp_features(pos, vel, cs, bpos, bcs, p, cfg)
    _log(f"features: n={sim.n}, flagged {int((ofl != 0).sum())}")
    assert 0 < (ofl != 0).sum() < sim.n
    assert np.array_equal(knn.cpu().numpy(), oknn)
    assert np.array_equal(flags.cpu().numpy(), ofl)
    assert np.array_equal(_bf16_to_f32(feat).reshape(-1, 96), of)
    oy, ofp = po.mlp_teacher(pos, vel, cs, bpos, bcs, p, cfg)
    assert np.array_equal(y.cpu().numpy(), oy)
    assert np.array_equal(fp.cpu().numpy(), ofp)
    assert np.abs(oy).max() > 0.1


def test_training_gradients_match_oracle(gpu):
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet, Trainer
    s, sim = _pile()
    net = MlpNet(seed=3)
    up = LearnedUpsampler(net)
    feat, flags, _ = up.features(sim)
    y, _, _ = up.teacher(sim)
    keep = torch.nonzero(flags == 0).flatten()[:512]
    x, t = feat[keep].contiguous(), y[keep].contiguous()
    B = x.shape[0]
    tr = Trainer(net, batch=B)
    gy = tr.forward(x).cpu().numpy()
    dW, db, sq = tr.grad(x, t)
    dW, db = dW.cpu().numpy(), db.cpu().numpy()
    W, b = net.host_bf16()
    xf, tf = _bf16_to_f32(x).reshape(B, 96), t.cpu().numpy()
    oy = po.mlp_forward(xf, W, b)
    assert np.abs(gy - oy).max() <= 1e-2 * np.abs(oy).max()
    loss, odW, odb = po.mlp_backward(xf, W, b, tf)
    gl = float(sq.sum().item()) / (B * 24)
    _log(f"gradient check: loss gpu {gl:.6e} oracle {loss:.6e}")
    assert gl == pytest.approx(loss, rel=1e-3)
    for l, (k_in, n_out) in enumerate(po.MLP_DIMS):
        sl = slice(po.MLP_WOFF[l], po.MLP_WOFF[l] + n_out * k_in)
        rel = np.linalg.norm(dW[sl] - odW[sl]) / np.linalg.norm(odW[sl])
        bl = slice(po.MLP_BOFF[l], po.MLP_BOFF[l] + n_out)
        relb = np.linalg.norm(db[bl] - odb[bl]) / max(np.linalg.norm(odb[bl]), 1e-30)
        _log(f"layer {l}: dW rel err {rel:.2e}, db rel err {relb:.2e}")
        assert rel < 1e-2 and relb < 1e-2


def test_adam_bitexact(gpu):
    from granular_gpu_amd.learned import MlpNet, Trainer
    net = MlpNet(seed=4)
    tr = Trainer(net, batch=16)
    rng = np.random.default_rng(1)
    w = net.w_master.cpu().numpy().copy()
    b = net.b_master.cpu().numpy().copy()
    mw, vw = np.zeros_like(w), np.zeros_like(w)
    mb, vb = np.zeros_like(b), np.zeros_like(b)
    for step in (1, 2):
        gw = rng.standard_normal(w.size).astype(np.float32) * 1e-3
        gb = rng.standard_normal(b.size).astype(np.float32) * 1e-3
        tr.dW.copy_(torch.from_numpy(gw))
        tr.db.copy_(torch.from_numpy(gb))
        tr.adam(lr=1e-3)
        w, mw, vw = po.mlp_adam(w, mw, vw, gw, 1e-3, step)
        b, mb, vb = po.mlp_adam(b, mb, vb, gb, 1e-3, step)
    torch.cuda.synchronize()
    assert np.array_equal(net.w_master.cpu().numpy(), w)
    assert np.array_equal(net.m_w.cpu().numpy(), mw) and np.array_equal(net.v_w.cpu().numpy(), vw)
    assert np.array_equal(net.b_master.cpu().numpy(), b)
    assert np.array_equal(net.bias.cpu().numpy(), b)
    # the bfloat16 copy is the rounded master
    u = w.view(np.uint32).astype(np.uint64)
    want = (((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)).view(np.float32)
    assert np.array_equal(net.host_bf16()[0], want)


def test_inference_matches_training_forward_and_fallback(gpu):
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet, Trainer
    s, sim = _pile()
    net = MlpNet(seed=5)
    up = LearnedUpsampler(net)
    flagged = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    fine = up.infer(sim, fallback_count=flagged)
    feat, flags, _ = up.features(sim)
    tr = Trainer(net, batch=sim.n)
    y = tr.forward(feat).cpu().numpy()
    pos, vel, cs, bpos, bcs, p = _host_state(s, sim)
    cfg = po.mlp_cfg(R)
    want = po.mlp_positions(pos, vel, y, cfg)
    _, tfp, _ = up.teacher(sim)
    fl = flags.cpu().numpy()
    got = fine.cpu().numpy()
    assert int(flagged.item()) == int((fl != 0).sum()) > 0
    assert np.array_equal(got[fl != 0], tfp.cpu().numpy()[fl != 0])
    d = np.abs(got[fl == 0] - want[fl == 0]).max()
    _log(f"inference vs training forward: max |dx| {d:.2e} m over {int((fl == 0).sum())} coarse")
    assert d < 1e-5


def test_inference_configs_agree(gpu):
    # waves per block and sample tiles per wave change the launch shape only
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet
    s, sim = _pile()
    up = LearnedUpsampler(MlpNet(seed=7))
    ref = None
    for w, t in [(4, 1), (4, 2), (4, 4), (8, 1), (8, 2), (8, 4), (16, 1)]:
        up.cfg.infer_waves, up.cfg.infer_tiles = w, t
        got = up.infer(sim).cpu().numpy()
        if ref is None:
            ref = got
        assert np.array_equal(got, ref), (w, t)


def test_training_gains_over_rigid_template(gpu):
    # train on snapshots of one run, score on another seed's dynamic phase
    from granular_gpu_amd.device import PlainSim
    from granular_gpu_amd.learned import LearnedUpsampler, MlpNet, Trainer, fine_error
    sim = PlainSim(scenes.plain_box(12, 1.0, r=R, dt=0.002, g=-9.8, seed=1))
    net = MlpNet(seed=6)
    up = LearnedUpsampler(net)
    x, t = up.collect(sim, snapshots=24, every=5, warmup=40)
    _log(f"training pairs: {x.shape[0]}")
    assert x.shape[0] > 20000
    tr = Trainer(net, batch=512)
    losses = tr.fit(x, t, steps=1500, lr=2e-3, lr_final=1e-4, seed=0)
    _log(f"loss {losses[:20].mean():.3e} -> {losses[-50:].mean():.3e}")
    # the untrained network outputs ~0 (the rigid template), so the first
    # losses are the rigid baseline's: training must cut them at least 4x
    assert losses[-50:].mean() < 0.25 * losses[:20].mean()
    r_fine = 0.01
    ev = PlainSim(scenes.plain_box(12, 1.0, r=R, dt=0.002, g=-9.8, seed=2))
    ev.step(90)
    for extra in (0, 40):
        ev.step(extra)
        ev.grid_build()
        _, tfp, _ = up.teacher(ev)
        e_learn = fine_error(up.infer(ev), tfp)
        e_rigid = fine_error(up.rigid(ev), tfp)
        _log(f"held-out seed 2, step {90 + extra}: learned {e_learn / r_fine:.4f} r_fine, "
             f"rigid {e_rigid / r_fine:.4f} r_fine")
        assert e_learn < 0.25 * r_fine
        assert e_learn < 0.5 * e_rigid


def _paired(side, seed, steps):
    from granular_gpu_amd.device import PlainSim
    coarse, fine = scenes.paired_boxes(side, 1.0, r=R, seed=seed)
    cs_, fs_ = PlainSim(coarse), PlainSim(fine)
    cs_.step(steps)
    fs_.step(steps)
    return coarse, fine, cs_, fs_


def test_fine_targets_bitexact(gpu):
    from granular_gpu_amd.learned import fine_targets
    _log("fine-run targets of a 10^3 coarse / 8000 fine pair at step 60")
    coarse, fine, csim, fsim = _paired(10, 1, 60)
    csim.grid_build()
    fsim.grid_build()
    y, d2, jm = fine_targets(csim, fsim)
    torch.cuda.synchronize()
    n, nf = csim.n, fsim.n
    pos, vel = csim.pos[:n].cpu().numpy(), csim.vel[:n].cpu().numpy()
    fp, fv = fsim.pos[:nf].cpu().numpy(), fsim.vel[:nf].cpu().numpy()
    fcs = fsim.cs.cpu().numpy()
    ry, rd2, rjm = po.mlp_fine_targets(pos, vel, fp, fv, fcs, fine.cell_length, fine.cells, R)
    assert np.array_equal(y.cpu().numpy(), ry)
    assert np.array_equal(d2.cpu().numpy(), rd2)
    assert np.array_equal(jm.cpu().numpy(), rjm)
    assert (rjm >= 0).all() and np.abs(ry).max() > 0


def test_training_on_fine_run_pairs_gains_over_rigid_and_teacher(gpu):
    # pairs from a coarse run (r = 0.02) and its fine twin (r = 0.01) seeded
    # identically; scored against the fine-run truth of another seed's pair
    from granular_gpu_amd.device import PlainSim
    from granular_gpu_amd.learned import (LearnedUpsampler, MlpNet, Trainer, collect_pairs,
                                          fine_targets)
    coarse, fine = scenes.paired_boxes(16, 1.0, r=R, seed=1)
    csim, fsim = PlainSim(coarse), PlainSim(fine)
    net = MlpNet(seed=6)
    up = LearnedUpsampler(net)
    x, t = collect_pairs(up, csim, fsim, snapshots=16, every=5, warmup=40)
    _log(f"training pairs: {x.shape[0]}")
    assert x.shape[0] > 20000
    tr = Trainer(net, batch=512)
    losses = tr.fit(x, t, steps=2000, lr=2e-3, lr_final=1e-4, seed=0)
    _log(f"loss {losses[:20].mean():.3e} -> {losses[-50:].mean():.3e}")
    assert losses[-50:].mean() < losses[:20].mean()
    _, _, ce, fe = _paired(16, 2, 80)
    lim = (1.0 * R) ** 2
    for extra in (0, 30):
        ce.step(extra)
        fe.step(extra)
        ce.grid_build()
        fe.grid_build()
        feat, flags, _ = up.features(ce)
        truth, d2, _ = fine_targets(ce, fe)
        keep = (flags == 0) & (d2 < lim).all(dim=1)
        yt = truth[keep]
        y_net = tr.forward(feat[keep].contiguous())
        y_teacher, _, _ = up.teacher(ce)
        y_teacher = y_teacher[keep]

        def rms(a):
            return float(torch.sqrt(((a - yt) ** 2).mean()).item())
        e_net, e_rigid, e_teacher = rms(y_net), rms(torch.zeros_like(yt)), rms(y_teacher)
        _log(f"held-out seed 2, step {80 + extra}, {int(keep.sum())} coarse: fine-velocity "
             f"rms error: learned {e_net:.4f}, rigid {e_rigid:.4f}, teacher {e_teacher:.4f} m/s")
        assert e_net < 0.9 * e_rigid and e_net < e_teacher
