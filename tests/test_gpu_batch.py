"""Batched scenes in one hipGraph launch (gg_batch_*): every replay equals
the same steps launched eagerly, bit for bit, with eager work and host
synchronisation between replays; the autotuner leaves the scenes untouched;
block sizes and branch counts never change results."""
import sys

import numpy as np
import pytest
import torch

from granular_gpu_amd import scenes

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sims(sides, mus, seed=1):
    from granular_gpu_amd.device import PlainSim
    out = []
    for k, (side, mu) in enumerate(zip(sides, mus)):
        s = scenes.batch_box(side, seed=seed + k)
        out.append(PlainSim(s, stab=3, solver=10, mu_s=mu, mu_k=0.8 * mu))
    return out


SIDES = [6, 11, 4, 9, 14, 7, 5, 12, 8, 10]
MUS = [0.3, 0.45, 0.5, 0.7, 0.35, 0.6, 0.4, 0.55, 0.65, 0.3]


def _state(sims):
    torch.cuda.synchronize()
    return [(s.pos[:s.n].cpu().numpy().copy(), s.vel[:s.n].cpu().numpy().copy()) for s in sims]


def test_batch_graph_replays_match_eager(gpu):
    from granular_gpu_amd.device import BatchSim
    a, b = _sims(SIDES, MUS), _sims(SIDES, MUS)
    _log(f"batch of {len(a)} scenes, n = {[s.n for s in a]}")
    ga = BatchSim(a, branches=3, steps_per_launch=2)
    for rep in range(3):
        ga.step(1)                      # graph: 2 steps of every scene
        for s in b:                     # eager: the same steps
            s.step(2)
        x = torch.randn(1 << 16, device="cuda")
        (x * 2).sum().item()            # unrelated eager work + host sync
        sa, sb = _state(a), _state(b)
        for k, ((pa, va), (pb, vb)) in enumerate(zip(sa, sb)):
            assert np.array_equal(pa, pb), (rep, k)
            assert np.array_equal(va, vb), (rep, k)
    ga.close()


def test_batch_eager_entry_matches_graph(gpu):
    from granular_gpu_amd.device import BatchSim
    a, b = _sims(SIDES[:5], MUS[:5]), _sims(SIDES[:5], MUS[:5])
    ga, gb = BatchSim(a, branches=2), BatchSim(b, branches=1)
    ga.step(3)
    gb.step_eager(3)
    for (pa, va), (pb, vb) in zip(_state(a), _state(b)):
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    ga.close()
    gb.close()


def test_block_sizes_and_branches_do_not_change_results(gpu):
    from granular_gpu_amd.device import BatchSim
    ref = None
    for br, sw, nl in [(1, 64, 128), (2, 128, 256), (4, 256, 128), (3, 512, 256)]:
        sims = _sims(SIDES[:6], MUS[:6])
        g = BatchSim(sims, branches=br, sweep_block=sw, nlist_block=nl)
        g.step(4)
        st = _state(sims)
        g.close()
        if ref is None:
            ref = st
        else:
            for (p0, v0), (p1, v1) in zip(ref, st):
                assert np.array_equal(p0, p1) and np.array_equal(v0, v1), (br, sw, nl)


def test_autotune_leaves_state_and_picks_valid_config(gpu):
    from granular_gpu_amd.device import BatchSim
    a, b = _sims(SIDES, MUS), _sims(SIDES, MUS)
    ga = BatchSim(a)
    ga.step(2)
    for s in b:
        s.step(2)
    before = _state(a)
    best = ga.autotune(trials=1)
    _log(f"autotune: {best}")
    assert best["branches"] in (1, 2, 4, 8) and best["sweep_block"] in (64, 128, 256, 512)
    assert best["nlist_block"] in (128, 256) and best["ms_per_launch"] > 0
    assert ga.config()["sweep_block"] == best["sweep_block"]
    after = _state(a)
    for (p0, v0), (p1, v1) in zip(before, after):
        assert np.array_equal(p0, p1) and np.array_equal(v0, v1)
    ga.step(2)
    for s in b:
        s.step(2)
    for (pa, va), (pb, vb) in zip(_state(a), _state(b)):
        assert np.array_equal(pa, pb) and np.array_equal(va, vb)
    ga.close()


def test_batch_rejects_bad_config(gpu):
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import BatchSim
    sims = _sims(SIDES[:2], MUS[:2])
    with pytest.raises(_lib.GGError):
        BatchSim(sims, branches=3)          # more branches than scenes
    with pytest.raises(_lib.GGError):
        BatchSim(sims, sweep_block=96)      # not a supported block size
