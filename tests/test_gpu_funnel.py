"""The funnel pour with radius-level merge/split rules (radius levels
0.005 / 0.01 / 0.02, i.e. masses 0.125 / 1 / 8; merge deeper than 3 r below
the free surface, split within 2 r of it): bit for bit against the oracle on
a 60k column over a 1 m cone, and at full size (1M particles over a cone of
400k boundary particles) through properties that hold at any size: mass
conservation, the particle count following the merge/split counts, finite
state inside the walls, and the column draining through the outlet."""
import sys

import numpy as np
import pytest
import torch

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DEPTH = dict(min_mass=0.125, split_mass=0.125, merge_dist=1.75, merge_depth=3.0,
             split_depth=2.0)
MAX_MASS = 8.0


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(scene, cap_factor, **lists):
    from granular_gpu_amd.device import AdaptiveSim
    sim = AdaptiveSim(scene, capacity=cap_factor * scene.n, max_mass=MAX_MASS, stab=3,
                      solver=5, **lists)
    sim.params.adaptive_rules = 1
    for k, v in DEPTH.items():
        setattr(sim.params, k, v)
    return sim


def test_funnel_depth_rules_bitexact_60k(gpu):
    steps = 100
    s = scenes.funnel_column(60000, 24000, cone_height=1.0)
    _log(f"funnel column n={s.n} boundary={s.boundary.shape[0]} cells={s.cells} "
         f"steps={steps}")
    sim = _sim(s, 4)
    sim.step(steps)
    got = sim.host_all()
    p = po.params_from(s, 3, 5, rules=dict(DEPTH, depth_rules=1))
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    st = po.AdaptiveState(s.pos, s.vel, s.mass, cap=sim.cap)
    stats = po.adaptive_steps(st, bpos, bcs, p, steps, max_mass=MAX_MASS)
    _log(f"merged {stats[:, 0].sum()} split {stats[:, 1].sum()} final n {st.n} "
         f"(gpu {sim.n})")
    assert stats[:, 0].sum() > 0 and stats[:, 1].sum() > 0
    assert np.array_equal(np.array(sim.stats, np.int32), stats)
    assert sim.n == st.n
    n = st.n
    for k, ref in (("pos", st.pos), ("vel", st.vel), ("mass", st.mass), ("adl", st.adl),
                   ("remove", st.remove), ("merge", st.merge), ("mcount", st.mcount)):
        assert np.array_equal(got[k], ref[:n]), k


def test_funnel_1m_cone_400k_depth_rules_properties(gpu):
    s = scenes.funnel_column()  # 1M particles, 400k cone, 979,843 cells
    _log(f"funnel column n={s.n} boundary={s.boundary.shape[0]} cells={s.cells}")
    assert s.n == 1_000_000 and s.boundary.shape[0] - 400_000 > 0
    sim = _sim(s, 4)
    m0 = float(s.mass.sum(dtype=np.float64))
    lo = np.zeros(3)
    hi = np.array(s.space)
    gap = 0.6  # the outlet's height above the floor
    below = []
    done = 0
    for chunk in (100, 100, 100, 100, 50, 50):
        sim.step(chunk)
        done += chunk
        torch.cuda.synchronize()
        n = sim.n
        pos = sim.pos[:n].cpu().numpy()
        vel = sim.vel[:n].cpu().numpy()
        mass = sim.mass[:n].cpu().numpy()
        stats = np.array(sim.stats, np.int64)
        below.append(int((pos[:, 1] < gap).sum()))
        _log(f"step {done}: n {n} merged {stats[:, 0].sum()} split {stats[:, 1].sum()} "
             f"below outlet {below[-1]} mass {mass.sum(dtype=np.float64):.3f}")
        assert np.isfinite(pos).all() and np.isfinite(vel).all()
        assert (mass > 0).all() and mass.max() < 2 * MAX_MASS
        assert abs(float(mass.sum(dtype=np.float64)) - m0) < 1e-5 * m0
        # the loose random column sprays a few particles out on its first
        # contacts; they leave through the walls' gaps and fall freely
        out_lo, out_hi = (pos < lo).sum(0), (pos > hi).sum(0)
        outside = int(((pos < lo) | (pos > hi)).any(1).sum())
        if outside:
            _log(f"  outside the space: below {out_lo.tolist()} above {out_hi.tolist()}")
        assert outside <= n // 1000, outside
    # every step's count follows from the previous one and its merge/split counts
    stats = np.array(sim.stats, np.int64)
    assert np.array_equal(stats[1:, 2], stats[:-1, 2] - stats[:-1, 0] + stats[:-1, 1])
    assert sim.n == stats[-1, 2] - stats[-1, 0] + stats[-1, 1]
    assert stats[:, 0].sum() > 0 and stats[:, 1].sum() > 0
    # both radius levels away from the unit mass are populated
    assert (mass <= 0.125 + 1e-6).sum() > 0 and (mass >= 2.0).sum() > 0
    # the column drains through the outlet: nothing below it before the fall
    # time of the axis particles (2.52 m, 0.72 s, step 359), then more and more
    assert below[0] == 0 and below[-1] > below[-2] > 0


def test_funnel_1m_lists_do_not_change_results(gpu):
    # 16 boundary slots overflow beside the cone (those particles scan their
    # stencils); 96 slots, 160 granular ones and the mid-step refresh use the
    # lists there instead: the same state bit for bit after 300 steps
    from granular_gpu_amd._lib import GG_FLAG_LIST_REFRESH
    s = scenes.funnel_column()
    out = []
    for lists in (dict(), dict(gmax=160, bmax=96, flags=GG_FLAG_LIST_REFRESH)):
        sim = _sim(s, 4, **lists)
        sim.step(300)
        torch.cuda.synchronize()
        n = sim.n
        out.append((n, sim.pos[:n].cpu().numpy(), sim.vel[:n].cpu().numpy(),
                    sim.mass[:n].cpu().numpy(), np.array(sim.stats, np.int64)))
        _log(f"lists {lists}: n {n}, merged {out[-1][4][:, 0].sum()}, "
             f"split {out[-1][4][:, 1].sum()}")
        del sim
        torch.cuda.empty_cache()
    (na, *a), (nb, *b) = out
    assert na == nb
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert a[3][:, 0].sum() > 0 and a[3][:, 1].sum() > 0
