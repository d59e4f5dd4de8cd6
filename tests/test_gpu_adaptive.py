"""GPU parity of the adaptive merge/split path: surface flags, surface
distance, merge decisions, blends, splits and compaction must equal the
oracle bit for bit, stage by stage and over whole adaptive steps."""
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


def _pile(side=20, steps=30):
    """A falling block from the oracle (plain steps, dt = 0.002): compact,
    with a deep interior (no surface particle within its stencil)."""
    s = scenes.plain_box(side, 1.0, dt=0.002, g=-9.8)
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    pos, vel, mass = po.plain_steps_f32(s.pos, s.vel, s.mass, bpos, bcs, p, steps)[:3]
    s.pos, s.vel, s.mass = pos, vel, mass
    return s


def _sim(scene, **kw):
    from granular_gpu_amd.device import AdaptiveSim
    return AdaptiveSim(scene, **kw)


def test_adaptive_stages_bitexact(gpu):
    import ctypes as C
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import _ptr, _stream
    _log("adaptive stages on a 20^3 block")
    s = _pile()
    sim = _sim(s)
    sim.grid_build()  # sorted state + cell table on the GPU
    torch.cuda.synchronize()
    n = sim.n
    pos = sim.pos[:n].cpu().numpy().copy()
    vel = sim.vel[:n].cpu().numpy().copy()
    mass = sim.mass[:n].cpu().numpy().copy()
    cs = sim.cs.cpu().numpy()
    p = po.params_from(s, 3, 5)
    lib = po.load()
    # surface
    _lib.check("gg_surface", sim.lib.gg_surface(C.byref(sim.set), _ptr(sim.cs),
                                                C.byref(sim.params), C.byref(sim.ws), _stream()))
    cover = np.zeros((n, 3), np.float32)
    surf = np.zeros(n, np.int32)
    lib.oracle_surface(n, pos.reshape(-1), mass, cs, p, cover.reshape(-1), surf)
    gsurf = sim.surface[:n].cpu().numpy()
    assert 0 < surf.sum() < n
    assert np.array_equal(gsurf, surf)
    # distance
    _lib.check("gg_surface_distance", sim.lib.gg_surface_distance(
        C.byref(sim.set), _ptr(sim.cs), C.byref(sim.ws), _stream()))
    sd = np.zeros(n, np.float32)
    lib.oracle_surface_distance(n, pos.reshape(-1), surf, cs, p, sd)
    assert np.array_equal(sim.sd[:n].cpu().numpy(), sd)
    assert (sd >= 100).sum() > 0, "pile has no deep particles"
    # merge mark
    _lib.check("gg_merge_mark", sim.lib.gg_merge_mark(C.byref(sim.set), C.byref(sim.ad),
                                                      _ptr(sim.cs), 4.0, C.byref(sim.ws),
                                                      _stream()))
    rem = np.zeros(n, np.int32)
    mer = np.zeros(n, np.float32)
    mv = np.zeros((n, 3), np.float32)
    lib.oracle_merge_mark(n, pos.reshape(-1), mass, vel.reshape(-1), surf, sd, rem, mer,
                          mv.reshape(-1), cs, p, 4.0, None, None)
    assert (rem == -1).sum() > 0
    assert np.array_equal(sim.remove[:n].cpu().numpy(), rem)
    assert np.array_equal(sim.merge[:n].cpu().numpy(), mer)
    assert np.array_equal(sim.mv[:n].cpu().numpy(), mv)
    # 40 blends
    mc = np.zeros(n, np.int32)
    adl = np.zeros(n, np.int32)
    for _ in range(40):
        _lib.check("gg_merge_blend", sim.lib.gg_merge_blend(C.byref(sim.set), C.byref(sim.ad), 40,
                                                            C.byref(sim.ws), _stream()))
        lib.oracle_merge_blend(n, mer, mass, vel.reshape(-1), mc, rem, mv.reshape(-1), 40, adl)
    assert (rem == 1).sum() > 0
    assert np.array_equal(sim.remove[:n].cpu().numpy(), rem)
    assert np.array_equal(sim.mass[:n].cpu().numpy(), mass)
    assert np.array_equal(sim.vel[:n].cpu().numpy(), vel)
    # make some surface particles heavy so that the split path runs
    heavy = np.nonzero(surf == 1)[0][::7]
    mass[heavy] = 4.0
    sim.mass[:n] = torch.from_numpy(mass).cuda()
    st = (C.c_int * 3)()
    bpos = sim.bpos.cpu().numpy()
    _lib.check("gg_split_compact", sim.lib.gg_split_compact(
        C.byref(sim.set), C.byref(sim.ad), _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs),
        C.byref(sim.params), C.byref(sim.ws), _stream(), st))
    cap = sim.cap
    P = np.zeros((cap, 3), np.float32); P[:n] = pos
    V = np.zeros((cap, 3), np.float32); V[:n] = vel
    M = np.zeros(cap, np.float32); M[:n] = mass
    A = np.zeros(cap, np.int32); A[:n] = sim_adl = adl
    R = np.zeros(cap, np.int32); R[:n] = rem
    ME = np.zeros(cap, np.float32); ME[:n] = mer
    MC = np.zeros(cap, np.int32); MC[:n] = mc
    ost = np.zeros(3, np.int32)
    n2 = lib.oracle_split_compact(n, cap, P.reshape(-1), V.reshape(-1), M, A, R, ME, MC, surf,
                                  None, np.ascontiguousarray(bpos).reshape(-1), cs,
                                  sim.bcs.cpu().numpy(), p, ost)
    _log(f"split stats gpu {list(st)} oracle {list(ost)} n {n} -> {n2}")
    assert sim.set.n == n2 and list(st) == list(ost) and ost[1] > 0
    torch.cuda.synchronize()
    assert np.array_equal(sim.pos[:n2].cpu().numpy(), P[:n2])
    assert np.array_equal(sim.vel[:n2].cpu().numpy(), V[:n2])
    assert np.array_equal(sim.mass[:n2].cpu().numpy(), M[:n2])
    assert np.array_equal(sim.adl[:n2].cpu().numpy(), A[:n2])
    assert np.array_equal(sim.remove[:n2].cpu().numpy(), R[:n2])
    assert np.array_equal(sim.merge[:n2].cpu().numpy(), ME[:n2])
    assert np.array_equal(sim.mcount[:n2].cpu().numpy(), MC[:n2])


@pytest.mark.parametrize("side,steps", [(16, 90)])
def test_adaptive_steps_bitexact(gpu, side, steps):
    _log(f"adaptive steps side={side} steps={steps}")
    s = scenes.plain_box(side, 1.0, dt=0.002, g=-9.8)
    sim = _sim(s, stab=3, solver=5)
    sim.step(steps)
    got = sim.host_all()
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    st = po.AdaptiveState(s.pos, s.vel, s.mass, cap=sim.cap)
    stats = po.adaptive_steps(st, bpos, bcs, p, steps)
    gstats = np.array(sim.stats, np.int32)
    _log(f"merged {stats[:, 0].sum()} split {stats[:, 1].sum()} final n {st.n}")
    assert stats[:, 0].sum() > 0
    assert np.array_equal(gstats, stats)
    assert sim.n == st.n
    n = st.n
    for k, ref in (("pos", st.pos), ("vel", st.vel), ("mass", st.mass), ("adl", st.adl),
                   ("remove", st.remove), ("merge", st.merge), ("mcount", st.mcount)):
        assert np.array_equal(got[k], ref[:n]), k


DEPTH = dict(min_mass=0.125, split_mass=0.125, merge_dist=1.75, merge_depth=3.0, split_depth=2.0)


def test_depth_rules_steps_bitexact_97k(gpu):
    # radius-level rules on a 46^3 block (97,336 particles) falling onto the
    # floor: deep particles merge up to mass 8, particles within two radii of
    # the surface split down to 0.125; whole adaptive steps against the oracle
    steps = 60
    _log(f"depth rules side=46 steps={steps}")
    s = scenes.plain_box(46, 1.0, dt=0.002, g=-9.8)
    sim = _sim(s, stab=3, solver=5, capacity=4 * s.n, max_mass=8.0)
    sim.params.adaptive_rules = 1
    for k, v in DEPTH.items():
        setattr(sim.params, k, v)
    sim.step(steps)
    got = sim.host_all()
    p = po.params_from(s, 3, 5, rules=dict(DEPTH, depth_rules=1))
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    st = po.AdaptiveState(s.pos, s.vel, s.mass, cap=sim.cap)
    stats = po.adaptive_steps(st, bpos, bcs, p, steps, max_mass=8.0)
    _log(f"merged {stats[:, 0].sum()} split {stats[:, 1].sum()} final n {st.n} (gpu {sim.n})")
    assert stats[:, 0].sum() > 0 and stats[:, 1].sum() > 0
    assert np.array_equal(np.array(sim.stats, np.int32), stats)
    assert sim.n == st.n
    n = st.n
    for k, ref in (("pos", st.pos), ("vel", st.vel), ("mass", st.mass), ("adl", st.adl),
                   ("remove", st.remove), ("merge", st.merge), ("mcount", st.mcount)):
        assert np.array_equal(got[k], ref[:n]), k
    # mass is conserved up to float rounding of the halvings and blends
    assert abs(float(st.mass[:n].sum(dtype=np.float64)) - float(s.mass.sum())) < 1e-3 * s.n
