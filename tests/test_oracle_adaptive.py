"""Pins the adaptive merge/split oracle (surface, surface distance, merge
mark/blend, split, compaction) with hand-worked cases, brute force and
conservation laws.  No GPU needed."""
import ctypes as C
import math

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def lib():
    return po.load()


def _scene(pos):
    s = scenes.plain_box(2, 1.0)
    s.pos = np.asarray(pos, np.float32).reshape(-1, 3)
    s.vel = np.zeros_like(s.pos)
    s.mass = np.ones(s.pos.shape[0], np.float32)
    return s


def _sorted(s):
    keys, perm, cs = po.grid_build(s.pos, s.cell_length, s.cells)
    return np.ascontiguousarray(s.pos[perm]), cs, perm


def _surface(lib, s, pos, cs, mass=None):
    n = pos.shape[0]
    mass = np.ones(n, np.float32) if mass is None else mass
    cover = np.zeros((n, 3), np.float32)
    surf = np.zeros(n, np.int32)
    lib.oracle_surface(n, pos.reshape(-1), mass, cs, po.params_from(s), cover.reshape(-1), surf)
    return cover, surf


def _cube(side, jitter, seed=3):
    return scenes.jittered_lattice(side, side, side, 0.02, (0.3, 0.3, 0.3), jitter, seed)


def test_surface_isolated_and_symmetric(lib):
    # isolated particle: no neighbours (n <= 5) -> surface
    s = _scene([[0.5, 0.5, 0.5]])
    pos, cs, _ = _sorted(s)
    assert _surface(lib, s, pos, cs)[1][0] == 1
    # centre with six axis neighbours at 2^-6: the normalised vectors are
    # exactly +-e_k, the cover vector sums to exactly 0, its normalisation is
    # NaN and no neighbour can vote -> surface although n = 6 > 5 (quirk)
    h = 2.0 ** -6
    off = [[0, 0, 0], [h, 0, 0], [-h, 0, 0], [0, h, 0], [0, -h, 0], [0, 0, h], [0, 0, -h]]
    s = _scene(0.5 + np.array(off))
    pos, cs, perm = _sorted(s)
    cover, surf = _surface(lib, s, pos, cs)
    centre = np.nonzero(np.all(np.isclose(pos, 0.5), axis=1))[0][0]
    assert np.all(cover[centre] == 0.0)
    assert surf[centre] == 1


def test_surface_jittered_block(lib):
    s = _scene(_cube(8, 0.001))
    pos, cs, perm = _sorted(s)
    _, surf = _surface(lib, s, pos, cs)
    # lattice coordinates of each sorted particle
    ijk = np.rint((pos - 0.3) / 0.02).astype(int)
    depth = np.minimum(ijk, 7 - ijk).min(1)
    assert np.all(surf[depth == 0] == 1)
    assert np.all(surf[depth >= 2] == 0)


def test_surface_distance_brute_force(lib):
    s = _scene(_cube(8, 0.001))
    pos, cs, perm = _sorted(s)
    _, surf = _surface(lib, s, pos, cs)
    n = pos.shape[0]
    sd = np.zeros(n, np.float32)
    lib.oracle_surface_distance(n, pos.reshape(-1), surf, cs, po.params_from(s), sd)
    L = float(s.cell_length)
    sp = pos[surf == 1].astype(np.float64)
    assert np.all(sd[surf == 1] == 0.0)
    for i in np.nonzero(surf == 0)[0]:
        d2 = ((sp - pos[i]) ** 2).sum(1).min()
        if d2 < L * L:  # nearest surface particle is inside the stencil
            assert abs(sd[i] - d2) <= 1e-7, (i, sd[i], d2)
        elif d2 > 4 * L * L:  # no surface particle can be in the stencil
            assert sd[i] == np.float32(1e5)


DEPTH = dict(depth_rules=1, min_mass=0.125, split_mass=0.125, merge_dist=1.75, merge_depth=3.0,
             split_depth=2.0)


def _merge(lib, s, pos, surf, sd, remove=None, mass=None, vel=None, rules=None, max_mass=4.0):
    n = pos.shape[0]
    _, _, cs = po.grid_build(pos, s.cell_length, s.cells)
    mass = np.ones(n, np.float32) if mass is None else np.asarray(mass, np.float32)
    vel = np.zeros((n, 3), np.float32) if vel is None else np.asarray(vel, np.float32)
    remove = np.zeros(n, np.int32) if remove is None else np.asarray(remove, np.int32)
    merge = np.zeros(n, np.float32)
    mv = np.zeros((n, 3), np.float32)
    cand = np.zeros(n, np.int32)
    dens = np.zeros(n, np.float32)
    lib.oracle_merge_mark(n, pos.reshape(-1), mass, vel.reshape(-1),
                          np.asarray(surf, np.int32), np.asarray(sd, np.float32), remove, merge,
                          mv.reshape(-1), cs, po.params_from(s, rules=rules), max_mass,
                          cand.ctypes.data, dens.ctypes.data)
    return remove, merge, mv, cand, dens


def test_merge_line_of_three(lib):
    # three deep particles 0.02 apart along z in one cell column (sorted by
    # index = z order): 0 -> 1 pairs; 1 is then taken; 2 has no j > 2
    s = _scene([[0.51, 0.51, 0.45], [0.51, 0.51, 0.47], [0.51, 0.51, 0.49]])
    pos, cs, perm = _sorted(s)
    assert np.all(np.diff(pos[:, 2]) > 0)
    vel = np.array([[1, 0, 0], [0, 2, 0], [0, 0, 3]], np.float32)
    rem, merge, mv, cand, dens = _merge(lib, s, pos, [0, 0, 0], [1e5] * 3, vel=vel)
    assert list(cand) == [1, 2, -1]
    assert list(rem) == [-1, -1, 0]
    assert list(merge) == [-1.0, 1.0, 0.0]
    vt = 0.5 * (vel[0] + vel[1])
    # local density of particle 0: itself and particle 1 within 0.035
    boost = 1.0 + (1.0 - min(1.0, 2.0 / 2 / 100.0)) * 0.1
    assert np.allclose(mv[0], -vt) and np.allclose(mv[1], vt * boost, rtol=1e-6)


def test_merge_two_claimers_one_target(lib):
    # 0 and 1 both pick 2 (the closest higher index); 0 wins, 1 releases
    s = _scene([[0.51, 0.51, 0.45], [0.51, 0.51, 0.452 + 0.03], [0.51, 0.51, 0.47]])
    pos, cs, perm = _sorted(s)
    d = lambda a, b: np.linalg.norm(pos[a] - pos[b])
    assert d(0, 1) > 0.035 or d(0, 2) < d(0, 1)
    rem, merge, mv, cand, dens = _merge(lib, s, pos, [0, 0, 0], [1e5] * 3)
    assert cand[0] == 2 and cand[1] == 2
    assert list(rem) == [-1, 0, -1]


def test_merge_partner_lock_quirk(lib):
    # the candidate test checks sd of i, not of j (Solver.cu:578): a target
    # with sd <= 100 is locked (remove = -1) without any merge data and i is
    # released (Solver.cu:609-649)
    s = _scene([[0.51, 0.51, 0.45], [0.51, 0.51, 0.47]])
    pos, cs, perm = _sorted(s)
    rem, merge, mv, cand, dens = _merge(lib, s, pos, [0, 0], [1e5, 50.0])
    assert cand[0] == 1
    assert list(rem) == [0, -1] and list(merge) == [0.0, 0.0]


def test_merge_respects_mass_limit_and_eligibility(lib):
    s = _scene([[0.51, 0.51, 0.45], [0.51, 0.51, 0.47]])
    pos, cs, perm = _sorted(s)
    rem, merge, *_ = _merge(lib, s, pos, [0, 0], [1e5, 1e5], mass=[2.5, 2.0])
    assert list(rem) == [0, 0]                     # 4.5 > max mass
    rem, merge, *_ = _merge(lib, s, pos, [1, 0], [1e5, 1e5])
    assert list(rem) == [0, 0]                     # surface particles do not merge
    rem, merge, *_ = _merge(lib, s, pos, [0, 0], [1e5, 1e5], remove=[-1, 0])
    assert list(rem) == [-1, 0]                    # already merging


def test_blend_conserves_mass_and_removes_donor(lib):
    n = 2
    merge = np.array([-1.0, 1.0], np.float32)
    mass = np.ones(n, np.float32)
    vel = np.zeros((n, 3), np.float32)
    mc = np.zeros(n, np.int32)
    rem = np.array([-1, -1], np.int32)
    mv = np.array([[-1, 0, 0], [1, 0, 0]], np.float32)
    adl = np.zeros(n, np.int32)
    for k in range(40):
        lib.oracle_merge_blend(n, merge, mass, vel.reshape(-1), mc, rem, mv.reshape(-1), 40, adl)
        assert abs(mass.sum() - 2.0) < 1e-5
        if k < 39:
            assert list(rem) == [-1, -1] and list(mc) == [k + 1] * 2
    assert list(rem) == [1, 0] and list(mc) == [0, 0]
    assert abs(mass[1] - 2.0) < 1e-5 and abs(mass[0]) < 1e-5
    assert list(merge) == [0.0, 0.0] and list(adl) == [0, 1]
    assert abs(vel[1, 0] - 1.0) < 1e-6


def _split(lib, s, pos, mass, surf, remove, bpos=None, cap=None, sd=None, rules=None):
    n = pos.shape[0]
    cap = cap or 2 * n
    P = np.zeros((cap, 3), np.float32); P[:n] = pos
    V = np.zeros((cap, 3), np.float32); V[:n] = 0.25
    M = np.zeros(cap, np.float32); M[:n] = mass
    adl = np.zeros(cap, np.int32)
    R = np.zeros(cap, np.int32); R[:n] = remove
    mer = np.zeros(cap, np.float32)
    mc = np.zeros(cap, np.int32)
    _, _, cs = po.grid_build(pos, s.cell_length, s.cells)
    if bpos is None:
        bpos = np.zeros((0, 3), np.float32)
        bcs = np.zeros_like(cs)
    else:
        bpos, bcs = po.sorted_boundary(bpos, s.cell_length, s.cells)
    stats = np.zeros(3, np.int32)
    sdv = None if sd is None else np.ascontiguousarray(sd, np.float32)
    n2 = lib.oracle_split_compact(n, cap, P.reshape(-1), V.reshape(-1), M, adl, R, mer, mc,
                                  np.asarray(surf, np.int32),
                                  None if sdv is None else sdv.ctypes.data, bpos.reshape(-1), cs,
                                  bcs, po.params_from(s, rules=rules), stats)
    return n2, P[:n2], V[:n2], M[:n2], adl[:n2], R[:n2], stats


def test_split_child_in_first_empty_cell(lib):
    s = _scene([[0.5, 0.5, 0.5]])
    L = float(s.cell_length)
    n2, P, V, M, adl, R, st = _split(lib, s, s.pos, [4.0], [1], [0])
    assert n2 == 2 and list(st) == [0, 1, 0]
    b = np.floor(np.float32(0.5) / np.float32(L))
    exp = (np.float32(b - 1) + np.float32(0.5)) * np.float32(L)   # stencil cell m = 0
    assert np.all(P[1] == np.float32(exp))
    assert list(M) == [2.0, 2.0] and list(adl) == [1, 1]
    assert np.all(V[1] == V[0])
    # a boundary particle within 2r cancels the split
    r = (3 * 4.0 / (4 * math.pi * 238732.4146)) ** (1 / 3)
    n2, *_ = _split(lib, s, s.pos, [4.0], [1], [0], bpos=np.array([[0.5 + 1.5 * r, 0.5, 0.5]],
                                                                  np.float32))
    assert n2 == 1
    # not on the surface, too light, or merging: no split
    for m, sf, rm in ((4.0, 0, 0), (2.0, 1, 0), (4.0, 1, -1)):
        assert _split(lib, s, s.pos, [m], [sf], [rm])[0] == 1


def test_compaction_keeps_order_and_appends_children(lib):
    s = _scene([[0.2, 0.2, 0.2], [0.4, 0.4, 0.4], [0.6, 0.6, 0.6], [0.8, 0.8, 0.8]])
    pos, cs, perm = _sorted(s)
    n2, P, V, M, adl, R, st = _split(lib, s, pos, [1.0, 4.0, 1.0, 4.0], [0, 1, 0, 1],
                                     [0, 0, 1, 0])
    # particle 2 removed; 1 and 3 split (children in parent order)
    assert list(st) == [1, 2, 0] and n2 == 5
    assert np.array_equal(P[:3], pos[[0, 1, 3]])
    assert list(M) == [1.0, 2.0, 2.0, 2.0, 2.0]
    # capacity limit: only room for one child -> the second candidate does not split
    n2, P, V, M, adl, R, st = _split(lib, s, pos, [1.0, 4.0, 1.0, 4.0], [0, 1, 0, 1],
                                     [0, 0, 1, 0], cap=4)
    assert n2 == 4 and list(st) == [1, 1, 1]
    assert list(M) == [1.0, 2.0, 4.0, 2.0]


def _r_surface(m):
    """The surface radius of mass m (float density, double pi, cbrt in
    double rounded once): what the depth rules scale by."""
    return np.float32((3.0 * float(np.float32(3.0 * m)) / 3.0 /
                       (4.0 * math.pi * float(np.float32(238732.4146)))) ** (1.0 / 3.0))


def test_depth_rules_merge_below_three_radii(lib):
    # two unit-mass particles (r = 0.01) 0.02 apart: partners within
    # 1.75 (r_i + r_j) = 0.035; merging needs both more than 3 r = 0.03
    # below the surface (sd > (3 r)^2, sd a squared distance)
    s = _scene([[0.51, 0.51, 0.45], [0.51, 0.51, 0.47]])
    pos, cs, perm = _sorted(s)
    r = float(_r_surface(1.0))
    assert abs(r - 0.01) < 1e-6
    deep, shallow = (3 * r * 1.01) ** 2, (3 * r * 0.99) ** 2
    rem, merge, *_ = _merge(lib, s, pos, [0, 0], [deep, deep], rules=DEPTH, max_mass=8.0)
    assert list(rem) == [-1, -1] and list(merge) == [-1.0, 1.0]
    rem, merge, *_ = _merge(lib, s, pos, [0, 0], [shallow, deep], rules=DEPTH, max_mass=8.0)
    assert list(rem) == [0, 0]                     # i too close to the surface
    rem, merge, *_ = _merge(lib, s, pos, [0, 0], [deep, shallow], rules=DEPTH, max_mass=8.0)
    assert list(rem) == [0, -1]                    # partner locked (reference quirk kept)
    # the reference rules would not merge at sd = 0.0009 (< 100)
    rem, *_ = _merge(lib, s, pos, [0, 0], [deep, deep])
    assert list(rem) == [0, 0]


def test_depth_rules_partner_distance_scales_with_radii(lib):
    # partners within 1.75 (r_i + r_j): 0.035 for two unit masses, 0.0556 for
    # two of mass 4 (r = 0.0159); at 0.036 (inside the neighbouring cell)
    # only the heavier pair merges; the merged mass may reach max_mass
    # (8 + 1 > 8 refused, 7 + 1 = 8 allowed)
    for dz, mass, pair in ((0.036, [4.0, 4.0], True), (0.036, [1.0, 1.0], False),
                           (0.030, [1.0, 1.0], True), (0.030, [8.0, 1.0], False),
                           (0.030, [7.0, 1.0], True)):
        s = _scene([[0.51, 0.51, 0.40], [0.51, 0.51, 0.40 + dz]])
        pos, cs, perm = _sorted(s)
        ri, rj = (float(_r_surface(m)) for m in mass)
        assert (np.linalg.norm(pos[0] - pos[1]) < 1.75 * (ri + rj)) == (pair or sum(mass) > 8)
        rem, *_ = _merge(lib, s, pos, [0, 0], [1e5, 1e5], mass=mass, rules=DEPTH, max_mass=8.0)
        assert list(rem) == ([-1, -1] if pair else [0, 0]), (dz, mass)


def test_depth_rules_split_within_two_radii(lib):
    # a particle of mass 1 (r = 0.01) splits when it is closer than 2 r to
    # the surface (sd < (2 r)^2), surface flag or not; mass 0.25 is the
    # lightest that splits (children of 0.125, the finest level)
    s = _scene([[0.5, 0.5, 0.5]])
    r = float(_r_surface(1.0))
    near, far = (2 * r * 0.99) ** 2, (2 * r * 1.01) ** 2
    n2, P, V, M, *_ = _split(lib, s, s.pos, [1.0], [0], [0], sd=[near], rules=DEPTH)
    assert n2 == 2 and list(M) == [0.5, 0.5]
    assert _split(lib, s, s.pos, [1.0], [1], [0], sd=[far], rules=DEPTH)[0] == 1
    assert _split(lib, s, s.pos, [0.25], [1], [0], sd=[0.0], rules=DEPTH)[0] == 2
    assert _split(lib, s, s.pos, [0.125], [1], [0], sd=[0.0], rules=DEPTH)[0] == 1


def test_adaptive_steps_conserve_mass():
    s = scenes.plain_box(14, 1.0, dt=0.002, g=-9.8)
    p = po.params_from(s, 3, 5)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    st = po.AdaptiveState(s.pos, s.vel, s.mass, cap=2 * s.n)
    total0 = st.mass[: st.n].sum(dtype=np.float64)
    stats = po.adaptive_steps(st, bpos, bcs, p, 120)
    assert stats[:, 0].sum() > 0, "no merges happened"
    # merges lose only the rounding of m/40 blends; splits are exact halves
    total = st.mass[: st.n].sum(dtype=np.float64)
    assert abs(total - total0) / total0 < 1e-4
    assert st.n == s.n - stats[:, 0].sum() + stats[:, 1].sum()
    assert np.isfinite(st.pos[: st.n]).all()


def test_regular_lattice_interior_is_not_surface(lib):
    # normalize uses an approximate rsqrt (as the reference's device rsqrtf
    # is): for a regular lattice the +-axis unit vectors of an interior
    # particle then do not cancel exactly, the cover vector is tiny but
    # nonzero and some axis neighbour lies within 60 degrees of it.  With a
    # correctly rounded 1/sqrt they cancel exactly and every particle would
    # be flagged surface (NaN vote).
    side = 6
    pos = np.array([[0.3 + 0.032 * i, 0.3 + 0.032 * j, 0.3 + 0.032 * k]
                    for i in range(side) for j in range(side) for k in range(side)], np.float32)
    s = _scene(pos)
    spos, cs, perm = _sorted(s)
    cover, surf = _surface(lib, s, spos, cs)
    q = np.rint((spos - 0.3) / 0.032).astype(int)
    interior = np.all((q > 0) & (q < side - 1), axis=1)
    shell = ~interior
    assert interior.sum() == (side - 2) ** 3
    assert surf[shell].all()                       # n <= 5 or an open side
    assert surf[interior].mean() < 0.5, surf[interior].mean()
    nz = np.any(cover[interior] != 0, axis=1)
    assert nz.mean() > 0.5
    # cover vectors of interior particles are rounding-sized
    assert np.abs(cover[interior]).max() < 1e-4
