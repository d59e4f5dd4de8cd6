"""Pins the CPU oracle of the plain step against what the mathematics fixes:
closed forms, library routines (stable sort, bincount/cumsum), brute force
on small inputs and invariants.  No GPU needed."""
import math

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po


@pytest.fixture(scope="module")
def lib():
    return po.load()


# ---------------------------------------------------------------- generators
def test_splitmix64_known_vector():
    # published SplitMix64 outputs for state 0 (Vigna's reference code)
    out = scenes.splitmix64(0, 3)
    assert [int(v) for v in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                     0x06C45D188009454F]


def test_box_walls_counts():
    # 2*ceil(1/0.0404)=50 per axis: 50*50*2 + 50*48*2 + 48*48*2
    assert scenes.box_walls((1, 1, 1)).shape == (14408, 3)
    assert scenes.box_walls((2.4, 2.4, 2.4)).shape == (84968, 3)
    w = scenes.box_walls((1, 1, 1))
    assert np.isclose(w[:, 1].min(), 0.005) and np.isclose(w[:, 1].max(), 0.995)


def test_plain_box_recipe():
    s = scenes.plain_box(32, 1.0)
    assert s.n == 32768 and s.cells == (25, 25, 25)
    lo = np.array([0.19, 0.125, 0.19])
    assert np.all(np.abs(s.pos.min(0) - lo) <= 0.0011)
    assert np.all(np.abs(s.pos.max(0) - (lo + 0.62)) <= 0.0011)
    m1 = scenes.plain_box(100, 2.4)
    assert m1.n == 1_000_000 and m1.cells == (60, 60, 60) and m1.boundary.shape[0] == 84968


# ---------------------------------------------------------------- grid build
def _hash_np(pos, L, cells):
    L = np.float32(L)
    f = np.floor(pos.astype(np.float32) / L)
    cx, cy, cz = cells
    ok = (f[:, 0] >= 0) & (f[:, 0] < cx) & (f[:, 1] >= 0) & (f[:, 1] < cy) & (f[:, 2] >= 0) & (
        f[:, 2] < cz)
    fi = np.where(ok[:, None], f, 0).astype(np.int64)
    key = (fi[:, 0] * cy + fi[:, 1]) * cz + fi[:, 2]
    return np.where(ok, key, cx * cy * cz).astype(np.int32)


def test_hash_special_cases(lib):
    L = np.float32(0.0404)
    cells = (25, 25, 25)
    pos = np.array([[0.0, 0.0, 0.0],              # corner cell 0
                    [L, 0.0, 0.0],                # exactly on a face -> next cell in x
                    [-1e-7, 0.5, 0.5],            # negative -> sentinel
                    [25 * L, 0.5, 0.5],           # x == cx*L -> sentinel
                    [np.nan, 0.5, 0.5],           # NaN -> sentinel
                    [0.5, np.inf, 0.5],           # inf -> sentinel
                    [0.999, 0.999, 0.999]],       # last cell
                   np.float32)
    keys = np.empty(len(pos), np.int32)
    lib.oracle_hash(pos.reshape(-1), len(pos), float(L), *cells, keys)
    C = 25 ** 3
    assert keys[0] == 0
    assert keys[1] == 1 * 25 * 25
    assert keys[2] == C and keys[3] == C and keys[4] == C and keys[5] == C
    assert keys[6] == C - 1


def test_grid_build_is_stable_counting_sort(lib):
    rng = np.random.default_rng(5)
    pos = rng.uniform(-0.05, 1.05, size=(20000, 3)).astype(np.float32)
    cells = (25, 25, 25)
    keys, perm, cs = po.grid_build(pos, 0.0404, cells)
    ref_keys = _hash_np(pos, 0.0404, cells)
    ref_perm = np.argsort(ref_keys, kind="stable")
    assert np.array_equal(perm, ref_perm)
    assert np.array_equal(keys, ref_keys[ref_perm])
    counts = np.bincount(ref_keys, minlength=25 ** 3 + 1)
    ref_cs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    assert np.array_equal(cs, ref_cs)


# ---------------------------------------------------------------- radii / exp
def test_radius_variants(lib):
    d = float(scenes.DENSITY_R001)
    for m in (0.5, 1.0, 2.0, 4.0):
        exact_int = (3 * m / (4 * math.pi * int(d))) ** (1 / 3)
        exact_flt = (3 * m / (4 * math.pi * float(np.float32(d)))) ** (1 / 3)
        for fn, ex in ((lib.oracle_radius_contact, exact_int),
                       (lib.oracle_radius_surface, exact_flt),
                       (lib.oracle_radius_split, exact_flt)):
            r = fn(m, d)
            assert abs(r - ex) <= 2 * np.spacing(np.float32(ex)), (fn, m, r, ex)
    assert abs(lib.oracle_radius_contact(1.0, d) - 0.01) < 1e-8


def test_expf_correctly_rounded(lib):
    xs = np.linspace(-5, 5, 257).astype(np.float32)
    for x in xs:
        assert lib.oracle_expf(float(x)) == np.float32(math.exp(float(x)))


# ---------------------------------------------------------------- dynamics
def _one_particle_scene(pos, vel, box=1.0, walls=True):
    s = scenes.plain_box(2, box)
    s.pos = np.asarray(pos, np.float32).reshape(-1, 3)
    s.vel = np.asarray(vel, np.float32).reshape(-1, 3)
    s.mass = np.ones(s.pos.shape[0], np.float32)
    if not walls:
        s.boundary = np.zeros((0, 3), np.float32)
    return s


def _run(s, steps, stab=3, solver=5, mu_s=0.0, mu_k=0.0, f64=False, g=None):
    p = po.params_from(s, stab, solver, mu_s, mu_k)
    if g is not None:
        p.gy = g
    if s.boundary.shape[0]:
        bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    else:
        bpos = np.zeros((0, 3), np.float32)
        bcs = np.zeros(s.cells[0] * s.cells[1] * s.cells[2] + 2, np.int32)
    if f64:
        return po.plain_steps_f64(s.pos, s.vel, s.mass, bpos, bcs, p, steps)
    return po.plain_steps_f32(s.pos, s.vel, s.mass, bpos, bcs, p, steps)


def test_free_fall_closed_form():
    s = _one_particle_scene([0.5, 0.6, 0.5], [0, 0, 0])
    dt = float(s.dt)
    g = float(s.gravity[1])
    for f64 in (False, True):
        pos, vel = _run(s, 10, f64=f64)[:2]
        k = 10
        y = 0.6 + dt * dt * g * k * (k + 1) / 2
        assert abs(pos[0, 1] - y) < 2e-7
        assert abs(vel[0, 1] - k * dt * g) < 1e-5
        assert pos[0, 0] == np.float32(0.5) and pos[0, 2] == np.float32(0.5)


def test_damping_and_sleep_closed_form():
    s = _one_particle_scene([[0.3, 0.5, 0.5], [0.7, 0.5, 0.5]], [[5.0, 0, 0], [0.005, 0, 0]])
    pos, vel = _run(s, 1, g=0.0)[:2]
    # fast particle: v *= 0.99 + 0.01*exp(-0.1|v|) (Solver.cu:227-243)
    f = 0.99 + 0.01 * math.exp(-0.5)
    assert abs(vel[0, 0] - 5.0 * f) < 5e-6
    assert abs(pos[0, 0] - (0.3 + float(s.dt) * 5.0)) < 1e-7
    # slow particle: speed < 0.01 -> velocity zeroed, position still advanced
    assert vel[1, 0] == 0.0
    assert abs(pos[1, 0] - (0.7 + float(s.dt) * 0.005)) < 1e-7


def test_two_particle_overlap_resolved_in_one_sweep():
    # head-on overlap of 0.004: symmetric gather moves each by half (n = 2)
    s = _one_particle_scene([[0.49, 0.5, 0.5], [0.506, 0.5, 0.5]], [[0, 0, 0]] * 2,
                            walls=False)
    for f64 in (False, True):
        pos = _run(s, 1, stab=1, solver=0, g=0.0, f64=f64)[0]
        d = pos[1, 0] - pos[0, 0]
        r = (3 / (4 * math.pi * 238732)) ** (1 / 3)
        assert abs(d - 2 * r) < 1e-7
        assert abs((pos[0, 0] + pos[1, 0]) / 2 - 0.498) < 1e-7


def test_particle_on_floor_particle():
    # floor particle of the BOX recipe at (x_w, 0.005, z_w); drop a particle
    # 0.015 above it: pushed to 0.01 + r_i from the wall particle.
    s = _one_particle_scene([0, 0, 0], [0, 0, 0])
    w = s.boundary
    k = np.argmin(np.abs(w[:, 0] - 0.5) + np.abs(w[:, 2] - 0.5) + w[:, 1])
    wx, wy, wz = (float(v) for v in w[k])
    s.pos = np.array([[wx, wy + 0.015, wz]], np.float32)
    pos = _run(s, 1, stab=1, solver=0, g=0.0)[0]
    r = (3 / (4 * math.pi * 238732)) ** (1 / 3)
    assert abs((pos[0, 1] - wy) - (0.01 + max(r, 0.01))) < 1e-7
    assert abs(pos[0, 0] - wx) < 1e-9 and abs(pos[0, 2] - wz) < 1e-9


def test_static_friction_stops_sliding_on_floor():
    s = _one_particle_scene([0, 0, 0], [0, 0, 0])
    w = s.boundary
    k = np.argmin(np.abs(w[:, 0] - 0.5) + np.abs(w[:, 2] - 0.5) + w[:, 1])
    wx, wy, wz = (float(v) for v in w[k])
    s.pos = np.array([[wx, wy + 0.0195, wz]], np.float32)
    s.vel = np.array([[0.05, 0.0, 0.0]], np.float32)
    pos_f = _run(s, 1, stab=0, solver=1, mu_s=10.0, mu_k=10.0, g=0.0)[0]
    pos_n = _run(s, 1, stab=0, solver=1, g=0.0)[0]
    pos_k = _run(s, 1, stab=0, solver=1, mu_s=0.0, mu_k=0.1, g=0.0)[0]
    # closed form of one solver sweep with the single floor contact:
    # x* = x + D, p = x* - x_b, x' = x* + ((rc-d)/d) p - f, f = friction of D
    dx = float(np.float32(s.dt)) * 0.05
    D = np.array([dx, 0.0, 0.0])
    pvec = np.array([wx + dx - wx, 0.0195, 0.0])
    pvec[1] = float(np.float32(wy + 0.0195)) - wy
    d = np.linalg.norm(pvec)
    rc = 0.02
    nrm = pvec / d
    t = D - D.dot(nrm) * nrm
    x_star = np.array([wx + dx, wy + pvec[1], wz])
    base = x_star + (rc - d) / d * pvec
    exp_n = base
    exp_f = base - t                                   # static: |t| < mu_s*delta
    c = min(0.1 * (rc - d) / np.linalg.norm(t), 1.0)   # kinetic with mu_k = 0.1
    exp_k = base - t * c
    assert np.abs(pos_n[0] - exp_n).max() < 1e-7
    assert np.abs(pos_f[0] - exp_f).max() < 1e-7
    assert np.abs(pos_k[0] - exp_k).max() < 1e-7
    # friction opposes the sliding direction
    assert pos_f[0, 0] < pos_k[0, 0] < pos_n[0, 0]


def test_thread_count_independent():
    lib = po.load()
    s = scenes.plain_box(12, 1.0)
    old = lib.oracle_get_threads()
    lib.oracle_set_threads(1)
    a = _run(s, 3)
    lib.oracle_set_threads(max(old, 4))
    b = _run(s, 3)
    lib.oracle_set_threads(old)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)


def _shadow(s, steps, stab=3, solver=10, mu_s=0.0, mu_k=0.0):
    p = po.params_from(s, stab, solver, mu_s, mu_k)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    return po.shadow_steps(s.pos, s.vel, s.mass, bpos, bcs, p, steps)


def test_shadow_fp32_half_is_the_free_fp32_run():
    s = scenes.plain_box(12, 1.0)
    p32, v32, _, _ = _shadow(s, 3, 3, 10, 0.5, 0.4)
    q32, w32 = _run(s, 3, 3, 10, 0.5, 0.4)[:2]
    assert np.array_equal(p32, q32) and np.array_equal(v32, w32)


@pytest.mark.parametrize("fric", [(0.0, 0.0), (0.5, 0.4)])
def test_fp32_within_tolerance_of_fp64(fric):
    """1e-5 r after 1 step, 1e-3 r after 10 steps (r = 0.01)."""
    s = scenes.plain_box(16, 1.0)
    p32, _, p64, _ = _shadow(s, 1, 3, 10, *fric)
    assert np.abs(p32 - p64).max() <= 1e-5 * 0.01
    p32, _, p64, _ = _shadow(s, 10, 3, 10, *fric)
    assert np.abs(p32 - p64).max() <= 1e-3 * 0.01


def test_fp64_free_run_close_to_fp32():
    """The independent FP64 run may take other discrete decisions at exact
    contact-threshold ties; it still agrees closely for almost every
    particle (compared by particle identity)."""
    s = scenes.plain_box(16, 1.0)
    p = po.params_from(s, 3, 10)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    p64, _, _, ids = po.plain_steps_f64(s.pos, s.vel, s.mass, bpos, bcs, p, 1)
    p32, _, p64s, _ = po.shadow_steps(s.pos, s.vel, s.mass, bpos, bcs, p, 1)
    keys, perm, _ = po.grid_build(s.pos, s.cell_length, s.cells)
    by_id32 = np.empty_like(p32)
    by_id32[perm] = p32
    by_id64 = np.empty_like(p64)
    by_id64[ids] = p64
    err = np.abs(by_id32 - by_id64).max(1)
    assert np.mean(err <= 1e-7) > 0.99
    assert err.max() < 1e-4


def test_nlist_equals_brute_force(lib):
    s = scenes.plain_box(10, 1.0)
    n = s.n
    p = po.params_from(s, skin=0.005)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    keys, perm, cs = po.grid_build(s.pos, s.cell_length, s.cells)
    pos = s.pos[perm]
    r = lib.oracle_radius_contact(1.0, float(s.density))
    base4 = np.concatenate([pos, np.full((n, 1), r, np.float32)], 1).astype(np.float32)
    off4 = np.zeros((n, 4), np.float32)
    off4[:, 1] = np.float32(float(s.dt) * float(s.dt) * float(s.gravity[1]))
    gmax, bmax = 64, 32
    gc = np.empty(n, np.int32); gl = np.empty(n * gmax, np.int32)
    bc = np.empty(n, np.int32); bl = np.empty(n * bmax, np.int32)
    base = np.empty(n, np.int32)
    lib.oracle_nlist_f32(n, base4.reshape(-1), off4.reshape(-1), bpos.shape[0],
                         bpos.reshape(-1), cs, bcs, p, gmax, bmax, gc, gl, bc, bl, base)
    xs = (pos.astype(np.float64) + off4[:, :3])
    d = np.linalg.norm(xs[:, None, :] - xs[None, :, :], axis=2)
    np.fill_diagonal(d, np.inf)
    for i in range(0, n, 37):
        brute = set(np.nonzero(d[i] < 2 * r + 0.005)[0].tolist())
        got = set(gl[i * gmax: i * gmax + gc[i]].tolist())
        assert got == brute, i
        assert list(gl[i * gmax: i * gmax + gc[i]]) == sorted(got)
