"""GPU parity of the plain step: the HIP path through the C-ABI must equal
the FP32 oracle bit for bit (cell ids, sort order, neighbour lists,
positions, velocities) and stay within tolerance of the decision-shadowed
FP64 oracle."""
import sys

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim(scene, **kw):
    from granular_gpu_amd.device import PlainSim
    return PlainSim(scene, **kw)


def _oracle_steps(scene, steps, stab, solver, mu_s=0.0, mu_k=0.0):
    p = po.params_from(scene, stab, solver, mu_s, mu_k)
    bpos, bcs = po.sorted_boundary(scene.boundary, scene.cell_length, scene.cells)
    return po.plain_steps_f32(scene.pos, scene.vel, scene.mass, bpos, bcs, p, steps)


def test_device_is_gfx950(gpu):
    assert gpu.gg_device_check() == 0


@pytest.mark.parametrize("case", ["lattice", "random_outside", "dense_cells", "huge_cells"])
def test_grid_build_bitexact(gpu, case):
    import torch
    _log(f"grid_build case={case}")
    s = scenes.plain_box(20, 1.0)
    rng = np.random.default_rng(7)
    if case == "random_outside":
        pos = rng.uniform(-0.1, 1.1, size=(30000, 3)).astype(np.float32)
        pos[::97, 0] = np.nan
        pos[5::101, 1] = np.inf
    elif case == "dense_cells":
        # one cell with 5000 (rank-sort path), cells with 40..2000 (LDS path)
        a = np.full((5000, 3), 0.5, np.float32) + rng.uniform(0, 0.005, (5000, 3)).astype(
            np.float32)
        b = np.full((2000, 3), 0.1, np.float32) + rng.uniform(0, 0.005, (2000, 3)).astype(
            np.float32)
        c = rng.uniform(0.2, 0.3, size=(3000, 3)).astype(np.float32)
        pos = np.concatenate([a, b, c])[rng.permutation(10000)]
    elif case == "huge_cells":
        # three cells of more than 4096 particles (one the sentinel) whose
        # members sit in different index ranges, with scattered strays
        a = np.full((6000, 3), 0.41, np.float32) + rng.uniform(0, 0.005, (6000, 3)).astype(
            np.float32)
        b = np.full((9000, 3), 0.77, np.float32) + rng.uniform(0, 0.005, (9000, 3)).astype(
            np.float32)
        out = rng.uniform(1.2, 1.5, size=(8000, 3)).astype(np.float32)
        rest = rng.uniform(0.0, 1.0, size=(7000, 3)).astype(np.float32)
        pos = np.concatenate([rest[:2000], a, rest[2000:4000], b, out, rest[4000:]])
        swap = rng.choice(pos.shape[0], size=600, replace=False)
        pos[swap] = pos[swap[::-1]]
    else:
        pos = s.pos
    s.pos = pos
    s.vel = rng.normal(size=pos.shape).astype(np.float32)
    s.mass = rng.uniform(0.5, 4, size=pos.shape[0]).astype(np.float32)
    sim = _sim(s)
    perm = torch.zeros(pos.shape[0], dtype=torch.int32, device="cuda")
    sim.grid_build(perm)
    torch.cuda.synchronize()
    keys, operm, cs = po.grid_build(pos, s.cell_length, s.cells)
    assert np.array_equal(perm.cpu().numpy(), operm)
    assert np.array_equal(sim.p2c[: s.n].cpu().numpy(), keys)
    assert np.array_equal(sim.cs.cpu().numpy(), cs)
    gp, gv, gm = sim.host()
    assert np.array_equal(gp, pos[operm], equal_nan=True)
    assert np.array_equal(gv, s.vel[operm])
    assert np.array_equal(gm, s.mass[operm])


@pytest.mark.parametrize("steps,stab,solver,fric,lists", [
    (1, 3, 5, (0.0, 0.0), True),
    (1, 3, 10, (0.5, 0.4), True),
    (10, 3, 10, (0.0, 0.0), True),
    (10, 3, 10, (0.5, 0.4), True),
    (10, 3, 10, (0.5, 0.4), False),
    (3, 0, 4, (0.0, 0.0), True),
    (3, 2, 0, (0.0, 0.0), True),
])
def test_step_bitexact_vs_fp32_oracle(gpu, steps, stab, solver, fric, lists):
    _log(f"step steps={steps} S={stab} I={solver} fric={fric} lists={lists}")
    s = scenes.plain_box(32, 1.0)  # 32,768 particles, 25^3 cells, 14,408 walls
    sim = _sim(s, stab=stab, solver=solver, mu_s=fric[0], mu_k=fric[1],
               skin=0.0075 if lists else 0.0)
    sim.step(steps)
    gp, gv, gm = sim.host()
    op, ov, om, oms, op2c, _ = _oracle_steps(s, steps, stab, solver, *fric)
    assert np.array_equal(sim.p2c[: s.n].cpu().numpy(), op2c)
    bad = np.nonzero(np.any(gp != op, axis=1))[0]
    assert bad.size == 0, (bad[:10], gp[bad[:3]], op[bad[:3]])
    assert np.array_equal(gv, ov)
    assert np.array_equal(gm, om)
    assert np.array_equal(sim.scaled_mass[: s.n].cpu().numpy(), oms)


def test_step_within_tolerance_of_fp64(gpu):
    _log("fp64 shadow tolerance, C1 + friction")
    s = scenes.plain_box(32, 1.0)
    p = po.params_from(s, 3, 10, 0.5, 0.4)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    for steps, tol in ((1, 1e-5 * 0.01), (10, 1e-3 * 0.01)):
        sim = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
        sim.step(steps)
        gp = sim.host()[0]
        _, _, p64, _ = po.shadow_steps(s.pos, s.vel, s.mass, bpos, bcs, p, steps)
        assert np.abs(gp.astype(np.float64) - p64).max() <= tol


def test_neighbour_lists_bitexact(gpu):
    import torch
    _log("neighbour lists")
    s = scenes.plain_box(24, 1.0)
    gmax, bmax = 48, 16
    sim = _sim(s, gmax=gmax, bmax=bmax, skin=0.0075)
    sim.grid_build()
    sim.predict()
    torch.cuda.synchronize()
    n = s.n
    cap = sim.cap + 1
    x0 = sim.view(0, n)
    off = sim.view(1, n)
    meta = sim.view(2, n, dtype=torch.int32)
    gl = sim.view(3, gmax * cap, dtype=torch.int32, width=1).reshape(gmax, cap)[:, :n]
    bl = sim.view(4, bmax * cap, dtype=torch.int32, width=1).reshape(bmax, cap)[:, :n]
    x0, off, meta = x0.cpu().numpy(), off.cpu().numpy(), meta.cpu().numpy()
    gl, bl = gl.cpu().numpy(), bl.cpu().numpy()
    p = po.params_from(s, skin=0.0075)
    bpos, bcs = po.sorted_boundary(s.boundary, s.cell_length, s.cells)
    cs = sim.cs.cpu().numpy()
    ogc = np.empty(n, np.int32); ogl = np.empty(n * gmax, np.int32)
    obc = np.empty(n, np.int32); obl = np.empty(n * bmax, np.int32)
    obase = np.empty(n, np.int32)
    po.load().oracle_nlist_f32(n, np.ascontiguousarray(x0).reshape(-1),
                               np.ascontiguousarray(off).reshape(-1), bpos.shape[0],
                               bpos.reshape(-1), cs, bcs, p, gmax, bmax, ogc, ogl, obc, obl,
                               obase)
    assert np.array_equal(meta[:, 0], ogc)
    assert np.array_equal(meta[:, 1], obc)
    assert ogc.max() <= gmax and obc.max() <= bmax
    ogl = ogl.reshape(n, gmax)
    obl = obl.reshape(n, bmax)
    for i in range(n):
        assert np.array_equal(gl[: ogc[i], i], ogl[i, : ogc[i]])
        assert np.array_equal(bl[: obc[i], i], obl[i, : obc[i]])


def test_fallback_paths_agree(gpu):
    # particles whose list is not valid for a sweep scan the stencil in their
    # own lane or by a whole wave: "auto" queues them in the block (this set
    # is small), "wave" forces the sweep-wide queue and its second pass, "lane"
    # scans every one in its lane; all must give the oracle's result
    import torch
    from granular_gpu_amd import _lib
    steps = 60
    _log(f"fallback paths, C1 + friction, {steps} steps")
    s = scenes.plain_box(32, 1.0)
    runs = {}
    for name, flags in (("auto", 0), ("lane", _lib.GG_FLAG_FALLBACK_LANE),
                        ("wave", _lib.GG_FLAG_FALLBACK_WAVE)):
        sim = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4, flags=flags)
        sim.step(steps)
        ctr = sim.view(8, 64, dtype=torch.int32, width=1).reshape(-1).cpu().numpy()
        runs[name] = (sim.host(), int(ctr[1:14].sum()))
        _log(f"{name}: {runs[name][1]} stencil scans in the last step")
    # (the scan count itself may vary between runs: a particle that flags
    # its stencil during a sweep races with readers in that sweep, which
    # changes the path they take but not their result)
    assert min(r[1] for r in runs.values()) > 0
    for name in ("lane", "wave"):
        for x, y in zip(runs[name][0], runs["auto"][0]):
            assert np.array_equal(x, y), name
    op, ov, _, _, _, _ = _oracle_steps(s, steps, 3, 10, 0.5, 0.4)
    assert np.array_equal(runs["wave"][0][0], op)
    assert np.array_equal(runs["wave"][0][1], ov)


def test_list_refresh_matches_oracle(gpu):
    # C1 with friction at dt 1/240: around step 100 the column is pushed
    # through the floor and lists lapse inside the step; with
    # GG_FLAG_LIST_REFRESH they are rebuilt after a sweep whose stencil-scan
    # count grew past the first sweep's by more than n / (8 R), R the sweeps
    # left.  The result must not change: the refreshed run equals the run
    # without refresh and the oracle bit for bit, and a refresh must have
    # happened (without one a step's scan counts never fall from one sweep
    # to the next: displacements and marks only grow within a step).
    import torch
    from granular_gpu_amd import _lib
    steps = 110
    s = scenes.plain_box(32, 1.0)
    n = s.n
    runs, drop = {}, 0
    for name, flags in (("plain", 0), ("refresh", _lib.GG_FLAG_LIST_REFRESH)):
        sim = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4, flags=flags)
        for _ in range(steps):
            sim.step(1)
            if flags:
                c = sim.view(8, 64, dtype=torch.int32, width=1).reshape(-1).cpu().numpy()[1:14]
                drop = max(drop, max(int(c[k]) - int(c[k + 1]) for k in range(1, 12)))
        runs[name] = sim.host()
    _log(f"list refresh, C1 + friction, {steps} steps: largest drop after a sweep "
         f"{drop} stencil scans (n = {n})")
    assert drop > n // 100
    for x, y in zip(runs["refresh"], runs["plain"]):
        assert np.array_equal(x, y)
    op, ov, _, _, _, _ = _oracle_steps(s, steps, 3, 10, 0.5, 0.4)
    assert np.array_equal(runs["refresh"][0], op)
    assert np.array_equal(runs["refresh"][1], ov)


def test_step_graph_replays_equal_eager_steps(gpu):
    # every replay of the captured step (eager work and a synchronisation
    # between replays) must equal the same steps launched kernel by kernel
    import torch
    _log("captured step (gg_step_graph)")
    s = scenes.plain_box(32, 1.0)
    a = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    b = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    g = b.graph(steps_per_launch=5)
    for r in range(4):
        g.launch()
        scratch = torch.ones(1 << 20, device="cuda").cumsum(0)  # eager work in between
        torch.cuda.synchronize()
        a.step(5)
        assert float(scratch[-1].item()) == float(1 << 20)
        for x, y in zip(a.host(), b.host()):
            assert np.array_equal(x, y), r
    g.close()
    op, ov, _, _, _, _ = _oracle_steps(s, 20, 3, 10, 0.5, 0.4)
    assert np.array_equal(b.host()[0], op) and np.array_equal(b.host()[1], ov)


def test_run_twice_bitwise_identical(gpu):
    _log("determinism")
    s = scenes.plain_box(32, 1.0)
    a = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    b = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    a.step(20)
    b.step(20)
    for x, y in zip(a.host(), b.host()):
        assert np.array_equal(x, y)


def test_edge_sizes(gpu):
    import ctypes as C
    from granular_gpu_amd import _lib
    _log("edge sizes")
    s = scenes.plain_box(2, 1.0)
    s.pos = s.pos[:1].copy()
    s.vel = np.array([[0.3, 0.0, 0.0]], np.float32)
    s.mass = s.mass[:1].copy()
    sim = _sim(s)
    sim.step(2)
    op = _oracle_steps(s, 2, 3, 5)[0]
    assert np.array_equal(sim.host()[0], op)
    # empty set
    sim.n = 0
    sim._refresh_set()
    sim.step(1)
    # capacity overflow is reported, not executed
    sim.n = sim.cap + 1
    sim._refresh_set()
    with pytest.raises(_lib.GGError) as e:
        sim.step(1)
    assert e.value.status == _lib.GG_ERR_CAPACITY
    sim.n = 1
    sim._refresh_set()
    # too many sweeps for the workspace
    sim.params.solver_iters = 50
    with pytest.raises(_lib.GGError):
        sim.step(1)


def test_m1_full_size_properties(gpu):
    import torch
    _log("M1 1M particles, 3 steps")
    s = scenes.plain_box(100, 2.4)
    sim = _sim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
    sim.step(3)
    torch.cuda.synchronize()
    keys = sim.p2c[: s.n].cpu().numpy()
    cs = sim.cs.cpu().numpy()
    pos, vel, mass = sim.host()
    assert np.all(np.diff(keys) >= 0)
    counts = np.bincount(keys, minlength=60 ** 3 + 1)
    assert np.array_equal(cs, np.concatenate([[0], np.cumsum(counts)]))
    assert np.isfinite(pos).all() and np.isfinite(vel).all()
    assert np.array_equal(np.sort(mass), np.ones(s.n, np.float32))
    # centre of mass is still in free fall (no floor contact yet): y drop of
    # dt^2 g k(k+1)/2 within jitter-contact noise
    dt, g = float(s.dt), float(s.gravity[1])
    drop = dt * dt * g * 3 * 4 / 2
    assert abs(pos[:, 1].mean() - (s.pos[:, 1].mean() + drop)) < 1e-5


def test_separate_stages_equal_fused_predict(gpu):
    # gg_add_external_force + gg_predict_positions + gg_mass_scaling equal
    # the fused predict pass bit for bit, and each equals its IEEE float32
    # formula evaluated by numpy; gg_sort_positions_by_key = numpy's stable
    # argsort of the keys.
    import ctypes as C

    import torch
    from granular_gpu_amd import _lib
    s = scenes.plain_box(12, 1.0, dt=0.002, g=-9.8)
    _log(f"separate stages n={s.n}")
    sim = _sim(s, stab=3, solver=5)
    sim.grid_build()
    sim.predict()
    xs = sim.view(5, s.n)[:, :3].cpu().numpy()
    ms_fused = sim.scaled_mass[:s.n].cpu().numpy().copy()
    pos0, vel0, mass0 = sim.host()
    lib = sim.lib
    dt = np.float32(s.dt)
    dv = _lib.gg_float3(*(float(dt * np.float32(g)) for g in s.gravity))
    _lib.check("gg_add_external_force", lib.gg_add_external_force(C.byref(sim.set), dv, None))
    out = torch.empty((s.n, 3), dtype=torch.float32, device=sim.dev)
    _lib.check("gg_predict_positions",
               lib.gg_predict_positions(C.byref(sim.set), float(dt), C.c_void_p(out.data_ptr()),
                                        None))
    _lib.check("gg_mass_scaling", lib.gg_mass_scaling(C.byref(sim.set), None))
    torch.cuda.synchronize()
    v1 = sim.vel[:s.n].cpu().numpy()
    want_v = vel0 + np.array([dv.x, dv.y, dv.z], np.float32)
    assert np.array_equal(v1, want_v)
    assert np.array_equal(out.cpu().numpy(), pos0 + dt * v1)
    assert np.array_equal(out.cpu().numpy(), xs)
    ms = sim.scaled_mass[:s.n].cpu().numpy()
    assert np.array_equal(ms, mass0 * np.exp(-pos0[:, 1].astype(np.float64)).astype(np.float32))
    assert np.array_equal(ms, ms_fused)
    # the positions alone sorted by given keys (out-of-range keys -> sentinel)
    rng = np.random.default_rng(3)
    C_ = sim.C
    keys = rng.integers(-3, C_ + 4, size=s.n).astype(np.int32)
    sim.p2c[:s.n] = torch.from_numpy(keys).to(sim.dev)
    cs = torch.zeros(C_ + 2, dtype=torch.int32, device=sim.dev)
    _lib.check("gg_sort_positions_by_key",
               lib.gg_sort_positions_by_key(C.byref(sim.set), C.c_void_p(cs.data_ptr()),
                                            C.byref(sim.ws), None))
    torch.cuda.synchronize()
    eff = np.where((keys < 0) | (keys > C_), C_, keys)
    order = np.argsort(eff, kind="stable")
    assert np.array_equal(sim.pos[:s.n].cpu().numpy(), pos0[order])
    assert np.array_equal(sim.p2c[:s.n].cpu().numpy(), keys)
    cnt = np.bincount(eff, minlength=C_ + 1)
    assert np.array_equal(cs.cpu().numpy()[:C_ + 2], np.concatenate([[0], np.cumsum(cnt)]))


@pytest.mark.parametrize("mode,count", [(0, 1 << 32), (1, 1 << 32), (2, 1 << 32)])
def test_short_sqrt_div_equal_ieee(gpu, mode, count):
    # the contact kernels' short square root / reciprocal / quotient equal
    # sqrtf and '/' on every input of their domains (mode 2: 2^32 pairs)
    import torch
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import _ptr, _stream
    _log(f"fast math check mode {mode}")
    bad = torch.zeros(1, dtype=torch.int64, device="cuda")
    _lib.check("gg_fast_math_check",
               _lib.load().gg_fast_math_check(mode, 0, count, _ptr(bad), _stream()))
    torch.cuda.synchronize()
    assert int(bad.item()) == 0
