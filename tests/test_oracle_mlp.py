"""Pins the learned-upsampler oracle: bfloat16 rounding against a bit-level
restatement, kNN-16 features against brute force, the teacher against a
closed form and Galilean invariance, the MLP forward against numpy, its
backward against finite differences, Adam against its first-step closed
form.  No GPU needed."""
import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

R = 0.02


def bf16_np(x):
    """Round-to-nearest-even to bfloat16, in numpy integer arithmetic."""
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(nan, (u | 0x00400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32)


def test_bf16_rounding():
    one = np.float32(1.0)
    cases = np.float32([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -7 + 2 ** -9), 3.14159265,
                        1e-30, -0.0, np.inf, -np.inf, 3.4028235e38, 6e-39])
    for c in cases:
        assert np.array_equal(np.float32(po.bf16_round(c)), bf16_np(c)), c
    assert po.bf16_round(1 + 2 ** -8) == one                 # tie to even (down)
    assert po.bf16_round(1 + 3 * 2 ** -8) == np.float32(1 + 2 ** -6)  # tie to even (up)
    assert np.isinf(po.bf16_round(3.4028235e38))             # overflow rounds to inf
    assert np.isnan(po.bf16_round(np.nan))
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(2000) * 10.0 ** rng.uniform(-20, 20, 2000)).astype(np.float32)
    got = np.array([po.bf16_round(v) for v in x], np.float32)
    assert np.array_equal(got, bf16_np(x))


def _cloud(n=300, seed=2, box=(0.4, 0.6)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(box[0], box[1], (n, 3)).astype(np.float32)
    vel = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    s = scenes.plain_box(2, 1.0)
    _, perm, cs = po.grid_build(pos, s.cell_length, s.cells)
    return s, np.ascontiguousarray(pos[perm]), np.ascontiguousarray(vel[perm]), cs


def _no_boundary(s):
    return np.zeros((0, 3), np.float32), np.zeros(s.cells[0] * s.cells[1] * s.cells[2] + 2,
                                                  np.int32)


def test_features_match_brute_force():
    s, pos, vel, cs = _cloud()
    bpos, bcs = _no_boundary(s)
    cfg = po.mlp_cfg(R)
    feat, flags, knn = po.mlp_features(pos, vel, cs, bpos, bcs, po.params_from(s), cfg)
    L = np.float32(s.cell_length)
    cell = np.floor(pos / L).astype(int)
    for i in range(pos.shape[0]):
        near = np.all(np.abs(cell - cell[i]) <= 1, axis=1)
        near[i] = False
        js = np.nonzero(near)[0]
        d = pos[js] - pos[i]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        order = np.lexsort((js, d2))[:16]
        want = np.full(16, -1)
        want[:len(order)] = js[order]
        assert np.array_equal(knn[i], want), i
        f = np.zeros((16, 6), np.float32)
        for slot, j in enumerate(want):
            if j < 0:
                continue
            f[slot, :3] = (pos[j] - pos[i]) * (np.float32(1.0) / np.float32(R))
            f[slot, 3:] = (vel[j] - vel[i]) * (np.float32(1.0) / np.float32(1.0))
        assert np.array_equal(feat[i], bf16_np(f.reshape(-1))), i
    assert (flags == 0).all()


def test_feature_flags():
    s, pos, vel, cs = _cloud(n=40)
    wall = np.float32([[0.5, 0.5, 0.5]])
    bpos, bcs = po.sorted_boundary(wall, s.cell_length, s.cells)
    pos = pos.copy()
    _, perm, cs = po.grid_build(pos, s.cell_length, s.cells)
    feat, flags, _ = po.mlp_features(pos, vel, cs, bpos, bcs, po.params_from(s), po.mlp_cfg(R))
    d = np.linalg.norm(pos - wall, axis=1)
    assert np.array_equal((flags & 1) == 1, d < 0.05)
    low = np.float32([[0.5, 0.05, 0.5], [0.03, 0.5, 0.5], [0.5, 0.5, 0.5]])
    _, perm, cs = po.grid_build(low, s.cell_length, s.cells)
    low = low[perm]
    _, fl, _ = po.mlp_features(low, np.zeros_like(low), cs, *_no_boundary(s), po.params_from(s),
                               po.mlp_cfg(R))
    assert list(fl) == [2 if (p[1] < 0.08 or p[0] < 0.05) else 0 for p in low]


def test_teacher_single_particle_closed_form():
    s = scenes.plain_box(2, 1.0)
    pos = np.float32([[0.5, 0.5, 0.5]])
    vel = np.float32([[0.3, -0.2, 0.1]])
    _, _, cs = po.grid_build(pos, s.cell_length, s.cells)
    y, fine = po.mlp_teacher(pos, vel, cs, *_no_boundary(s), po.params_from(s), po.mlp_cfg(R))
    d = 0.5 * np.sqrt(3.0) * R
    w = (1.0 - d * d * 493.8271605) ** 3
    alpha = 1.0 - w
    for k in range(8):
        assert y[0, 3 * k] == pytest.approx(0.0, abs=1e-6)
        assert y[0, 3 * k + 2] == pytest.approx(0.0, abs=1e-6)
        assert y[0, 3 * k + 1] == pytest.approx(alpha * -9.8 * 0.002, rel=1e-4)
        t = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1]) - 0.5
        x_f = pos[0] + R * t
        want = x_f + 0.002 * (vel[0] + y[0, 3 * k:3 * k + 3])
        assert np.allclose(fine[0, 3 * k:3 * k + 3], want, atol=1e-7)
    # the positions of the learned path from the teacher's own targets
    assert np.allclose(po.mlp_positions(pos, vel, y, po.mlp_cfg(R)), fine, atol=1e-7)


def test_teacher_is_galilean_invariant():
    s, pos, vel, cs = _cloud(n=200)
    vel = (vel * 0.2).astype(np.float32)
    bpos, bcs = _no_boundary(s)
    y0, _ = po.mlp_teacher(pos, vel, cs, bpos, bcs, po.params_from(s), po.mlp_cfg(R))
    y1, _ = po.mlp_teacher(pos, vel + np.float32([0.7, -0.3, 0.2]), cs, bpos, bcs,
                           po.params_from(s), po.mlp_cfg(R))
    assert np.abs(y0).max() > 0.05
    assert np.abs(y1 - y0).max() < 2e-6


def _fine_sorted(fpos, fvel, s):
    _, perm, fcs = po.grid_build(fpos, s.cell_length, s.cells)
    return fpos[perm], fvel[perm], fcs, perm


def test_fine_targets_of_the_paired_start_are_the_children():
    # at the start of a paired run the fine particles sit exactly on the
    # template points: each slot matches its own child at distance 0
    coarse, fine = scenes.paired_boxes(5, 1.0, r=R)
    rng = np.random.default_rng(4)
    cvel = rng.uniform(-1, 1, (coarse.n, 3)).astype(np.float32)
    fvel = rng.uniform(-1, 1, (fine.n, 3)).astype(np.float32)
    fp, fv, fcs, perm = _fine_sorted(fine.pos, fvel, fine)
    y, d2, jm = po.mlp_fine_targets(coarse.pos, cvel, fp, fv, fcs, fine.cell_length,
                                    fine.cells, R)
    # the template points in float32 as the oracle forms them
    q = coarse.pos[:, None, :] + np.float32(R) * scenes.TEMPLATE8.astype(np.float32)[None]
    assert np.array_equal(q.reshape(-1, 3), fine.pos)  # float64 and float32 forms agree
    assert (d2 == 0).all()
    assert np.array_equal(perm[jm], np.arange(fine.n).reshape(-1, 8))
    assert np.array_equal(y.reshape(-1, 8, 3), fvel.reshape(-1, 8, 3) - cvel[:, None, :])


def test_fine_targets_match_brute_force_nearest():
    s, pos, vel, _ = _cloud(n=150)
    rng = np.random.default_rng(9)
    nf = 3000
    fpos = rng.uniform(0.3, 0.7, (nf, 3)).astype(np.float32)
    fvel = rng.uniform(-1, 1, (nf, 3)).astype(np.float32)
    # duplicated fine particles: equal distances resolve to the lower index
    fpos[1500:1600] = fpos[:100]
    fp, fv, fcs, _ = _fine_sorted(fpos, fvel, s)
    y, d2, jm = po.mlp_fine_targets(pos, vel, fp, fv, fcs, s.cell_length, s.cells, R)
    T = scenes.TEMPLATE8.astype(np.float32)
    checked = 0
    for i in range(pos.shape[0]):
        for k in range(8):
            q = pos[i] + np.float32(R) * T[k]
            d = fp - q
            dd = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
            j = int(np.argmin(dd))  # first of the minima: the lower index
            if np.sqrt(dd[j]) >= s.cell_length:
                continue  # the nearest may lie outside the 27-cell stencil
            checked += 1
            assert jm[i, k] == j and d2[i, k] == dd[j]
            assert np.array_equal(y[i, 3 * k:3 * k + 3], fv[j] - vel[i])
    assert checked > 1000


def _net(seed=1, scale=0.15):
    rng = np.random.default_rng(seed)
    W = np.zeros(po.MLP_W, np.float32)
    b = np.zeros(po.MLP_B, np.float32)
    for l, (k_in, n_out) in enumerate(po.MLP_DIMS):
        n_used = 24 if l == 3 else n_out
        w = rng.uniform(-1, 1, (n_out, k_in)) * scale * np.sqrt(6.0 / k_in)
        w[n_used:] = 0.0
        W[po.MLP_WOFF[l]:po.MLP_WOFF[l] + n_out * k_in] = w.reshape(-1)
        b[po.MLP_BOFF[l]:po.MLP_BOFF[l] + n_used] = rng.uniform(-0.1, 0.1, n_used)
    return W, b


def _np_forward(x, W, b, rnd):
    h = rnd(x.astype(np.float32)).astype(np.float64)
    for l, (k_in, n_out) in enumerate(po.MLP_DIMS):
        w = rnd(W[po.MLP_WOFF[l]:po.MLP_WOFF[l] + n_out * k_in]).reshape(n_out, k_in)
        z = (h @ w.astype(np.float64).T).astype(np.float32) + b[po.MLP_BOFF[l]:po.MLP_BOFF[l] + n_out]
        if l < 3:
            h = rnd(np.maximum(z, 0)).astype(np.float64)
        else:
            return z[:, :24]


@pytest.mark.parametrize("exact", [1, 0])
def test_forward_matches_numpy(exact):
    W, b = _net(scale=1.0)
    x = np.random.default_rng(3).standard_normal((50, 96)).astype(np.float32)
    rnd = (lambda a: np.asarray(a, np.float32)) if exact else bf16_np
    y = po.mlp_forward(x, W, b, exact=exact)
    want = _np_forward(x, W, b, rnd)
    assert np.abs(y).max() > 0.1
    np.testing.assert_allclose(y, want, rtol=2e-5, atol=2e-6)


def test_backward_matches_finite_differences():
    W, b = _net(scale=1.0)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, 96)).astype(np.float32)
    t = rng.standard_normal((4, 24)).astype(np.float32)
    loss, dW, db = po.mlp_backward(x, W, b, t, exact=1)
    h = 1e-2
    picks = []
    for l, (k_in, n_out) in enumerate(po.MLP_DIMS):
        n_used = 24 if l == 3 else n_out
        for _ in range(8):
            n, k = rng.integers(n_used), rng.integers(k_in)
            picks.append(("W", po.MLP_WOFF[l] + n * k_in + k))
        for _ in range(3):
            picks.append(("b", po.MLP_BOFF[l] + rng.integers(n_used)))
    checked = 0
    for kind, idx in picks:
        Wp, Wm, bp, bm = W.copy(), W.copy(), b.copy(), b.copy()
        if kind == "W":
            Wp[idx] += h
            Wm[idx] -= h
            g = dW[idx]
        else:
            bp[idx] += h
            bm[idx] -= h
            g = db[idx]
        lp = po.mlp_backward(x, Wp, bp, t, exact=1)[0]
        lm = po.mlp_backward(x, Wm, bm, t, exact=1)[0]
        fd = (lp - lm) / (2 * h)
        if abs(fd) < 1e-5 and abs(g) < 1e-5:
            continue
        assert g == pytest.approx(fd, rel=3e-2, abs=3e-5), (kind, idx)
        checked += 1
    assert checked > 20
    # the rounded (GPU) semantics stays close to the exact gradient
    _, dWr, dbr = po.mlp_backward(x, W, b, t, exact=0)
    rel = np.linalg.norm(dWr - dW) / np.linalg.norm(dW)
    assert rel < 0.05


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(5)
    w = rng.standard_normal(100).astype(np.float32)
    g = rng.standard_normal(100).astype(np.float32)
    w1, m1, v1 = po.mlp_adam(w, np.zeros(100), np.zeros(100), g, lr=1e-3, step=1)
    np.testing.assert_allclose(w1, w - 1e-3 * np.sign(g), atol=2e-7 + 1e-3 * 1e-6)
    one = np.float32(1.0)
    np.testing.assert_allclose(m1, (one - np.float32(0.9)) * g, rtol=1e-6)
    np.testing.assert_allclose(v1, (one - np.float32(0.999)) * (g * g), rtol=1e-6)
