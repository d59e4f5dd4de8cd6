"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "_build" / "liboracle.so"


class oracle_params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gx", C.c_float), ("gy", C.c_float), ("gz", C.c_float),
        ("cell_length", C.c_float), ("cx", C.c_int), ("cy", C.c_int), ("cz", C.c_int),
        ("density", C.c_float), ("stab_iters", C.c_int), ("solver_iters", C.c_int),
        ("mu_s", C.c_float), ("mu_k", C.c_float), ("boundary_radius", C.c_float),
        ("contact_radius_min", C.c_float), ("sleep_speed", C.c_float),
        ("damp_speed", C.c_float), ("damping", C.c_float), ("adaptive_damping", C.c_float),
        ("skin", C.c_float), ("own_only", C.c_int),
        ("depth_rules", C.c_int), ("min_mass", C.c_float), ("split_mass", C.c_float),
        ("merge_dist", C.c_float), ("merge_depth", C.c_float), ("split_depth", C.c_float),
    ]


class oracle_upsample_params(C.Structure):
    _fields_ = [(k, C.c_float) for k in (
        "dt", "gx", "gy", "gz", "r9", "boundary_radius", "repulsion", "min_weight",
        "near_damping", "ground_y", "ground_damping", "max_velocity", "floor_y",
        "floor_friction")]


class oracle_animation(C.Structure):
    _fields_ = [
        ("is_animating", C.c_int), ("animation_time", C.c_float), ("start_index", C.c_int),
        ("num_particles", C.c_int), ("translation_direction", C.c_float * 3),
        ("translation_speed", C.c_float), ("translation_complete", C.c_int),
        ("rotation_angle", C.c_float), ("rotation_speed", C.c_float),
        ("rotation_center", C.c_float * 3), ("translation_time", C.c_float),
        ("rotation_limit", C.c_float), ("center_from_marked", C.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    P = C.POINTER(oracle_params)
    lib.oracle_hash.argtypes = [fp, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, ip]
    lib.oracle_grid_build.argtypes = [fp, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, ip,
                                      ip, ip]
    lib.oracle_plain_steps_f32.argtypes = [C.c_int, fp, fp, fp, C.c_int, fp, ip, P, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p]
    lib.oracle_plain_steps_f64.argtypes = [C.c_int, dp, dp, dp, C.c_int, dp, ip, P, C.c_int,
                                           C.c_void_p, C.c_void_p]
    lib.oracle_shadow_steps.argtypes = [C.c_int, fp, fp, fp, dp, dp, dp, C.c_int, fp, dp, ip, P,
                                        C.c_int, C.c_void_p]
    lib.oracle_sweep_f32.argtypes = [C.c_int, fp, fp, fp, C.c_int, fp, ip, ip, P, C.c_int]
    lib.oracle_nlist_f32.argtypes = [C.c_int, fp, fp, C.c_int, fp, ip, ip, P, C.c_int, C.c_int,
                                     ip, ip, ip, ip, ip]
    vp = C.c_void_p
    lib.oracle_upsample_params_default.argtypes = [C.POINTER(oracle_upsample_params)]
    lib.oracle_upsample_params_default.restype = None
    lib.oracle_upsample_plain.argtypes = [C.c_int, fp, fp, C.c_int, fp, fp, ip, C.c_int, fp, ip, P,
                                          C.POINTER(oracle_upsample_params)]
    lib.oracle_surface.argtypes = [C.c_int, fp, fp, ip, P, fp, ip]
    lib.oracle_surface_distance.argtypes = [C.c_int, fp, ip, ip, P, fp]
    lib.oracle_merge_mark.argtypes = [C.c_int, fp, fp, fp, ip, fp, ip, fp, fp, ip, P, C.c_float,
                                      vp, vp]
    lib.oracle_merge_blend.argtypes = [C.c_int, fp, fp, fp, ip, ip, fp, C.c_int, ip]
    lib.oracle_split_compact.argtypes = [C.c_int, C.c_int, fp, fp, fp, ip, ip, fp, ip, ip, vp,
                                         fp, ip, ip, P, ip]
    lib.oracle_split_compact.restype = C.c_int
    lib.oracle_adaptive_steps.argtypes = [C.POINTER(C.c_int), C.c_int, fp, fp, fp, ip, ip, fp,
                                          ip, fp, ip, fp, C.c_int, fp, ip, P, C.c_int, C.c_int,
                                          C.c_float, ip]
    for nm in ("oracle_radius_contact", "oracle_radius_surface", "oracle_radius_split"):
        getattr(lib, nm).argtypes = [C.c_float, C.c_float]
        getattr(lib, nm).restype = C.c_float
    lib.oracle_expf.argtypes = [C.c_float]
    lib.oracle_expf.restype = C.c_float
    lib.oracle_rotation_cs.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.oracle_rotation_cs.restype = None
    lib.oracle_animate.argtypes = [C.c_int, fp, ip, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float]
    lib.oracle_animate.restype = None
    lib.oracle_update_animation.argtypes = [C.POINTER(oracle_animation), C.c_int, fp, ip,
                                            C.c_float]
    lib.oracle_center_of_mass.argtypes = [C.c_int, fp, fp, fp]
    lib.oracle_center_of_mass.restype = None
    lib.oracle_explode.argtypes = [C.c_int, fp, fp, fp, fp, C.c_float]
    lib.oracle_explode.restype = None
    lib.oracle_pick_sphere.argtypes = [C.c_int, fp, fp, C.c_float, ip]
    lib.oracle_move_picked.argtypes = [fp, ip, C.c_int, fp, fp, fp]
    lib.oracle_move_picked.restype = None
    lib.oracle_set_threads.argtypes = [C.c_int]
    lib.oracle_get_threads.restype = C.c_int
    _lib = lib
    return lib


def params_from(scene, stab=3, solver=5, mu_s=0.0, mu_k=0.0, skin=0.005,
                own_only=0, rules=None) -> oracle_params:
    """Oracle parameters for a scenes.Scene (reference constants elsewhere).
    rules: None for the reference merge/split rules, or a dict of
    depth_rules, min_mass, split_mass, merge_dist, merge_depth, split_depth."""
    p = oracle_params()
    p.dt = float(scene.dt)
    p.gx, p.gy, p.gz = (float(v) for v in scene.gravity)
    p.cell_length = float(scene.cell_length)
    p.cx, p.cy, p.cz = scene.cells
    p.density = float(scene.density)
    p.stab_iters = stab
    p.solver_iters = solver
    p.mu_s = mu_s
    p.mu_k = mu_k
    p.boundary_radius = 0.01
    p.contact_radius_min = 0.01
    p.sleep_speed = 0.01
    p.damp_speed = 1.0
    p.damping = 0.99
    p.adaptive_damping = 0.98
    p.skin = skin
    p.own_only = own_only
    # reference merge / split rules (Solver.cu:526-827)
    p.depth_rules, p.min_mass, p.split_mass = 0, 1.0, 2.0
    p.merge_dist, p.merge_depth, p.split_depth = 0.035, 3.0, 2.0
    for k, v in (rules or {}).items():
        setattr(p, k, v)
    return p


def grid_build(pos: np.ndarray, cell_length, cells):
    lib = load()
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n = pos.shape[0]
    C_ = cells[0] * cells[1] * cells[2]
    keys = np.empty(n, np.int32)
    perm = np.empty(n, np.int32)
    cs = np.empty(C_ + 2, np.int32)
    lib.oracle_grid_build(pos.reshape(-1), n, float(cell_length), *cells, keys, perm, cs)
    return keys, perm, cs


def sorted_boundary(bpos: np.ndarray, cell_length, cells):
    """Boundary positions sorted by cell (stable) and its cell table."""
    keys, perm, cs = grid_build(bpos, cell_length, cells)
    return np.ascontiguousarray(bpos[perm]), cs


def plain_steps_f32(pos, vel, mass, bpos_sorted, bcs, params, nsteps, adaptive_last=None):
    lib = load()
    pos = np.ascontiguousarray(pos, np.float32).copy()
    vel = np.ascontiguousarray(vel, np.float32).copy()
    mass = np.ascontiguousarray(mass, np.float32).copy()
    n = pos.shape[0]
    ms = np.empty(n, np.float32)
    p2c = np.empty(n, np.int32)
    adl = None if adaptive_last is None else np.ascontiguousarray(adaptive_last, np.int32).copy()
    rc = lib.oracle_plain_steps_f32(
        n, pos.reshape(-1), vel.reshape(-1), mass, bpos_sorted.shape[0],
        np.ascontiguousarray(bpos_sorted, np.float32).reshape(-1), bcs, C.byref(params), nsteps,
        ms.ctypes.data, p2c.ctypes.data, None if adl is None else adl.ctypes.data)
    assert rc == 0
    return pos, vel, mass, ms, p2c, adl


def plain_steps_f64(pos, vel, mass, bpos_sorted, bcs, params, nsteps):
    """Independent FP64 run; returns (pos, vel, mass, ids) with ids the
    original index of each sorted particle."""
    lib = load()
    pos = np.ascontiguousarray(pos, np.float64).copy()
    vel = np.ascontiguousarray(vel, np.float64).copy()
    mass = np.ascontiguousarray(mass, np.float64).copy()
    n = pos.shape[0]
    ids = np.arange(n, dtype=np.int32)
    rc = lib.oracle_plain_steps_f64(
        n, pos.reshape(-1), vel.reshape(-1), mass, bpos_sorted.shape[0],
        np.ascontiguousarray(bpos_sorted, np.float64).reshape(-1), bcs, C.byref(params), nsteps,
        None, ids.ctypes.data)
    assert rc == 0
    return pos, vel, mass, ids


def shadow_steps(pos, vel, mass, bpos_sorted, bcs, params, nsteps):
    """Lockstep FP32 + decision-shadowed FP64 run.
    Returns (pos32, vel32, pos64, vel64) in the same (sorted) order."""
    lib = load()
    p32 = np.ascontiguousarray(pos, np.float32).copy()
    v32 = np.ascontiguousarray(vel, np.float32).copy()
    m32 = np.ascontiguousarray(mass, np.float32).copy()
    p64 = p32.astype(np.float64)
    v64 = v32.astype(np.float64)
    m64 = m32.astype(np.float64)
    b32 = np.ascontiguousarray(bpos_sorted, np.float32)
    b64 = b32.astype(np.float64)
    rc = lib.oracle_shadow_steps(p32.shape[0], p32.reshape(-1), v32.reshape(-1), m32,
                                 p64.reshape(-1), v64.reshape(-1), m64, b32.shape[0],
                                 b32.reshape(-1), b64.reshape(-1), bcs, C.byref(params), nsteps,
                                 None)
    assert rc == 0
    return p32, v32, p64, v64


class AdaptiveState:
    """Host copy of everything the adaptive path reads and writes."""

    def __init__(self, pos, vel, mass, cap=None):
        n = pos.shape[0]
        self.cap = cap or n
        cap = self.cap
        self.n = n
        self.pos = np.zeros((cap, 3), np.float32); self.pos[:n] = pos
        self.vel = np.zeros((cap, 3), np.float32); self.vel[:n] = vel
        self.mass = np.zeros(cap, np.float32); self.mass[:n] = mass
        self.adl = np.zeros(cap, np.int32)
        self.remove = np.zeros(cap, np.int32)
        self.merge = np.zeros(cap, np.float32)
        self.mcount = np.zeros(cap, np.int32)
        self.mv = np.zeros((cap, 3), np.float32)
        self.surface = np.zeros(cap, np.int32)
        self.sd = np.zeros(cap, np.float32)

    def copy(self):
        import copy
        return copy.deepcopy(self)


def adaptive_steps(st: AdaptiveState, bpos_sorted, bcs, params, nsteps, merge_first=0,
                   max_mass=4.0):
    lib = load()
    nn = C.c_int(st.n)
    stats = np.zeros(3 * max(nsteps, 1), np.int32)
    rc = lib.oracle_adaptive_steps(
        C.byref(nn), st.cap, st.pos.reshape(-1), st.vel.reshape(-1), st.mass, st.adl, st.remove,
        st.merge, st.mcount, st.mv.reshape(-1), st.surface, st.sd, bpos_sorted.shape[0],
        np.ascontiguousarray(bpos_sorted, np.float32).reshape(-1), bcs, C.byref(params), nsteps,
        merge_first, max_mass, stats)
    assert rc == 0
    st.n = nn.value
    return stats.reshape(-1, 3)


def upsample_params():
    u = oracle_upsample_params()
    load().oracle_upsample_params_default(C.byref(u))
    return u


def upsample_plain(fpos, fvel, cpos, cvel, cs, bpos_sorted, bcs, params, u=None, steps=1):
    """In-place copies: returns updated (fpos, fvel) after `steps` passes
    against fixed coarse data."""
    lib = load()
    u = u or upsample_params()
    fp_ = np.ascontiguousarray(fpos, np.float32).copy()
    fv_ = np.ascontiguousarray(fvel, np.float32).copy()
    for _ in range(steps):
        lib.oracle_upsample_plain(fp_.shape[0], fp_.reshape(-1), fv_.reshape(-1), cpos.shape[0],
                                  np.ascontiguousarray(cpos, np.float32).reshape(-1),
                                  np.ascontiguousarray(cvel, np.float32).reshape(-1), cs,
                                  bpos_sorted.shape[0],
                                  np.ascontiguousarray(bpos_sorted, np.float32).reshape(-1), bcs,
                                  C.byref(params), C.byref(u))
    return fp_, fv_


class oracle_mlp_cfg(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("radius", "vscale", "guard_boundary", "guard_floor",
                                         "guard_side")]


MLP_W = 49152     # bf16 weights of the 4 layers, [n_out][k_in] each, layer 4 padded to 32 rows
MLP_B = 416       # fp32 biases
MLP_WOFF = (0, 12288, 28672, 45056)
MLP_BOFF = (0, 128, 256, 384)
MLP_DIMS = ((96, 128), (128, 128), (128, 128), (128, 32))


def mlp_cfg(radius=0.02, vscale=1.0, guard_boundary=0.05, guard_floor=0.08, guard_side=0.05):
    return oracle_mlp_cfg(radius, vscale, guard_boundary, guard_floor, guard_side)


def _mlp_protos(lib):
    f32 = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    P = C.POINTER(oracle_params)
    lib.oracle_bf16_round.argtypes = [C.c_float]
    lib.oracle_bf16_round.restype = C.c_float
    lib.oracle_mlp_features.argtypes = [C.c_int, f32, f32, i32, C.c_int, f32, i32, P,
                                        C.POINTER(oracle_mlp_cfg), f32, i32, C.c_void_p]
    lib.oracle_mlp_teacher.argtypes = [C.c_int, f32, f32, i32, f32, i32, P,
                                       C.POINTER(oracle_upsample_params),
                                       C.POINTER(oracle_mlp_cfg), C.c_void_p, C.c_void_p]
    lib.oracle_mlp_positions.argtypes = [C.c_int, f32, f32, f32, C.POINTER(oracle_mlp_cfg),
                                         C.c_float, f32]
    lib.oracle_mlp_fine_targets.argtypes = [C.c_int, f32, f32, f32, f32, i32, C.c_float,
                                            C.c_int, C.c_int, C.c_int, C.c_float, f32, f32,
                                            i32]
    lib.oracle_mlp_forward.argtypes = [C.c_int, f32, f32, f32, f32, C.c_int]
    lib.oracle_mlp_backward.argtypes = [C.c_int, f32, f32, f32, f32, f32, f32, C.c_int]
    lib.oracle_mlp_backward.restype = C.c_double
    lib.oracle_mlp_adam.argtypes = [C.c_int, f32, f32, f32, f32, C.c_float, C.c_float,
                                    C.c_float, C.c_float, C.c_float, C.c_float]


def _f(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if shape is None else a.reshape(shape)


def mlp_features(pos, vel, cs, bpos_sorted, bcs, params, cfg):
    lib = load()
    _mlp_protos(lib)
    n = pos.shape[0]
    feat = np.zeros((n, 96), np.float32)
    flags = np.zeros(n, np.int32)
    knn = np.zeros((n, 16), np.int32)
    lib.oracle_mlp_features(n, _f(pos).reshape(-1), _f(vel).reshape(-1), cs,
                            bpos_sorted.shape[0], _f(bpos_sorted).reshape(-1), bcs,
                            C.byref(params), C.byref(cfg), feat.reshape(-1), flags,
                            knn.ctypes.data)
    return feat, flags, knn


def mlp_teacher(pos, vel, cs, bpos_sorted, bcs, params, cfg, u=None):
    lib = load()
    _mlp_protos(lib)
    n = pos.shape[0]
    u = u or upsample_params()
    y = np.zeros((n, 24), np.float32)
    fine = np.zeros((n, 24), np.float32)
    lib.oracle_mlp_teacher(n, _f(pos).reshape(-1), _f(vel).reshape(-1), cs,
                           _f(bpos_sorted).reshape(-1), bcs, C.byref(params), C.byref(u),
                           C.byref(cfg), y.ctypes.data, fine.ctypes.data)
    return y, fine


def mlp_fine_targets(pos, vel, fpos_sorted, fvel_sorted, fcs, f_cell_length, f_cells, radius):
    """Fine-run targets of every coarse particle: (y n x 24, d2 n x 8,
    matched fine index n x 8) over the fine set sorted by its cell table."""
    lib = load()
    _mlp_protos(lib)
    n = pos.shape[0]
    y = np.zeros((n, 24), np.float32)
    d2 = np.zeros((n, 8), np.float32)
    jm = np.zeros((n, 8), np.int32)
    lib.oracle_mlp_fine_targets(n, _f(pos).reshape(-1), _f(vel).reshape(-1),
                                _f(fpos_sorted).reshape(-1), _f(fvel_sorted).reshape(-1),
                                np.ascontiguousarray(fcs, np.int32), float(f_cell_length),
                                *[int(c) for c in f_cells], float(radius), y.reshape(-1),
                                d2.reshape(-1), jm.reshape(-1))
    return y, d2, jm


def mlp_positions(pos, vel, y, cfg, dt=0.002):
    lib = load()
    _mlp_protos(lib)
    n = pos.shape[0]
    fine = np.zeros((n, 24), np.float32)
    lib.oracle_mlp_positions(n, _f(pos).reshape(-1), _f(vel).reshape(-1), _f(y).reshape(-1),
                             C.byref(cfg), dt, fine.reshape(-1))
    return fine


def mlp_forward(x, W, b, exact=0):
    lib = load()
    _mlp_protos(lib)
    B = x.shape[0]
    y = np.zeros((B, 24), np.float32)
    lib.oracle_mlp_forward(B, _f(x).reshape(-1), _f(W), _f(b), y.reshape(-1), exact)
    return y


def mlp_backward(x, W, b, t, exact=0):
    lib = load()
    _mlp_protos(lib)
    B = x.shape[0]
    dW = np.zeros(MLP_W, np.float32)
    db = np.zeros(MLP_B, np.float32)
    loss = lib.oracle_mlp_backward(B, _f(x).reshape(-1), _f(W), _f(b), _f(t).reshape(-1), dW,
                                   db, exact)
    return loss, dW, db


def mlp_adam(w, m, v, g, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
    lib = load()
    _mlp_protos(lib)
    # bias corrections as the device library forms them: in double from the
    # float32 betas, rounded once
    b1d, b2d = float(np.float32(beta1)), float(np.float32(beta2))
    c1 = float(np.float32(1.0 / (1.0 - b1d ** step)))
    c2 = float(np.float32(1.0 / (1.0 - b2d ** step)))
    w, m, v = (np.ascontiguousarray(a, np.float32).copy() for a in (w, m, v))
    lib.oracle_mlp_adam(w.size, w, m, v, _f(g), lr, beta1, beta2, eps, c1, c2)
    return w, m, v


def bf16_round(x):
    lib = load()
    _mlp_protos(lib)
    return np.float32(lib.oracle_bf16_round(float(x)))


def animation_state(**kw) -> oracle_animation:
    """An animation state with the reference's AnimationState defaults
    (main.cu:100-111) and its 2.5 s / 45 degree limits, overridden by kw."""
    a = oracle_animation()
    a.translation_direction[:] = (-0.1, 0.0, 0.0)
    a.translation_speed = 5.0
    a.rotation_speed = 90.0
    a.translation_time = 2.5
    a.rotation_limit = 45.0
    for k, v in kw.items():
        if k in ("translation_direction", "rotation_center"):
            getattr(a, k)[:] = tuple(float(x) for x in v)
        else:
            setattr(a, k, v)
    return a


def rotation_cs(angle_deg):
    c, s = C.c_float(), C.c_float()
    load().oracle_rotation_cs(float(angle_deg), C.byref(c), C.byref(s))
    return np.float32(c.value), np.float32(s.value)


def animate(pos, animated, translation, angle_deg=0.0, center=(0.0, 0.0)):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    animated = np.ascontiguousarray(animated, np.int32)
    load().oracle_animate(pos.shape[0], pos.reshape(-1), animated,
                          *(float(v) for v in translation), float(angle_deg),
                          float(center[0]), float(center[1]))
    return pos


def update_animation(state: oracle_animation, pos, animated, dt):
    """One updateAnimation(dt) call on pos (modified in place); returns
    True if the positions moved."""
    assert pos.dtype == np.float32 and pos.flags.c_contiguous
    return bool(load().oracle_update_animation(
        C.byref(state), pos.shape[0], pos.reshape(-1),
        np.ascontiguousarray(animated, np.int32), float(dt)))


def center_of_mass(pos, mass):
    pos = np.ascontiguousarray(pos, np.float32)
    com = np.zeros(3, np.float32)
    load().oracle_center_of_mass(pos.shape[0], pos.reshape(-1),
                                 np.ascontiguousarray(mass, np.float32), com)
    return com


def explode(pos, vel, mass, com, force):
    """Velocities after the explosion impulse about `com`."""
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32).copy()
    load().oracle_explode(pos.shape[0], pos.reshape(-1), vel.reshape(-1),
                          np.ascontiguousarray(mass, np.float32),
                          np.ascontiguousarray(com, np.float32), float(force))
    return vel


def pick_sphere(pos, center, radius):
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.zeros(max(pos.shape[0], 1), np.int32)
    c = load().oracle_pick_sphere(pos.shape[0], pos.reshape(-1),
                                  np.ascontiguousarray(center, np.float32), float(radius), out)
    return out[:c].copy()


def move_picked(pos, idx, delta, lo=(0.01, 0.01, 0.01), hi=(0.99, 0.99, 0.99)):
    pos = np.ascontiguousarray(pos, np.float32).copy()
    idx = np.ascontiguousarray(idx, np.int32)
    f3 = lambda v: np.ascontiguousarray(v, np.float32)
    load().oracle_move_picked(pos.reshape(-1), idx, idx.shape[0], f3(delta), f3(lo), f3(hi))
    return pos
