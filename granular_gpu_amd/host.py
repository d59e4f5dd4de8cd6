"""ctypes bindings of the host C++ API's C-ABI (libgranular_host.so, declared
in include/granular_host.h): the GranularSystem class behind a handle, the
scene JSON reader and the scene builders.

HostSystem mirrors the reference's GranularSystem (GranularSystem.hpp:8-56):
the constructor runs one step, step() returns 1 and frame times accumulate.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libgranular_host.so"
_lib = None

FIELDS = {"pos": (0, np.float32, 3), "vel": (1, np.float32, 3), "mass": (2, np.float32, 1),
          "particle_2_cell": (3, np.int32, 1), "surface": (4, np.int32, 1),
          "surface_distance": (5, np.float32, 1), "adaptive_last_step": (6, np.int32, 1),
          "is_animated": (7, np.int32, 1)}
SETS = {"granular": 0, "boundary": 1, "fine": 2}


class ggh_scene(C.Structure):
    _fields_ = [("space_size", C.c_float * 3), ("dt", C.c_float), ("gravity", C.c_float * 3),
                ("sph_spacing", C.c_float), ("init_spacing", C.c_float),
                ("smoothing_radius", C.c_float), ("cell_length", C.c_float),
                ("cell_size", C.c_int * 3), ("density", C.c_float), ("box_size", C.c_float * 3),
                ("boundary_translation", C.c_float * 3), ("particle_translation", C.c_float * 3),
                ("scene", C.c_int), ("is_adaptive", C.c_int), ("obj_file", C.c_char * 256),
                ("save_file", C.c_char * 256)]


class ggh_animation(C.Structure):
    _fields_ = [("is_animating", C.c_int), ("animation_time", C.c_float),
                ("start_index", C.c_int), ("num_particles", C.c_int),
                ("translation_direction", C.c_float * 3), ("translation_speed", C.c_float),
                ("translation_complete", C.c_int), ("rotation_angle", C.c_float),
                ("rotation_speed", C.c_float), ("rotation_center", C.c_float * 3),
                ("translation_time", C.c_float), ("rotation_limit", C.c_float),
                ("center_from_marked", C.c_int)]


class HostError(RuntimeError):
    pass


def load() -> C.CDLL:
    """Loads libgranular_host.so (built by __graft_entry__.build()); raises
    when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise HostError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
    from . import _lib as gpu
    gpu.load()  # resolves libgranular_gpu.so first
    lib = C.CDLL(str(_LIB_PATH))
    fp, ip, vp = C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_void_p
    S = C.POINTER(ggh_scene)
    sig = {
        "ggh_last_error": (C.c_char_p, []),
        "ggh_load_scene": (C.c_int, [C.c_char_p, S]),
        "ggh_parse_scene": (C.c_int, [C.c_char_p, S]),
        "ggh_scene_lattice": (C.c_int, [S, vp, C.c_int]),
        "ggh_scene_walls": (C.c_int, [S, vp, C.c_int]),
        "ggh_upsampled_cloud": (C.c_int, [vp, C.c_int, C.c_float, C.c_int, vp, C.c_int]),
        "ggh_animation_default": (None, [C.POINTER(ggh_animation)]),
        "ggh_scene_boundary": (C.c_int, [S, vp, vp, C.c_int, C.POINTER(ggh_animation)]),
        "ggh_system_animate": (C.c_int, [vp, fp, C.c_float, fp]),
        "ggh_system_update_animation": (C.c_int, [vp, C.POINTER(ggh_animation), C.c_float]),
        "ggh_system_explode": (C.c_int, [vp, C.c_float, fp]),
        "ggh_system_pick": (C.c_int, [vp, fp, C.c_float, ip, C.c_int]),
        "ggh_system_move_picked": (C.c_int, [vp, ip, C.c_int, fp, fp, fp]),
        "ggh_system_create": (vp, [vp, C.c_int, vp, C.c_int, vp, vp, C.c_int, fp, C.c_float,
                                   C.c_float, fp, ip, C.c_float, C.c_float, C.c_int, C.c_int,
                                   C.c_int]),
        "ggh_system_destroy": (None, [vp]),
        "ggh_system_step": (C.c_int, [vp, C.c_int]),
        "ggh_system_size": (C.c_int, [vp, C.c_int]),
        "ggh_system_read": (C.c_int, [vp, C.c_int, C.c_int, vp, C.c_int]),
        "ggh_system_translate_animated": (C.c_int, [vp, C.c_float, C.c_float, C.c_float]),
        "ggh_system_frame_times": (C.c_int, [vp, fp, C.c_int]),
        "ggh_system_particle_stats": (C.c_int, [vp, ip, C.c_int]),
        "ggh_system_sweep_scans": (C.c_int, [vp, ip, C.c_int]),
        "ggh_system_save_stats": (C.c_int, [vp, C.c_char_p]),
        "ggh_system_set_solver": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float,
                                            C.c_int]),
        "ggh_system_set_lists": (C.c_int, [vp, C.c_int, C.c_int, C.c_float]),
        "ggh_system_set_merge_rules": (C.c_int, [vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                                 C.c_float, C.c_float, C.c_float]),
        "ggh_system_set_learned": (C.c_int, [vp, vp, vp, C.c_float, C.c_float, C.c_float,
                                             C.c_float, C.c_float]),
        "ggh_darray_selftest": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
        "ggh_solver_stages_selftest": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
        "ggh_write_vtk": (C.c_int, [C.c_char_p, C.c_char_p, vp, C.c_int]),
        "ggh_system_save_upsampled_vtk": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int]),
        "ggh_system_save_boundary_vtk": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_int]),
        "ggh_system_save_mass_vtk": (C.c_int, [vp, C.c_char_p, C.c_int, ip]),
        "ggh_checkpoint_write": (C.c_int, [C.c_char_p, vp, vp, vp, vp, C.c_int, vp, vp, C.c_int,
                                           vp, vp, C.c_int]),
        "ggh_system_save_checkpoint": (C.c_int, [vp, C.c_char_p]),
        "ggh_checkpoint_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, vp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _err(lib) -> str:
    m = lib.ggh_last_error()
    return m.decode() if m else "unknown error"


def _chk(lib, rc, what):
    if rc is None or (isinstance(rc, int) and rc < 0):
        raise HostError(f"{what}: {_err(lib)}")
    return rc


def _f3(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 3))


def load_scene(path) -> ggh_scene:
    lib = load()
    s = ggh_scene()
    _chk(lib, lib.ggh_load_scene(str(path).encode(), C.byref(s)), "ggh_load_scene")
    return s


def parse_scene(text: str) -> ggh_scene:
    lib = load()
    s = ggh_scene()
    _chk(lib, lib.ggh_parse_scene(text.encode(), C.byref(s)), "ggh_parse_scene")
    return s


def _emit(fn, *args) -> np.ndarray:
    lib = load()
    n = _chk(lib, fn(*args, None, 0), fn.__name__)
    out = np.zeros((n, 3), np.float32)
    if n:
        _chk(lib, fn(*args, out.ctypes.data, n), fn.__name__)
    return out


def scene_lattice(s: ggh_scene) -> np.ndarray:
    return _emit(load().ggh_scene_lattice, C.byref(s))


def scene_walls(s: ggh_scene) -> np.ndarray:
    return _emit(load().ggh_scene_walls, C.byref(s))


def animation_default() -> ggh_animation:
    """The reference's AnimationState defaults (main.cu:100-111)."""
    a = ggh_animation()
    load().ggh_animation_default(C.byref(a))
    return a


def scene_boundary(s: ggh_scene):
    """Boundary set of the scene: (positions (n, 3), animated markers (n,),
    animation state).  EXCAVATOR / FUNNEL add the obj_file point cloud when
    that file exists."""
    lib = load()
    n = _chk(lib, lib.ggh_scene_boundary(C.byref(s), None, None, 0, None), "ggh_scene_boundary")
    pos = np.zeros((n, 3), np.float32)
    animated = np.zeros(n, np.int32)
    anim = ggh_animation()
    _chk(lib, lib.ggh_scene_boundary(C.byref(s), pos.ctypes.data, animated.ctypes.data, n,
                                     C.byref(anim)), "ggh_scene_boundary")
    return pos, animated, anim


def upsampled_cloud(coarse, radius: float, per: int = 25) -> np.ndarray:
    c = _f3(coarse)
    return _emit(load().ggh_upsampled_cloud, c.ctypes.data, c.shape[0], float(radius), int(per))


def darray_selftest(n: int) -> str:
    lib = load()
    buf = C.create_string_buffer(512)
    rc = lib.ggh_darray_selftest(int(n), buf, 512)
    msg = buf.value.decode()
    if rc != 0:
        raise HostError(f"DArray self-test (n={n}) failed: {msg}")
    return msg


def solver_stages_selftest(n: int) -> str:
    """C++ self-test of the Solver's separate stage methods on n particles."""
    lib = load()
    buf = C.create_string_buffer(512)
    rc = lib.ggh_solver_stages_selftest(int(n), buf, 512)
    msg = buf.value.decode()
    if rc != 0:
        raise HostError(f"Solver stage self-test (n={n}) failed: {msg}")
    return msg


def write_vtk(path, points, title="Upsampled Granular Particles") -> int:
    """The reference's legacy-ASCII VTK point file (NaN points dropped, y/z
    swapped); returns the number of points written."""
    lib = load()
    p = _f3(points)
    return _chk(lib, lib.ggh_write_vtk(str(path).encode(), title.encode(), p.ctypes.data,
                                       p.shape[0]), "ggh_write_vtk")


def write_checkpoint(path, pos, vel, mass, adaptive_last_step, boundary, boundary_animated,
                     fine_pos, fine_vel):
    lib = load()
    p, v, b, fp, fv = _f3(pos), _f3(vel), _f3(boundary), _f3(fine_pos), _f3(fine_vel)
    m = np.ascontiguousarray(mass, np.float32)
    a = np.ascontiguousarray(adaptive_last_step, np.int32)
    ba = np.ascontiguousarray(boundary_animated, np.int32)
    _chk(lib, lib.ggh_checkpoint_write(str(path).encode(), p.ctypes.data, v.ctypes.data,
                                       m.ctypes.data, a.ctypes.data, p.shape[0], b.ctypes.data,
                                       ba.ctypes.data, b.shape[0], fp.ctypes.data,
                                       fv.ctypes.data, fp.shape[0]), "ggh_checkpoint_write")


def read_checkpoint(path) -> dict:
    """All arrays of a checkpoint file, keyed (set, field)."""
    lib = load()
    out = {}
    for which, field in (("granular", "pos"), ("granular", "vel"), ("granular", "mass"),
                         ("granular", "adaptive_last_step"), ("boundary", "pos"),
                         ("boundary", "is_animated"), ("fine", "pos"), ("fine", "vel")):
        code, dt, w = FIELDS[field]
        n = _chk(lib, lib.ggh_checkpoint_read(str(path).encode(), SETS[which], code, None, 0),
                 "ggh_checkpoint_read")
        a = np.zeros((n, w) if w > 1 else n, dt)
        _chk(lib, lib.ggh_checkpoint_read(str(path).encode(), SETS[which], code, a.ctypes.data,
                                          n), "ggh_checkpoint_read")
        out[(which, field)] = a
    return out


class HostSystem:
    """One GranularSystem (the reference's host class) behind its C handle."""

    def __init__(self, pos, boundary, fine, space_size, cell_length, dt, gravity, cell_size,
                 density, upsampled_radius=0.005, is_move_boundary=False, is_adaptive=0,
                 boundary_animated=None, verbose=False):
        lib = load()
        self.lib = lib
        p, b, f = _f3(pos), _f3(boundary), _f3(fine)
        anim = None
        if boundary_animated is not None:
            anim = np.ascontiguousarray(boundary_animated, np.int32)
        sp = (C.c_float * 3)(*[float(v) for v in space_size])
        g = (C.c_float * 3)(*[float(v) for v in gravity])
        cs = (C.c_int * 3)(*[int(v) for v in cell_size])
        h = lib.ggh_system_create(p.ctypes.data, p.shape[0], b.ctypes.data, b.shape[0],
                                  None if anim is None else anim.ctypes.data, f.ctypes.data,
                                  f.shape[0], sp, float(cell_length), float(dt), g, cs,
                                  float(density), float(upsampled_radius),
                                  int(bool(is_move_boundary)), int(is_adaptive), int(verbose))
        if not h:
            raise HostError(f"ggh_system_create: {_err(lib)}")
        self.h = C.c_void_p(h)

    @classmethod
    def from_scene(cls, scene, fine=None, **kw):
        """From a granular_gpu_amd.scenes.Scene (fine particles optional)."""
        f = np.zeros((0, 3), np.float32) if fine is None else fine
        return cls(scene.pos, scene.boundary, f, scene.space, scene.cell_length, scene.dt,
                   scene.gravity, scene.cells, scene.density, **kw)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ggh_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps=1) -> int:
        return _chk(self.lib, self.lib.ggh_system_step(self.h, int(nsteps)), "ggh_system_step")

    def size(self, which="granular") -> int:
        k = {"granular": 0, "boundary": 1, "fine": 2, "total": 3}[which]
        return self.lib.ggh_system_size(self.h, k)

    def read(self, which: str, field: str, count: int | None = None) -> np.ndarray:
        code, dt, w = FIELDS[field]
        n = self.size(which) if count is None else count
        out = np.zeros((n, w) if w > 1 else n, dt)
        _chk(self.lib, self.lib.ggh_system_read(self.h, SETS[which], code, out.ctypes.data, n),
             "ggh_system_read")
        return out

    def translate_animated(self, d) -> int:
        return _chk(self.lib, self.lib.ggh_system_translate_animated(self.h, *map(float, d)),
                    "ggh_system_translate_animated")

    def animate(self, translation, angle_deg=0.0, center=(0.0, 0.0, 0.0)):
        """animateParticles on the boundary: translate the animated particles,
        then rotate them by angle_deg about z through center."""
        t = (C.c_float * 3)(*[float(v) for v in translation])
        c = (C.c_float * 3)(*[float(v) for v in center])
        _chk(self.lib, self.lib.ggh_system_animate(self.h, t, float(angle_deg), c),
             "ggh_system_animate")

    def explode(self, force: float = 50.0) -> np.ndarray:
        """Solver::trigger_explosion on the granular set; returns the centre
        of mass it used."""
        com = (C.c_float * 3)()
        _chk(self.lib, self.lib.ggh_system_explode(self.h, float(force), com),
             "ggh_system_explode")
        return np.array(com, np.float32)

    def pick(self, center, radius: float) -> np.ndarray:
        """Indices of the granular particles within radius of center,
        ascending."""
        c = (C.c_float * 3)(*[float(v) for v in center])
        n = _chk(self.lib, self.lib.ggh_system_pick(self.h, c, float(radius), None, 0),
                 "ggh_system_pick")
        out = np.zeros(max(n, 1), np.int32)
        _chk(self.lib, self.lib.ggh_system_pick(self.h, c, float(radius),
                                                out.ctypes.data_as(C.POINTER(C.c_int)), n),
             "ggh_system_pick")
        return out[:n]

    def move_picked(self, indices, delta, lo=(0.01, 0.01, 0.01), hi=(0.99, 0.99, 0.99)) -> int:
        idx = np.ascontiguousarray(indices, np.int32)
        f3 = lambda v: (C.c_float * 3)(*[float(x) for x in v])
        return _chk(self.lib, self.lib.ggh_system_move_picked(
            self.h, idx.ctypes.data_as(C.POINTER(C.c_int)), idx.shape[0], f3(delta), f3(lo),
            f3(hi)), "ggh_system_move_picked")

    def update_animation(self, anim: ggh_animation, dt: float) -> bool:
        """One updateAnimation(dt) call advancing `anim`; True if the
        boundary moved."""
        return bool(_chk(self.lib, self.lib.ggh_system_update_animation(
            self.h, C.byref(anim), float(dt)), "ggh_system_update_animation"))

    def frame_times(self) -> np.ndarray:
        n = self.lib.ggh_system_frame_times(self.h, None, 0)
        out = np.zeros(n, np.float32)
        self.lib.ggh_system_frame_times(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), n)
        return out

    def particle_stats(self) -> np.ndarray:
        n = self.lib.ggh_system_particle_stats(self.h, None, 0)
        out = np.zeros((n, 3), np.int32)
        self.lib.ggh_system_particle_stats(self.h, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out

    def sweep_scans(self) -> np.ndarray:
        """Particles that scanned their whole stencil in each sweep of the
        last step."""
        k = self.lib.ggh_system_sweep_scans(self.h, None, 0)
        out = np.zeros(max(k, 0), np.int32)
        rc = self.lib.ggh_system_sweep_scans(self.h, out.ctypes.data_as(C.POINTER(C.c_int)), k)
        if rc < 0:
            _chk(self.lib, rc, "ggh_system_sweep_scans")
        return out

    def save_stats(self, path):
        _chk(self.lib, self.lib.ggh_system_save_stats(self.h, str(path).encode()),
             "ggh_system_save_stats")

    def save_upsampled_vtk(self, save_file, frame, is_adaptive=False) -> int:
        return _chk(self.lib, self.lib.ggh_system_save_upsampled_vtk(
            self.h, str(save_file).encode(), int(bool(is_adaptive)), int(frame)),
            "ggh_system_save_upsampled_vtk")

    def save_boundary_vtk(self, save_file, frame, centred=False) -> int:
        """The animated boundary particles as a VTK file; centred uses the
        fine set's centroid and lowest point."""
        return _chk(self.lib, self.lib.ggh_system_save_boundary_vtk(
            self.h, str(save_file).encode(), int(frame), int(bool(centred))),
            "ggh_system_save_boundary_vtk")

    def save_mass_vtk(self, save_file, frame) -> np.ndarray:
        counts = np.zeros(3, np.int32)
        _chk(self.lib, self.lib.ggh_system_save_mass_vtk(
            self.h, str(save_file).encode(), int(frame),
            counts.ctypes.data_as(C.POINTER(C.c_int))), "ggh_system_save_mass_vtk")
        return counts

    def save_checkpoint(self, path):
        _chk(self.lib, self.lib.ggh_system_save_checkpoint(self.h, str(path).encode()),
             "ggh_system_save_checkpoint")

    def set_solver(self, stab=3, solver=5, mu_s=0.0, mu_k=0.0, flags=0):
        _chk(self.lib, self.lib.ggh_system_set_solver(self.h, stab, solver, mu_s, mu_k, flags),
             "ggh_system_set_solver")

    def set_lists(self, gmax=48, bmax=16, skin=0.0075):
        """Solver::set_neighbor_lists: list capacities and skin."""
        _chk(self.lib, self.lib.ggh_system_set_lists(self.h, int(gmax), int(bmax), float(skin)),
             "ggh_system_set_lists")

    def set_merge_rules(self, depth_rules=1, min_mass=0.125, split_mass=0.125, merge_dist=1.75,
                        merge_depth=3.0, split_depth=2.0, max_mass=8.0):
        """GranularSystem::set_merge_rules (defaults: the radius-level rules
        for masses 0.125 .. 8)."""
        _chk(self.lib, self.lib.ggh_system_set_merge_rules(
            self.h, int(depth_rules), min_mass, split_mass, merge_dist, merge_depth, split_depth,
            max_mass), "ggh_system_set_merge_rules")

    def set_learned(self, w_bf16: np.ndarray, bias: np.ndarray, radius=0.02, vscale=1.0,
                    guard_boundary=0.05, guard_floor=0.08, guard_side=0.05):
        """GranularSystem::set_learned_upsampler: w_bf16 the 49,152 bfloat16
        bit patterns (uint16), bias the 416 float biases."""
        w = np.ascontiguousarray(w_bf16, np.uint16)
        b = np.ascontiguousarray(bias, np.float32)
        if w.size != 49152 or b.size != 416:
            raise ValueError("set_learned: 49,152 weights and 416 biases expected")
        _chk(self.lib, self.lib.ggh_system_set_learned(
            self.h, w.ctypes.data, b.ctypes.data, radius, vscale, guard_boundary, guard_floor,
            guard_side), "ggh_system_set_learned")
