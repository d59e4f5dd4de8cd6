"""The C-ABI library loads without a GPU and exports every entry point that
include/granular_gpu.h declares (no compute calls)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    txt = (ROOT / "include" / "granular_gpu.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gg_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = header_functions()
    for must in ("gg_grid_build", "gg_step", "gg_contact_sweep", "gg_predict", "gg_finalize",
                 "gg_boundary_build", "gg_workspace_bytes"):
        assert must in names


def test_library_exports_all_header_symbols():
    from granular_gpu_amd import _lib
    lib = _lib.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (gg_[a-z0-9_]+)", out))
    assert set(header_functions()) <= exported


def test_binding_matches_header():
    from granular_gpu_amd import _lib
    assert set(_lib.declared_symbols()) <= set(header_functions())


def test_host_only_calls():
    from granular_gpu_amd import _lib
    lib = _lib.load()
    assert lib.gg_version().startswith(b"granular_gpu_amd")
    p = _lib.default_params()
    assert p.stab_iters == 3 and p.solver_iters == 5
    assert abs(p.density - 238732.4146) < 0.1
    g = _lib.gg_grid(0.0404, 25, 25, 25)
    assert lib.gg_workspace_bytes(32768, 14408, g, 13, 48, 16) > 0
    assert lib.gg_workspace_bytes(-1, 0, g, 13, 48, 16) == 0
    assert lib.gg_status_string(-2) == b"capacity exceeded"
    # invalid workspace is rejected before any device work
    ws = _lib.gg_workspace()
    assert lib.gg_workspace_init(C.byref(ws), None) == _lib.GG_ERR_INVALID


def test_motion_and_extras_reject_bad_arguments_on_host():
    # argument checks of the animation, explosion and picking entry points
    # return before any device work
    from granular_gpu_amd import _lib
    lib = _lib.load()

    class F3(C.Structure):
        _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]
    P = C.POINTER(_lib.gg_particles)
    lib.gg_boundary_animate.argtypes = [P, F3, C.c_float, F3, C.c_void_p]
    lib.gg_explosion_scratch_bytes.restype = C.c_size_t
    lib.gg_explosion_scratch_bytes.argtypes = [C.c_int]
    lib.gg_explosion.argtypes = [P, C.c_float, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    lib.gg_pick_scratch_bytes.restype = C.c_size_t
    lib.gg_pick_scratch_bytes.argtypes = [C.c_int]
    lib.gg_pick_sphere.argtypes = [P, F3, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_size_t, C.c_void_p]
    lib.gg_move_picked.argtypes = [P, C.c_void_p, C.c_void_p, C.c_int, F3, F3, F3, C.c_void_p]
    z = F3(0.0, 0.0, 0.0)
    s = _lib.gg_particles()
    assert lib.gg_boundary_animate(None, z, 10.0, z, None) == _lib.GG_ERR_INVALID
    assert lib.gg_boundary_animate(C.byref(s), z, 10.0, z, None) == _lib.GG_OK  # n = 0
    s.n = 5
    assert lib.gg_boundary_animate(C.byref(s), z, 10.0, z, None) == _lib.GG_ERR_INVALID
    assert lib.gg_explosion_scratch_bytes(1000) >= 4 * 8 * 4
    assert lib.gg_explosion(C.byref(s), 50.0, None, 0, None, None) == _lib.GG_ERR_INVALID
    s.pos = s.vel = s.mass = 1
    assert lib.gg_explosion(C.byref(s), 50.0, 1, 0, None, None) == _lib.GG_ERR_WORKSPACE
    assert lib.gg_pick_scratch_bytes(1000) >= 2 * 4 * 1000
    assert lib.gg_pick_sphere(C.byref(s), z, 1.0, 1, None, 1, 1 << 20, None) == \
        _lib.GG_ERR_INVALID
    assert lib.gg_move_picked(C.byref(s), 1, 1, 6, z, z, z, None) == _lib.GG_ERR_INVALID
    assert lib.gg_move_picked(C.byref(s), 1, 1, -1, z, z, z, None) == _lib.GG_ERR_INVALID


def host_header_functions():
    txt = (ROOT / "include" / "granular_host.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ggh_[a-z0-9_]+)\s*\(", txt)))


def test_host_library_exports_all_header_symbols():
    from granular_gpu_amd import host
    lib = host.load()
    names = host_header_functions()
    assert "ggh_system_create" in names and "ggh_load_scene" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(host._LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    assert set(names) <= set(re.findall(r"\bT (ggh_[a-z0-9_]+)", out))


def test_cpp_headers_compile_standalone(tmp_path):
    # every public C++ header of the host API compiles on its own
    for h in sorted((ROOT / "include" / "granular").glob("*.hpp")):
        src = tmp_path / f"inc_{h.stem}.cpp"
        src.write_text(f'#include "granular/{h.name}"\nint main() {{ return 0; }}\n')
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", f"-I{ROOT / 'include'}", str(src)], check=True)


def test_scenes_and_fine_targets_reject_bad_arguments_on_host():
    # the multi-scene step, its graph and the fine-run targets check their
    # arguments (the scene table against the workspace's cell count) before
    # any device work
    import ctypes as C
    from granular_gpu_amd import _lib
    lib = _lib.load()
    g = _lib.gg_grid(0.0404, 9, 1, 1)  # T - 1 = 9 cells: two 2x2x1 scenes (4 + 1 each)
    ws = _lib.gg_workspace(C.c_void_p(16).value, 1 << 30, 10, 10, g, 8, 48, 16)
    desc = (_lib.gg_scene_desc * 2)(_lib.gg_scene_desc(2, 2, 1, 0, 0.5, 0.4),
                                    _lib.gg_scene_desc(2, 2, 1, 5, 0.5, 0.4))
    sc = _lib.gg_scenes(2, 16, C.addressof(desc), 16, 16, 1)
    sset = _lib.gg_particles()
    bset = _lib.gg_particles()
    p = _lib.default_params()
    bad = _lib.GG_ERR_INVALID if hasattr(_lib, "GG_ERR_INVALID") else -1
    # a scene table whose cell ranges do not add up to the workspace's
    desc[1].cell_base = 4
    assert lib.gg_step_scenes(C.byref(sset), C.c_void_p(16), C.byref(bset), C.c_void_p(16),
                              C.byref(sc), C.byref(p), C.byref(ws), None) == bad
    desc[1].cell_base = 5
    desc[1].cx = 3  # 3x2x1 + 1 overflows the 10 cells
    assert lib.gg_scenes_boundary_build(C.byref(bset), C.c_void_p(16), C.byref(sc), C.byref(ws),
                                        None) == bad
    desc[1].cx = 2
    desc[0].mu_s = -1.0  # negative friction
    h = C.c_void_p()
    assert lib.gg_scenes_graph_create(C.byref(sset), C.c_void_p(16), C.byref(bset),
                                      C.c_void_p(16), C.byref(sc), C.byref(p), C.byref(ws), 1,
                                      C.byref(h)) == bad
    assert not h.value
    assert lib.gg_scenes_graph_launch(None, None) == bad
    lib.gg_scenes_graph_destroy(None)
    # no host table
    sc2 = _lib.gg_scenes(2, 16, None, 16, 16, 1)
    assert lib.gg_step_scenes(C.byref(sset), C.c_void_p(16), C.byref(bset), C.c_void_p(16),
                              C.byref(sc2), C.byref(p), C.byref(ws), None) == bad
    # fine-run targets: missing outputs, bad grid, bad radius
    cfg = _lib.gg_mlp_config()
    lib.gg_mlp_config_default(C.byref(cfg))
    fg = _lib.gg_grid(0.0404, 4, 4, 4)
    assert lib.gg_mlp_fine_targets(C.byref(sset), C.byref(sset), C.c_void_p(16), fg,
                                   C.byref(cfg), None, None, None, None) == bad
    assert lib.gg_mlp_fine_targets(C.byref(sset), C.byref(sset), C.c_void_p(16),
                                   _lib.gg_grid(0.0, 4, 4, 4), C.byref(cfg), C.c_void_p(16),
                                   None, None, None) == bad
    cfg.radius = 0.0
    assert lib.gg_mlp_fine_targets(C.byref(sset), C.byref(sset), C.c_void_p(16), fg,
                                   C.byref(cfg), C.c_void_p(16), None, None, None) == bad
