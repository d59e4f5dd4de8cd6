"""ctypes binding of libgranular_gpu.so (the C-ABI in include/granular_gpu.h).

The library is the product: every compute call goes to its HIP kernels.  If
the shared object is missing this module raises immediately; there is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libgranular_gpu.so"

GG_OK = 0
GG_ERR_INVALID = -1
GG_ERR_CAPACITY = -2
GG_ERR_HIP = -3
GG_ERR_WORKSPACE = -4
GG_ERR_NO_DEVICE = -5
GG_FLAG_NO_NLIST = 1
GG_FLAG_NO_XCD_REMAP = 2
GG_FLAG_FALLBACK_LANE = 4
GG_FLAG_FALLBACK_WAVE = 8
GG_FLAG_LIST_REFRESH = 16
STAGES = ["sort", "predict", "nlist", "sweep", "finalize", "stab_sweep", "adaptive", "upsample",
          "upsample_sort", "mlp", "mlp_features", "mlp_fallback", "sweep_queue"]


class gg_float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class gg_grid(C.Structure):
    _fields_ = [("cell_length", C.c_float), ("cx", C.c_int), ("cy", C.c_int), ("cz", C.c_int)]


class gg_params(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gx", C.c_float), ("gy", C.c_float), ("gz", C.c_float),
        ("density", C.c_float), ("stab_iters", C.c_int), ("solver_iters", C.c_int),
        ("mu_s", C.c_float), ("mu_k", C.c_float), ("boundary_radius", C.c_float),
        ("contact_radius_min", C.c_float), ("sleep_speed", C.c_float),
        ("damp_speed", C.c_float), ("damping", C.c_float), ("adaptive_damping", C.c_float),
        ("skin", C.c_float), ("flags", C.c_int), ("sweep_block", C.c_int),
        ("nlist_block", C.c_int), ("adaptive_rules", C.c_int), ("min_mass", C.c_float),
        ("split_mass", C.c_float), ("merge_dist", C.c_float), ("merge_depth", C.c_float),
        ("split_depth", C.c_float),
    ]


class gg_particles(C.Structure):
    _fields_ = [
        ("n", C.c_int), ("capacity", C.c_int), ("pos", C.c_void_p), ("vel", C.c_void_p),
        ("mass", C.c_void_p), ("scaled_mass", C.c_void_p), ("surface", C.c_void_p),
        ("surface_distance", C.c_void_p), ("num_surface", C.c_void_p),
        ("particle_2_cell", C.c_void_p), ("is_animated", C.c_void_p),
        ("adaptive_last_step", C.c_void_p),
    ]


class gg_adaptive(C.Structure):
    _fields_ = [("remove", C.c_void_p), ("merge", C.c_void_p), ("merge_velocity", C.c_void_p),
                ("merge_count", C.c_void_p)]


class gg_upsample_params(C.Structure):
    _fields_ = [(k, C.c_float) for k in (
        "dt", "gx", "gy", "gz", "r9", "boundary_radius", "repulsion", "min_weight",
        "near_damping", "ground_y", "ground_damping", "max_velocity", "floor_y",
        "floor_friction")]


class gg_workspace(C.Structure):
    _fields_ = [
        ("base", C.c_void_p), ("bytes", C.c_size_t), ("capacity", C.c_int),
        ("boundary_capacity", C.c_int), ("grid", gg_grid), ("max_sweeps", C.c_int),
        ("gmax", C.c_int), ("bmax", C.c_int),
    ]


class gg_mlp_config(C.Structure):
    _fields_ = [(k, C.c_float) for k in ("radius", "vscale", "guard_boundary", "guard_floor",
                                         "guard_side")] + [("infer_waves", C.c_int),
                                                           ("infer_tiles", C.c_int),
                                                           ("select_chunk", C.c_int)]


class gg_mlp_weights(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("w_bf16", "bias", "w_master", "m_w", "v_w",
                                          "b_master", "m_b", "v_b")]


class gg_batch_scene(C.Structure):
    _fields_ = [("set", gg_particles), ("cell_start", C.c_void_p), ("boundary", gg_particles),
                ("bcell_start", C.c_void_p), ("params", gg_params), ("ws", gg_workspace)]


class gg_scene_desc(C.Structure):
    _fields_ = [("cx", C.c_int), ("cy", C.c_int), ("cz", C.c_int), ("cell_base", C.c_int),
                ("mu_s", C.c_float), ("mu_k", C.c_float)]


class gg_scenes(C.Structure):
    _fields_ = [("count", C.c_int), ("desc", C.c_void_p), ("host_desc", C.c_void_p),
                ("scene_of", C.c_void_p), ("bscene_of", C.c_void_p), ("friction", C.c_int)]


class gg_batch_config(C.Structure):
    _fields_ = [("branches", C.c_int), ("sweep_block", C.c_int), ("nlist_block", C.c_int),
                ("steps_per_launch", C.c_int)]


P = C.POINTER
VP = C.c_void_p

# name -> (restype, argtypes); mirrors include/granular_gpu.h
_PROTOS = {
    "gg_version": (C.c_char_p, []),
    "gg_status_string": (C.c_char_p, [C.c_int]),
    "gg_device_check": (C.c_int, []),
    "gg_params_default": (None, [P(gg_params)]),
    "gg_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, gg_grid, C.c_int, C.c_int, C.c_int]),
    "gg_workspace_init": (C.c_int, [P(gg_workspace), VP]),
    "gg_grid_build": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, VP, P(gg_workspace), VP]),
    "gg_boundary_build": (C.c_int, [P(gg_particles), VP, P(gg_workspace), VP]),
    "gg_step": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, P(gg_particles), VP, P(gg_params),
                          P(gg_workspace), VP]),
    "gg_step_noalloc": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, P(gg_particles), VP,
                                  gg_grid, P(gg_params), VP]),
    "gg_predict": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, P(gg_params),
                             P(gg_workspace), VP]),
    "gg_contact_sweep": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, P(gg_params), C.c_int,
                                   P(gg_workspace), VP]),
    "gg_finalize": (C.c_int, [P(gg_particles), P(gg_params), P(gg_workspace), VP]),
    "gg_workspace_view": (VP, [P(gg_workspace), C.c_int]),
    "gg_add_external_force": (C.c_int, [P(gg_particles), gg_float3, VP]),
    "gg_predict_positions": (C.c_int, [P(gg_particles), C.c_float, VP, VP]),
    "gg_mass_scaling": (C.c_int, [P(gg_particles), VP]),
    "gg_sort_positions_by_key": (C.c_int, [P(gg_particles), VP, P(gg_workspace), VP]),
    "gg_surface": (C.c_int, [P(gg_particles), VP, P(gg_params), P(gg_workspace), VP]),
    "gg_surface_distance": (C.c_int, [P(gg_particles), VP, P(gg_workspace), VP]),
    "gg_merge_mark": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, C.c_float, P(gg_workspace),
                                VP]),
    "gg_merge_blend": (C.c_int, [P(gg_particles), P(gg_adaptive), C.c_int, P(gg_workspace), VP]),
    "gg_split_compact": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, P(gg_particles), VP,
                                   P(gg_params), P(gg_workspace), VP, P(C.c_int)]),
    "gg_adaptive_sampling": (C.c_int, [P(gg_particles), P(gg_adaptive), VP, P(gg_particles), VP,
                                       P(gg_params), C.c_float, C.c_int, C.c_int,
                                       P(gg_workspace), VP, P(C.c_int)]),
    "gg_upsample_params_default": (None, [P(gg_upsample_params)]),
    "gg_upsampled_sort": (C.c_int, [P(gg_particles), VP, P(gg_workspace), VP]),
    "gg_upsample_plain": (C.c_int, [P(gg_particles), P(gg_particles), VP, P(gg_particles), VP,
                                    gg_grid, P(gg_upsample_params), C.c_int, VP]),
    "gg_batch_create": (C.c_int, [P(gg_batch_scene), C.c_int, P(gg_batch_config), P(VP)]),
    "gg_batch_step": (C.c_int, [VP, VP]),
    "gg_batch_step_eager": (C.c_int, [VP, VP]),
    "gg_batch_get_config": (C.c_int, [VP, P(gg_batch_config)]),
    "gg_batch_set_config": (C.c_int, [VP, P(gg_batch_config)]),
    "gg_batch_autotune": (C.c_int, [VP, C.c_int, VP, P(gg_batch_config), P(C.c_float)]),
    "gg_batch_destroy": (None, [VP]),
    "gg_scenes_boundary_build": (C.c_int, [P(gg_particles), VP, P(gg_scenes), P(gg_workspace),
                                           VP]),
    "gg_step_scenes": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, P(gg_scenes),
                                 P(gg_params), P(gg_workspace), VP]),
    "gg_scenes_graph_create": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP,
                                         P(gg_scenes), P(gg_params), P(gg_workspace), C.c_int,
                                         P(VP)]),
    "gg_scenes_graph_launch": (C.c_int, [VP, VP]),
    "gg_scenes_graph_autotune": (C.c_int, [VP, C.c_int, VP, P(C.c_int), P(C.c_int),
                                           P(C.c_float)]),
    "gg_scenes_graph_destroy": (None, [VP]),
    "gg_step_graph_create": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, P(gg_params),
                                       P(gg_workspace), C.c_int, P(VP)]),
    "gg_step_graph_launch": (C.c_int, [VP, VP]),
    "gg_step_graph_destroy": (None, [VP]),
    "gg_mlp_config_default": (None, [P(gg_mlp_config)]),
    "gg_mlp_features": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, gg_grid,
                                  P(gg_mlp_config), VP, VP, VP, VP]),
    "gg_mlp_teacher": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, gg_grid,
                                 P(gg_upsample_params), P(gg_mlp_config), VP, VP, VP, VP, VP]),
    "gg_mlp_fine_targets": (C.c_int, [P(gg_particles), P(gg_particles), VP, gg_grid,
                                      P(gg_mlp_config), VP, VP, VP, VP]),
    "gg_mlp_train_scratch_bytes": (C.c_size_t, [C.c_int]),
    "gg_mlp_forward": (C.c_int, [P(gg_mlp_weights), VP, C.c_int, VP, VP, C.c_size_t, VP]),
    "gg_mlp_grad": (C.c_int, [P(gg_mlp_weights), VP, VP, C.c_int, VP, VP, VP, VP, C.c_size_t,
                              VP]),
    "gg_mlp_adam": (C.c_int, [P(gg_mlp_weights), VP, VP, C.c_float, C.c_float, C.c_float,
                              C.c_float, C.c_int, VP]),
    "gg_mlp_sync_weights": (C.c_int, [P(gg_mlp_weights), VP]),
    "gg_upsample_mlp_scratch_bytes": (C.c_size_t, [C.c_int]),
    "gg_upsample_mlp": (C.c_int, [P(gg_particles), VP, P(gg_particles), VP, gg_grid,
                                  P(gg_upsample_params), P(gg_mlp_config), VP, VP, VP, VP, VP,
                                  C.c_size_t, VP, VP]),
    "gg_mfma_probe": (C.c_int, [VP, VP, VP, VP]),
    "gg_bf16_round_check": (C.c_int, [C.c_uint64, C.c_uint64, VP, VP]),
    "gg_fast_math_check": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, VP, VP]),
    "gg_compact_scratch_bytes": (C.c_size_t, [C.c_longlong]),
    "gg_compact": (C.c_int, [VP, VP, C.c_longlong, C.c_int, VP, VP, VP, C.c_size_t, VP]),
    "gg_fill_u32": (C.c_int, [VP, C.c_size_t, C.c_uint32, VP]),
    "gg_profiler_enable": (C.c_int, [C.c_int]),
    "gg_profiler_pause": (C.c_int, [C.c_int]),
    "gg_profiler_collect": (C.c_int, [P(C.c_double), P(C.c_int), C.c_int]),
}

_lib = None


def declared_symbols() -> list[str]:
    """Names of the entry points this binding declares."""
    return sorted(_PROTOS)


def load() -> C.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} not found: build it with `make -C granular_gpu_amd/csrc` "
            "(or __graft_entry__.build()); there is no non-HIP fallback")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_GLOBAL if hasattr(os, "RTLD_GLOBAL") else 0)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class GGError(RuntimeError):
    def __init__(self, fn: str, status: int):
        msg = load().gg_status_string(status)
        super().__init__(f"{fn} failed: {status} ({msg.decode() if msg else '?'})")
        self.status = status


def check(fn: str, status: int) -> None:
    if status != GG_OK:
        raise GGError(fn, status)


def default_params() -> gg_params:
    p = gg_params()
    load().gg_params_default(C.byref(p))
    return p
