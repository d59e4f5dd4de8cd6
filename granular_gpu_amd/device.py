"""Device-side driver over the C-ABI: owns the device arrays of one scene
(torch tensors used purely as device allocations) and calls the library's
stage and step entry points on torch's current HIP stream.

Array roles mirror the reference's GranularSystem (GranularSystem.hpp:60-90):
a granular set, a static boundary set, their cell tables and a workspace.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, load  # noqa: F401


def _ptr(t: torch.Tensor | None):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class PlainSim:
    """Plain (non-adaptive) scene on one GPU."""

    def __init__(self, scene, stab=3, solver=5, mu_s=0.0, mu_k=0.0, skin=0.0075, gmax=48,
                 bmax=16, flags=0, capacity=None, device="cuda", max_sweeps=None):
        lib = load()
        self.lib = lib
        self.scene = scene
        dev = torch.device(device)
        self.dev = dev
        n = scene.n
        cap = capacity or n
        nb = scene.boundary.shape[0]
        self.n, self.cap, self.nb = n, cap, nb
        f32, i32 = torch.float32, torch.int32
        self.pos = torch.zeros((cap, 3), dtype=f32, device=dev)
        self.vel = torch.zeros((cap, 3), dtype=f32, device=dev)
        self.mass = torch.zeros(cap, dtype=f32, device=dev)
        self.scaled_mass = torch.zeros(cap, dtype=f32, device=dev)
        self.p2c = torch.zeros(cap, dtype=i32, device=dev)
        self.adl = torch.zeros(cap, dtype=i32, device=dev)
        self.is_animated = torch.zeros(cap, dtype=i32, device=dev)
        self.pos[:n] = torch.from_numpy(scene.pos).to(dev)
        self.vel[:n] = torch.from_numpy(scene.vel).to(dev)
        self.mass[:n] = torch.from_numpy(scene.mass).to(dev)
        self.bpos = torch.from_numpy(np.ascontiguousarray(scene.boundary)).to(dev).contiguous()
        if nb == 0:
            self.bpos = torch.zeros((1, 3), dtype=f32, device=dev)
        self.bvel = torch.zeros_like(self.bpos)
        self.bmass = torch.ones(self.bpos.shape[0], dtype=f32, device=dev)
        self.bp2c = torch.zeros(self.bpos.shape[0], dtype=i32, device=dev)
        self.banim = torch.zeros(self.bpos.shape[0], dtype=i32, device=dev)
        cx, cy, cz = scene.cells
        self.C = cx * cy * cz
        self.cs = torch.zeros(self.C + 2, dtype=i32, device=dev)
        self.bcs = torch.zeros(self.C + 2, dtype=i32, device=dev)
        self.grid = _lib.gg_grid(float(scene.cell_length), cx, cy, cz)
        p = _lib.default_params()
        p.dt = float(scene.dt)
        p.gx, p.gy, p.gz = (float(v) for v in scene.gravity)
        p.density = float(scene.density)
        p.stab_iters, p.solver_iters = stab, solver
        p.mu_s, p.mu_k = mu_s, mu_k
        p.skin = skin
        p.flags = flags
        self.params = p
        ms = max_sweeps or (stab + solver)
        nbytes = lib.gg_workspace_bytes(cap, max(nb, 1), self.grid, ms, gmax, bmax)
        if nbytes == 0:
            raise ValueError("invalid workspace request")
        self.ws_buf = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.ws = _lib.gg_workspace(self.ws_buf.data_ptr(), int(nbytes), cap, max(nb, 1),
                                    self.grid, ms, gmax, bmax)
        check("gg_workspace_init", lib.gg_workspace_init(C.byref(self.ws), _stream()))
        self.set = _lib.gg_particles()
        self._refresh_set()
        self.bset = _lib.gg_particles(nb, self.bpos.shape[0], _ptr(self.bpos), _ptr(self.bvel),
                                      _ptr(self.bmass), None, None, None, None, _ptr(self.bp2c),
                                      _ptr(self.banim), None)
        check("gg_boundary_build",
              lib.gg_boundary_build(C.byref(self.bset), _ptr(self.bcs), C.byref(self.ws),
                                    _stream()))

    def _refresh_set(self):
        s = self.set
        s.n, s.capacity = self.n, self.cap
        s.pos, s.vel, s.mass = _ptr(self.pos), _ptr(self.vel), _ptr(self.mass)
        s.scaled_mass = _ptr(self.scaled_mass)
        s.particle_2_cell = _ptr(self.p2c)
        s.is_animated = _ptr(self.is_animated)
        s.adaptive_last_step = _ptr(self.adl)

    def step(self, nsteps=1):
        for _ in range(nsteps):
            check("gg_step", self.lib.gg_step(C.byref(self.set), None, _ptr(self.cs),
                                              C.byref(self.bset), _ptr(self.bcs),
                                              C.byref(self.params), C.byref(self.ws), _stream()))

    def graph(self, steps_per_launch=1) -> "StepGraph":
        """gg_step of this scene captured into a hipGraph (gg_step_graph)."""
        return StepGraph(self, steps_per_launch)

    # stage-level entry points ------------------------------------------------
    def grid_build(self, perm_out: torch.Tensor | None = None):
        check("gg_grid_build", self.lib.gg_grid_build(C.byref(self.set), None, _ptr(self.cs),
                                                      _ptr(perm_out), C.byref(self.ws),
                                                      _stream()))

    def predict(self):
        check("gg_predict", self.lib.gg_predict(C.byref(self.set), _ptr(self.cs),
                                                C.byref(self.bset), _ptr(self.bcs),
                                                C.byref(self.params), C.byref(self.ws),
                                                _stream()))

    def sweep(self, k):
        check("gg_contact_sweep",
              self.lib.gg_contact_sweep(C.byref(self.set), _ptr(self.cs), C.byref(self.bset),
                                        _ptr(self.bcs), C.byref(self.params), k,
                                        C.byref(self.ws), _stream()))

    def finalize(self):
        check("gg_finalize", self.lib.gg_finalize(C.byref(self.set), C.byref(self.params),
                                                  C.byref(self.ws), _stream()))

    def view(self, which: int, count: int, dtype=torch.float32, width=4) -> torch.Tensor:
        """Copy of a workspace buffer (see gg_workspace_view)."""
        ptr = self.lib.gg_workspace_view(C.byref(self.ws), which)
        if not ptr:
            raise RuntimeError("invalid workspace view")
        off = ptr - self.ws_buf.data_ptr()
        esz = torch.empty((), dtype=dtype).element_size()
        raw = self.ws_buf[off: off + count * width * esz]
        return raw.view(dtype).reshape(count, width).clone()

    # host copies -------------------------------------------------------------
    def host(self):
        torch.cuda.synchronize()
        n = self.n
        return (self.pos[:n].cpu().numpy().copy(), self.vel[:n].cpu().numpy().copy(),
                self.mass[:n].cpu().numpy().copy())


class AdaptiveSim(PlainSim):
    """Scene with adaptive merge/split (reference is_adaptive == 1): the plain
    step, then surface flags, distance to the surface and
    Solver::adaptive_sampling (GranularSystem.cu:420-432)."""

    def __init__(self, scene, capacity=None, max_mass=4.0, blend=40, merge_first=False, **kw):
        cap = capacity or 2 * scene.n
        super().__init__(scene, capacity=cap, **kw)
        dev = self.dev
        i32, f32 = torch.int32, torch.float32
        self.remove = torch.zeros(cap, dtype=i32, device=dev)
        self.merge = torch.zeros(cap, dtype=f32, device=dev)
        self.mcount = torch.zeros(cap, dtype=i32, device=dev)
        self.mv = torch.zeros((cap, 3), dtype=f32, device=dev)
        self.surface = torch.zeros(cap, dtype=i32, device=dev)
        self.sd = torch.zeros(cap, dtype=f32, device=dev)
        self.num_surface = torch.zeros(cap, dtype=i32, device=dev)
        self.max_mass = max_mass
        self.blend = blend
        self.merge_first = merge_first
        self.calls = 0
        self.ad = _lib.gg_adaptive(_ptr(self.remove), _ptr(self.merge), _ptr(self.mv),
                                   _ptr(self.mcount))
        self.stats = []
        self._refresh_set()

    def _refresh_set(self):
        super()._refresh_set()
        if hasattr(self, "surface"):
            self.set.surface = _ptr(self.surface)
            self.set.surface_distance = _ptr(self.sd)
            self.set.num_surface = _ptr(self.num_surface)

    def step(self, nsteps=1):
        lib = self.lib
        st = (C.c_int * 3)()
        for _ in range(nsteps):
            s = _stream()
            check("gg_step", lib.gg_step(C.byref(self.set), C.byref(self.ad), _ptr(self.cs),
                                         C.byref(self.bset), _ptr(self.bcs),
                                         C.byref(self.params), C.byref(self.ws), s))
            check("gg_surface", lib.gg_surface(C.byref(self.set), _ptr(self.cs),
                                               C.byref(self.params), C.byref(self.ws), s))
            check("gg_surface_distance", lib.gg_surface_distance(C.byref(self.set), _ptr(self.cs),
                                                                 C.byref(self.ws), s))
            do_merge = 1 if (self.calls > 0 or self.merge_first) else 0
            n_before = self.set.n
            check("gg_adaptive_sampling",
                  lib.gg_adaptive_sampling(C.byref(self.set), C.byref(self.ad), _ptr(self.cs),
                                           C.byref(self.bset), _ptr(self.bcs),
                                           C.byref(self.params), self.max_mass, do_merge,
                                           self.blend, C.byref(self.ws), s, st))
            self.n = self.set.n
            self.calls += 1
            self.stats.append((st[0], st[1], n_before))

    def host_all(self):
        torch.cuda.synchronize()
        n = self.n
        g = lambda t: t[:n].cpu().numpy().copy()
        return dict(pos=g(self.pos), vel=g(self.vel), mass=g(self.mass), adl=g(self.adl),
                    remove=g(self.remove), merge=g(self.merge), mcount=g(self.mcount),
                    surface=g(self.surface), sd=g(self.sd))


class FineSet:
    """Fine (upsampled) particles of one scene: positions, velocities, their
    own cell table and a sort-only workspace (the reference's _upsampled set
    and _cell_start_upsampled, GranularSystem.hpp:60-90)."""

    def __init__(self, sim: PlainSim, pos: np.ndarray, vel: np.ndarray | None = None):
        lib = sim.lib
        self.lib, self.sim = lib, sim
        dev = sim.dev
        n = int(pos.shape[0])
        self.n = n
        self.pos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).to(dev)
        v = np.zeros_like(pos, np.float32) if vel is None else vel
        self.vel = torch.from_numpy(np.ascontiguousarray(v, np.float32)).to(dev)
        self.p2c = torch.zeros(max(n, 1), dtype=torch.int32, device=dev)
        self.cs = torch.zeros(sim.C + 2, dtype=torch.int32, device=dev)
        nbytes = lib.gg_workspace_bytes(max(n, 1), 0, sim.grid, 0, 0, 0)
        if nbytes == 0:
            raise ValueError("invalid fine workspace request")
        self.ws_buf = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.ws = _lib.gg_workspace(self.ws_buf.data_ptr(), int(nbytes), max(n, 1), 0, sim.grid,
                                    0, 0, 0)
        check("gg_workspace_init", lib.gg_workspace_init(C.byref(self.ws), _stream()))
        self.set = _lib.gg_particles(n, max(n, 1), _ptr(self.pos), _ptr(self.vel), None, None,
                                     None, None, None, _ptr(self.p2c), None, None)
        self.up = _lib.gg_upsample_params()
        lib.gg_upsample_params_default(C.byref(self.up))

    def sort(self):
        check("gg_upsampled_sort", self.lib.gg_upsampled_sort(C.byref(self.set), _ptr(self.cs),
                                                              C.byref(self.ws), _stream()))

    def upsample(self, flags=0):
        s = self.sim
        check("gg_upsample_plain",
              self.lib.gg_upsample_plain(C.byref(self.set), C.byref(s.set), _ptr(s.cs),
                                         C.byref(s.bset), _ptr(s.bcs), s.grid, C.byref(self.up),
                                         flags, _stream()))

    def host(self):
        torch.cuda.synchronize()
        return self.pos.cpu().numpy().copy(), self.vel.cpu().numpy().copy()


class StepGraph:
    """A captured fixed-N step of one PlainSim (gg_step_graph_*): each launch
    advances the scene by steps_per_launch steps."""

    def __init__(self, sim: PlainSim, steps_per_launch=1):
        self.lib, self.sim = sim.lib, sim
        h = C.c_void_p()
        check("gg_step_graph_create", self.lib.gg_step_graph_create(
            C.byref(sim.set), _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs),
            C.byref(sim.params), C.byref(sim.ws), steps_per_launch, C.byref(h)))
        self.h = h

    def launch(self, n=1):
        for _ in range(n):
            check("gg_step_graph_launch", self.lib.gg_step_graph_launch(self.h, _stream()))

    def close(self):
        if self.h:
            self.lib.gg_step_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchSim:
    """Independent plain scenes advanced together by one hipGraph launch
    (gg_batch_*).  `sims` are PlainSim objects; their arrays, cell tables
    and workspaces are used in place."""

    def __init__(self, sims, branches=None, sweep_block=0, nlist_block=0, steps_per_launch=1):
        lib = load()
        self.lib, self.sims = lib, list(sims)
        arr = (_lib.gg_batch_scene * len(self.sims))()
        for k, s in enumerate(self.sims):
            arr[k].set = s.set
            arr[k].cell_start = _ptr(s.cs)
            arr[k].boundary = s.bset
            arr[k].bcell_start = _ptr(s.bcs)
            arr[k].params = s.params
            arr[k].ws = s.ws
        cfg = _lib.gg_batch_config(branches or min(4, len(self.sims)), sweep_block, nlist_block,
                                   steps_per_launch)
        h = C.c_void_p()
        check("gg_batch_create", lib.gg_batch_create(arr, len(self.sims), C.byref(cfg),
                                                     C.byref(h)))
        self.h = h
        self._arr = arr

    def step(self, launches=1):
        for _ in range(launches):
            check("gg_batch_step", self.lib.gg_batch_step(self.h, _stream()))

    def step_eager(self, launches=1):
        for _ in range(launches):
            check("gg_batch_step_eager", self.lib.gg_batch_step_eager(self.h, _stream()))

    def config(self):
        c = _lib.gg_batch_config()
        check("gg_batch_get_config", self.lib.gg_batch_get_config(self.h, C.byref(c)))
        return dict(branches=c.branches, sweep_block=c.sweep_block, nlist_block=c.nlist_block,
                    steps_per_launch=c.steps_per_launch)

    def set_config(self, **kw):
        c = self.config()
        c.update(kw)
        cfg = _lib.gg_batch_config(c["branches"], c["sweep_block"], c["nlist_block"],
                                   c["steps_per_launch"])
        check("gg_batch_set_config", self.lib.gg_batch_set_config(self.h, C.byref(cfg)))

    def autotune(self, trials=2):
        best = _lib.gg_batch_config()
        ms = C.c_float()
        check("gg_batch_autotune", self.lib.gg_batch_autotune(self.h, trials, _stream(),
                                                              C.byref(best), C.byref(ms)))
        return dict(branches=best.branches, sweep_block=best.sweep_block,
                    nlist_block=best.nlist_block, steps_per_launch=best.steps_per_launch,
                    ms_per_launch=ms.value)

    def close(self):
        if getattr(self, "h", None):
            self.lib.gg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SceneBatch:
    """Independent plain scenes stored back to back in one set and advanced
    by one launch per stage (gg_step_scenes; captured: gg_scenes_graph_*).
    `scenes` are scenes.Scene objects sharing cell length, dt, gravity and
    density; `friction` is a list of (mu_s, mu_k) per scene.  Scene s's
    particles are the rows first[s]:first[s + 1] of the set (cell-sorted
    within the scene after every step)."""

    def __init__(self, scenes, friction, stab=3, solver=5, skin=0.0075, gmax=48, bmax=16,
                 flags=0, device="cuda"):
        lib = load()
        self.lib = lib
        self.scenes = list(scenes)
        s0 = self.scenes[0]
        for sc in self.scenes:
            if (np.float32(sc.cell_length) != np.float32(s0.cell_length)
                    or np.float32(sc.dt) != np.float32(s0.dt)
                    or tuple(sc.gravity) != tuple(s0.gravity)
                    or np.float32(sc.density) != np.float32(s0.density)):
                raise ValueError("SceneBatch: scenes must share cell length, dt, gravity and "
                                 "density")
        if len(friction) != len(self.scenes):
            raise ValueError("SceneBatch: one (mu_s, mu_k) per scene")
        fr = [(float(a), float(b)) for a, b in friction]
        with_f = [a > 0.0 or b > 0.0 for a, b in fr]
        if any(with_f) and not all(with_f):
            raise ValueError("SceneBatch: friction on every scene or on none")
        dev = torch.device(device)
        self.dev = dev
        ns = [sc.n for sc in self.scenes]
        nbs = [sc.boundary.shape[0] for sc in self.scenes]
        self.first = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
        self.bfirst = np.concatenate([[0], np.cumsum(nbs)]).astype(np.int64)
        n, nb = int(self.first[-1]), int(self.bfirst[-1])
        self.n, self.nb = n, nb
        desc = (_lib.gg_scene_desc * len(self.scenes))()
        base = 0
        for k, sc in enumerate(self.scenes):
            cx, cy, cz = sc.cells
            desc[k] = _lib.gg_scene_desc(cx, cy, cz, base, fr[k][0], fr[k][1])
            base += cx * cy * cz + 1
        self.total_cells = base
        self.host_desc = desc
        raw = np.frombuffer(bytes(desc), dtype=np.uint8)
        self.desc_dev = torch.from_numpy(raw.copy()).to(dev)
        f32, i32 = torch.float32, torch.int32
        cat = lambda xs: np.ascontiguousarray(np.concatenate(xs), np.float32)
        self.pos = torch.from_numpy(cat([sc.pos for sc in self.scenes])).to(dev)
        self.vel = torch.from_numpy(cat([sc.vel for sc in self.scenes])).to(dev)
        self.mass = torch.from_numpy(cat([sc.mass for sc in self.scenes])).to(dev)
        self.scaled_mass = torch.zeros(n, dtype=f32, device=dev)
        self.p2c = torch.zeros(n, dtype=i32, device=dev)
        self.adl = torch.zeros(n, dtype=i32, device=dev)
        self.scene_of = torch.from_numpy(
            np.repeat(np.arange(len(self.scenes), dtype=np.int32), ns)).to(dev)
        self.bpos = torch.from_numpy(cat([sc.boundary for sc in self.scenes])).to(dev)
        self.bvel = torch.zeros_like(self.bpos)
        self.bmass = torch.ones(max(nb, 1), dtype=f32, device=dev)
        self.bp2c = torch.zeros(max(nb, 1), dtype=i32, device=dev)
        self.banim = torch.zeros(max(nb, 1), dtype=i32, device=dev)
        self.bscene_of = torch.from_numpy(
            np.repeat(np.arange(len(self.scenes), dtype=np.int32), nbs)).to(dev)
        self.cs = torch.zeros(base + 1, dtype=i32, device=dev)
        self.bcs = torch.zeros(base + 1, dtype=i32, device=dev)
        self.grid = _lib.gg_grid(float(s0.cell_length), base - 1, 1, 1)
        p = _lib.default_params()
        p.dt = float(s0.dt)
        p.gx, p.gy, p.gz = (float(v) for v in s0.gravity)
        p.density = float(s0.density)
        p.stab_iters, p.solver_iters = stab, solver
        p.skin = skin
        p.flags = flags
        self.params = p
        nbytes = lib.gg_workspace_bytes(max(n, 1), max(nb, 1), self.grid, stab + solver, gmax,
                                        bmax)
        if nbytes == 0:
            raise ValueError("invalid workspace request")
        self.ws_buf = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        self.ws = _lib.gg_workspace(self.ws_buf.data_ptr(), int(nbytes), max(n, 1), max(nb, 1),
                                    self.grid, stab + solver, gmax, bmax)
        check("gg_workspace_init", lib.gg_workspace_init(C.byref(self.ws), _stream()))
        self.set = _lib.gg_particles(n, n, _ptr(self.pos), _ptr(self.vel), _ptr(self.mass),
                                     _ptr(self.scaled_mass), None, None, None, _ptr(self.p2c),
                                     None, _ptr(self.adl))
        self.bset = _lib.gg_particles(nb, max(nb, 1), _ptr(self.bpos), _ptr(self.bvel),
                                      _ptr(self.bmass), None, None, None, None, _ptr(self.bp2c),
                                      _ptr(self.banim), None)
        self.sc = _lib.gg_scenes(len(self.scenes), self.desc_dev.data_ptr(), C.addressof(desc),
                                 self.scene_of.data_ptr(), self.bscene_of.data_ptr(),
                                 1 if any(with_f) else 0)
        check("gg_scenes_boundary_build",
              lib.gg_scenes_boundary_build(C.byref(self.bset), _ptr(self.bcs), C.byref(self.sc),
                                           C.byref(self.ws), _stream()))
        self.graph = None

    def step(self, nsteps=1):
        for _ in range(nsteps):
            check("gg_step_scenes",
                  self.lib.gg_step_scenes(C.byref(self.set), _ptr(self.cs), C.byref(self.bset),
                                          _ptr(self.bcs), C.byref(self.sc),
                                          C.byref(self.params), C.byref(self.ws), _stream()))

    def capture(self, steps_per_launch=1):
        """The step captured into a hipGraph (gg_scenes_graph_create)."""
        self.close_graph()
        h = C.c_void_p()
        check("gg_scenes_graph_create", self.lib.gg_scenes_graph_create(
            C.byref(self.set), _ptr(self.cs), C.byref(self.bset), _ptr(self.bcs),
            C.byref(self.sc), C.byref(self.params), C.byref(self.ws), steps_per_launch,
            C.byref(h)))
        self.graph = h

    def launch(self, launches=1):
        for _ in range(launches):
            check("gg_scenes_graph_launch",
                  self.lib.gg_scenes_graph_launch(self.graph, _stream()))

    def autotune(self, trials=2):
        sw, nl, ms = C.c_int(), C.c_int(), C.c_float()
        check("gg_scenes_graph_autotune", self.lib.gg_scenes_graph_autotune(
            self.graph, trials, _stream(), C.byref(sw), C.byref(nl), C.byref(ms)))
        return dict(sweep_block=sw.value, nlist_block=nl.value, ms_per_launch=ms.value)

    def scene(self, k):
        """Host copies (pos, vel, mass) of scene k's particles."""
        torch.cuda.synchronize()
        a, b = int(self.first[k]), int(self.first[k + 1])
        return (self.pos[a:b].cpu().numpy().copy(), self.vel[a:b].cpu().numpy().copy(),
                self.mass[a:b].cpu().numpy().copy())

    def close_graph(self):
        if getattr(self, "graph", None):
            self.lib.gg_scenes_graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close_graph()
        except Exception:
            pass
