"""gg_step_noalloc (the plain step with its workspace in the library's
static device array) against gg_step (caller-owned workspace) through the
C-ABI: two calls in a row, the second on a smaller set, each bit for bit;
a request beyond the array's size is refused before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from granular_gpu_amd import scenes

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc_step(sim, steps):
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import _ptr, _stream
    for _ in range(steps):
        _lib.check("gg_step_noalloc", sim.lib.gg_step_noalloc(
            C.byref(sim.set), None, _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs), sim.grid,
            C.byref(sim.params), _stream()))


def test_noalloc_step_equals_step(gpu):
    from granular_gpu_amd.device import PlainSim
    # first call: the 32,768-particle box; second: a smaller box, whose
    # workspace layout differs, so the array holds the first call's data
    for side, steps in ((32, 6), (14, 4)):
        s = scenes.plain_box(side, 1.0)
        _log(f"noalloc vs gg_step: n={s.n}, {steps} steps")
        a = PlainSim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
        b = PlainSim(s, stab=3, solver=10, mu_s=0.5, mu_k=0.4)
        a.step(steps)
        _noalloc_step(b, steps)
        torch.cuda.synchronize()
        for x, y in zip(a.host(), b.host()):
            assert np.array_equal(x, y)
        assert np.array_equal(a.cs.cpu().numpy(), b.cs.cpu().numpy())
        assert np.array_equal(a.p2c[:s.n].cpu().numpy(), b.p2c[:s.n].cpu().numpy())


def test_noalloc_refuses_oversize_request(gpu):
    from granular_gpu_amd import _lib
    from granular_gpu_amd.device import PlainSim, _ptr, _stream
    s = scenes.plain_box(44, 1.2)  # 85,184 particles: workspace > 32 MiB
    sim = PlainSim(s, stab=3, solver=10)
    need = sim.lib.gg_workspace_bytes(s.n, s.boundary.shape[0], sim.grid, 13, 48, 16)
    assert need > (32 << 20)
    before = sim.pos[:s.n].clone()
    rc = sim.lib.gg_step_noalloc(C.byref(sim.set), None, _ptr(sim.cs), C.byref(sim.bset),
                                 _ptr(sim.bcs), sim.grid, C.byref(sim.params), _stream())
    assert rc == _lib.GG_ERR_CAPACITY
    torch.cuda.synchronize()
    assert torch.equal(before, sim.pos[:s.n])
