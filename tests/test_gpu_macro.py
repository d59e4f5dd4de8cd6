"""Macroscopic statistics of long runs: the GPU's FP32 run against the
independent FP64 oracle (its own decisions, no shadowing), each within 2%:

- the centre-of-mass trajectory of the 32,768-particle plain box (r = 0.01,
  dt = 1/240, 3 + 10 sweeps, friction 0.5 / 0.4) over 500 steps;
- the outflow rate through the funnel: 16,000 particles of r = 0.005 poured
  into a 30-degree cone of 12,000 boundary particles (outlet 0.1), the rate
  being the slope of the count below the outlet over steps 300..800;
- the angle of repose of a fixed-base heap: a 28 x 50 x 28 column of
  r = 0.005 particles collapsing on a disk of radius 0.2 whose rim lets the
  excess fall off; the angle from the volume the disk retains,
  tan(theta) = 3 V / (pi R^3), a cone on the disk's base, with V the
  retained particles' volume over a packing fraction of 0.6.

Trajectories of single particles part ways within a few hundred steps
(FP32 and FP64 round differently and granular contacts amplify it); these
statistics average over thousands of particles.  The slope of the heap's
surface is noisier at this size (the surface is a few hundred particles
and keeps creeping); it is logged beside the volume angle and held to
10%."""
import math
import sys

import numpy as np
import pytest

from granular_gpu_amd import scenes
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gpu(scene, mu_s, mu_k, solver=10):
    from granular_gpu_amd.device import PlainSim
    return PlainSim(scene, stab=3, solver=solver, mu_s=mu_s, mu_k=mu_k)


class _F64:
    """The FP64 oracle run of a scene, advanced in chunks."""

    def __init__(self, scene, mu_s, mu_k, solver=10):
        self.p = po.params_from(scene, 3, solver, mu_s, mu_k)
        self.bpos, self.bcs = po.sorted_boundary(scene.boundary, scene.cell_length, scene.cells)
        self.pos = scene.pos.astype(np.float64)
        self.vel = scene.vel.astype(np.float64)
        self.mass = scene.mass.astype(np.float64)

    def step(self, k):
        self.pos, self.vel, self.mass, _ = po.plain_steps_f64(
            self.pos, self.vel, self.mass, self.bpos, self.bcs, self.p, k)
        return self.pos


def test_com_trajectory_plain_box_500_steps(gpu):
    s = scenes.plain_box(32, 1.0, dt=1.0 / 240.0, g=-9.81)
    sim = _gpu(s, 0.5, 0.4)
    ref = _F64(s, 0.5, 0.4)
    c0 = s.pos.astype(np.float64).mean(0)
    c32, c64 = [], []
    for k in range(10):
        sim.step(50)
        c32.append(sim.host()[0].astype(np.float64).mean(0))
        c64.append(ref.step(50).mean(0))
        _log(f"step {50 * (k + 1)}: com y fp32 {c32[-1][1]:.6f} fp64 {c64[-1][1]:.6f}")
    c32, c64 = np.array(c32), np.array(c64)
    drop = np.abs(c64 - c0).max()
    dev = np.abs(c32 - c64).max()
    _log(f"com: largest difference {dev:.3e} m against a displacement of {drop:.4f} m "
         f"({100 * dev / drop:.2f}%)")
    assert drop > 0.05
    assert dev <= 0.02 * drop


def test_funnel_outflow_rate(gpu):
    s = scenes.funnel_pour(n=16000, cone_count=12000, r=0.005, cell_length=0.0202)
    outlet = 0.3 * 2.0
    sim = _gpu(s, 0.5, 0.4)
    ref = _F64(s, 0.5, 0.4)
    sim.step(300)
    ref.step(300)
    steps, o32, o64 = [], [], []
    for k in range(11):
        if k:
            sim.step(50)
            ref.step(50)
        steps.append(300 + 50 * k)
        o32.append(int((sim.host()[0][:, 1] < outlet).sum()))
        o64.append(int((ref.pos[:, 1] < outlet).sum()))
    r32 = np.polyfit(steps, o32, 1)[0]
    r64 = np.polyfit(steps, o64, 1)[0]
    _log(f"outflow below the outlet: fp32 {o32} fp64 {o64}")
    _log(f"outflow rate steps 300..800: fp32 {r32:.4f} fp64 {r64:.4f} particles/step "
         f"({100 * abs(r32 - r64) / r64:.2f}%)")
    assert r64 > 1.0
    assert abs(r32 - r64) <= 0.02 * r64


def _volume_angle(pos, c, R, r, base_y=0.11, phi=0.6):
    on = (np.hypot(pos[:, 0] - c, pos[:, 2] - c) < R) & (pos[:, 1] > base_y - 0.01)
    v = on.sum() * (4.0 / 3.0) * math.pi * r ** 3 / phi
    return int(on.sum()), math.degrees(math.atan(3.0 * v / (math.pi * R ** 3)))


def test_fixed_base_heap_angle_of_repose(gpu):
    R, box, r = 0.2, 0.8, 0.005
    s = scenes.disk_heap(28, 50, r=r, disk_radius=R, box=box)
    c = 0.5 * box
    sim = _gpu(s, 0.5, 0.4)
    ref = _F64(s, 0.5, 0.4)
    for k in range(8):
        sim.step(100)
        ref.step(100)
        p32 = sim.host()[0]
        n32, a32 = _volume_angle(p32, c, R, r)
        n64, a64 = _volume_angle(ref.pos, c, R, r)
        s32 = scenes.heap_angle(p32, (c, c), 0.11, R)
        s64 = scenes.heap_angle(ref.pos, (c, c), 0.11, R)
        _log(f"step {100 * (k + 1)}: on the disk fp32 {n32} fp64 {n64}; volume angle "
             f"{a32:.3f} / {a64:.3f} deg; surface slope {s32:.2f} / {s64:.2f} deg")
    assert n64 > 2000
    assert abs(a32 - a64) <= 0.02 * a64
    assert abs(s32 - s64) <= 0.10 * s64
