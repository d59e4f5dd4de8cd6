"""Learned upsampler: device weights, data from plain coarse runs, training
on the GPU through the library's MFMA kernels, and inference.

Every computation runs in libgranular_gpu.so (gg_mlp_*, gg_upsample_mlp);
torch tensors here are device allocations and index plumbing (batch
gathers).  The definitions (features, teacher, network, mixed precision)
are stated in include/granular_gpu.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import check, load
from .device import _ptr, _stream

DIMS = ((96, 128), (128, 128), (128, 128), (128, 32))
WOFF = (0, 12288, 28672, 45056)
BOFF = (0, 128, 256, 384)
NW, NB = 49152, 416


def default_config(radius=0.02) -> _lib.gg_mlp_config:
    c = _lib.gg_mlp_config()
    load().gg_mlp_config_default(C.byref(c))
    c.radius = radius
    return c


def default_upsample_params() -> _lib.gg_upsample_params:
    u = _lib.gg_upsample_params()
    load().gg_upsample_params_default(C.byref(u))
    return u


def init_master(seed=0) -> tuple[np.ndarray, np.ndarray]:
    """He-uniform FP32 master weights (rows 24..31 of the last layer zero),
    zero biases."""
    rng = np.random.default_rng(seed)
    W = np.zeros(NW, np.float32)
    for l, (k_in, n_out) in enumerate(DIMS):
        a = np.sqrt(6.0 / k_in)
        w = rng.uniform(-a, a, (n_out, k_in)).astype(np.float32)
        if l == 3:
            w[24:] = 0.0
            w *= 0.1
        W[WOFF[l]:WOFF[l] + n_out * k_in] = w.reshape(-1)
    return W, np.zeros(NB, np.float32)


class MlpNet:
    """Device weights: bf16 copy and biases for inference, FP32 masters and
    Adam moments for training."""

    def __init__(self, seed=0, device="cuda", master=None):
        self.lib = load()
        W, b = master if master is not None else init_master(seed)
        f32 = dict(dtype=torch.float32, device=device)
        self.w_master = torch.from_numpy(np.ascontiguousarray(W)).to(device)
        self.b_master = torch.from_numpy(np.ascontiguousarray(b)).to(device)
        self.m_w, self.v_w = torch.zeros(NW, **f32), torch.zeros(NW, **f32)
        self.m_b, self.v_b = torch.zeros(NB, **f32), torch.zeros(NB, **f32)
        self.w_bf16 = torch.zeros(NW, dtype=torch.int16, device=device)
        self.bias = torch.zeros(NB, **f32)
        self.step = 0
        self.struct = _lib.gg_mlp_weights(*[_ptr(t).value for t in (
            self.w_bf16, self.bias, self.w_master, self.m_w, self.v_w, self.b_master,
            self.m_b, self.v_b)])
        check("gg_mlp_sync_weights", self.lib.gg_mlp_sync_weights(C.byref(self.struct),
                                                                  _stream()))

    def host_bf16(self) -> tuple[np.ndarray, np.ndarray]:
        """The bfloat16 weights (as float32 values) and biases."""
        torch.cuda.synchronize()
        w = self.w_bf16.cpu().numpy().astype(np.uint16).astype(np.uint32) << 16
        return w.view(np.float32).copy(), self.bias.cpu().numpy().copy()


class Trainer:
    """Minibatch Adam on (features, targets) held on the device."""

    def __init__(self, net: MlpNet, batch=512):
        self.net, self.lib, self.B = net, net.lib, batch
        dev = net.w_master.device
        nbytes = int(self.lib.gg_mlp_train_scratch_bytes(batch))
        self.scratch = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.dW = torch.zeros(NW, dtype=torch.float32, device=dev)
        self.db = torch.zeros(NB, dtype=torch.float32, device=dev)
        self.sq = torch.zeros(batch, dtype=torch.float32, device=dev)

    def grad(self, x: torch.Tensor, t: torch.Tensor):
        B = x.shape[0]
        assert B <= self.B and x.dtype == torch.int16 and x.shape[1] == 96
        assert t.dtype == torch.float32 and t.shape == (B, 24)
        check("gg_mlp_grad", self.lib.gg_mlp_grad(
            C.byref(self.net.struct), _ptr(x), _ptr(t), B, _ptr(self.dW), _ptr(self.db),
            _ptr(self.sq), _ptr(self.scratch), self.scratch.numel(), _stream()))
        return self.dW, self.db, self.sq[:B]

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """The training forward over x, in chunks of the trainer's batch."""
        n = x.shape[0]
        y = torch.empty((n, 24), dtype=torch.float32, device=x.device)
        for a in range(0, n, self.B):
            b = min(n, a + self.B)
            check("gg_mlp_forward", self.lib.gg_mlp_forward(
                C.byref(self.net.struct), _ptr(x[a:b]), b - a, _ptr(y[a:b]),
                _ptr(self.scratch), self.scratch.numel(), _stream()))
        return y

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.net.step += 1
        check("gg_mlp_adam", self.lib.gg_mlp_adam(C.byref(self.net.struct), _ptr(self.dW),
                                                  _ptr(self.db), lr, beta1, beta2, eps,
                                                  self.net.step, _stream()))

    def fit(self, x: torch.Tensor, t: torch.Tensor, steps=1000, lr=1e-3, seed=0,
            log_every=0, lr_final=None):
        """Adam over random minibatches (seeded); learning rate decays
        geometrically to lr_final.  Returns the per-step mean losses."""
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = x.shape[0]
        losses = torch.zeros(steps, dtype=torch.float32, device=x.device)
        lr_final = lr if lr_final is None else lr_final
        for k in range(steps):
            idx = torch.randint(0, n, (self.B,), generator=g).to(x.device)
            xb, tb = x[idx].contiguous(), t[idx].contiguous()
            _, _, sq = self.grad(xb, tb)
            losses[k] = sq.sum() / (self.B * 24)
            self.adam(lr * (lr_final / lr) ** (k / max(steps - 1, 1)))
            if log_every and (k + 1) % log_every == 0:
                print(f"# mlp step {k + 1}: loss {losses[k].item():.3e}", flush=True)
        return losses.cpu().numpy()


class LearnedUpsampler:
    """Features, teacher and inference for the coarse set of a PlainSim."""

    def __init__(self, net: MlpNet, cfg=None, up=None):
        self.net, self.lib = net, net.lib
        self.cfg = cfg or default_config()
        self.up = up or default_upsample_params()
        self._scratch = None

    def features(self, sim):
        n = sim.n
        dev = sim.pos.device
        feat = torch.empty((n, 96), dtype=torch.int16, device=dev)
        flags = torch.empty(n, dtype=torch.int32, device=dev)
        knn = torch.empty((n, 16), dtype=torch.int32, device=dev)
        check("gg_mlp_features", self.lib.gg_mlp_features(
            C.byref(sim.set), _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs), sim.grid,
            C.byref(self.cfg), _ptr(feat), _ptr(flags), _ptr(knn), _stream()))
        return feat, flags, knn

    def teacher(self, sim, flags=None):
        n = sim.n
        dev = sim.pos.device
        y = torch.empty((n, 24), dtype=torch.float32, device=dev)
        fp = torch.empty((n, 24), dtype=torch.float32, device=dev)
        fv = torch.empty((n, 24), dtype=torch.float32, device=dev)
        check("gg_mlp_teacher", self.lib.gg_mlp_teacher(
            C.byref(sim.set), _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs), sim.grid,
            C.byref(self.up), C.byref(self.cfg), _ptr(flags), _ptr(y), _ptr(fp), _ptr(fv),
            _stream()))
        return y, fp, fv

    def infer(self, sim, fine_pos=None, fine_vel=None, fallback_count=None):
        n = sim.n
        dev = sim.pos.device
        need = int(self.lib.gg_upsample_mlp_scratch_bytes(n))
        if self._scratch is None or self._scratch.numel() < need:
            self._scratch = torch.empty(need, dtype=torch.uint8, device=dev)
        if fine_pos is None:
            fine_pos = torch.empty((n, 24), dtype=torch.float32, device=dev)
        check("gg_upsample_mlp", self.lib.gg_upsample_mlp(
            C.byref(sim.set), _ptr(sim.cs), C.byref(sim.bset), _ptr(sim.bcs), sim.grid,
            C.byref(self.up), C.byref(self.cfg), _ptr(self.net.w_bf16), _ptr(self.net.bias),
            _ptr(fine_pos), _ptr(fine_vel), _ptr(self._scratch), self._scratch.numel(),
            _ptr(fallback_count), _stream()))
        return fine_pos

    def rigid(self, sim):
        """Fine positions with y = 0: the template carried rigidly with the
        parent (the method without the network)."""
        T = torch.tensor([[(k >> 2) & 1, (k >> 1) & 1, k & 1] for k in range(8)],
                         dtype=torch.float32, device=sim.pos.device) - 0.5
        x = sim.pos[:sim.n][:, None, :] + self.cfg.radius * T[None]
        return (x + self.up.dt * sim.vel[:sim.n][:, None, :]).reshape(sim.n, 24)

    def collect(self, sim, snapshots, every, warmup=0):
        """Training pairs from a plain coarse run: after `warmup` steps, every
        `every` steps, the features and teacher targets of the particles
        that are not flagged for fallback."""
        xs, ts = [], []
        sim.step(warmup)
        for _ in range(snapshots):
            sim.step(every)
            sim.grid_build()  # cell table of the current positions
            feat, flags, _ = self.features(sim)
            y, _, _ = self.teacher(sim)
            keep = flags == 0
            xs.append(feat[keep])
            ts.append(y[keep])
        return torch.cat(xs), torch.cat(ts)


def fine_targets(coarse_sim, fine_sim, cfg=None):
    """Training targets of every coarse particle from the fine run
    (gg_mlp_fine_targets): y (n x 24), d2 (n x 8), match (n x 8).  Both
    sims must have current cell tables (grid_build after their last step)."""
    cfg = cfg or default_config()
    n = coarse_sim.n
    dev = coarse_sim.pos.device
    y = torch.empty((n, 24), dtype=torch.float32, device=dev)
    d2 = torch.empty((n, 8), dtype=torch.float32, device=dev)
    jm = torch.empty((n, 8), dtype=torch.int32, device=dev)
    check("gg_mlp_fine_targets", load().gg_mlp_fine_targets(
        C.byref(coarse_sim.set), C.byref(fine_sim.set), _ptr(fine_sim.cs), fine_sim.grid,
        C.byref(cfg), _ptr(y), _ptr(d2), _ptr(jm), _stream()))
    return y, d2, jm


def collect_pairs(up: "LearnedUpsampler", coarse_sim, fine_sim, snapshots, every, warmup=0,
                  max_match=1.0):
    """Training pairs from a coarse run and its fine twin stepped side by
    side (scenes.paired_boxes): after `warmup` steps, every `every` steps,
    the coarse features and the fine-run targets of the coarse particles
    that are not flagged for fallback and whose 8 matched fine particles all
    lie within max_match coarse radii of their template points (where the
    fine run has not drifted away from the coarse one).  Returns (x, t)."""
    xs, ts = [], []
    lim = (max_match * up.cfg.radius) ** 2
    for _ in range(warmup):
        coarse_sim.step()
        fine_sim.step()
    for _ in range(snapshots):
        coarse_sim.step(every)
        fine_sim.step(every)
        coarse_sim.grid_build()
        fine_sim.grid_build()
        feat, flags, _ = up.features(coarse_sim)
        y, d2, _ = fine_targets(coarse_sim, fine_sim, up.cfg)
        keep = (flags == 0) & (d2 < lim).all(dim=1)
        xs.append(feat[keep])
        ts.append(y[keep])
    return torch.cat(xs), torch.cat(ts)


def fine_error(a: torch.Tensor, b: torch.Tensor) -> float:
    """RMS distance between two fine-particle sets slot by slot (n x 24)."""
    d = (a - b).reshape(-1, 3)
    return float(torch.sqrt((d * d).sum(dim=1).mean()).item())
