#!/usr/bin/env python3
"""Benchmark of the plain granular PBD step on one MI355X.

Workload: 1,000,000 particles (r = 0.01) on a 100^3 jittered lattice (seed 1)
in a 2.4^3 m box (60^3 cells of L = 0.0404, 84,968 wall particles), dt =
0.002 s and g = -9.8 (the reference scenes' values, scenes/*.json), 3
stabilisation + 10 solver sweeps, friction 0.5/0.4, FP32.  (With dt = 1/240
the reference's constraint averaging cannot hold a 100-layer column: most
particles are pushed through the floor lattice and leave the grid.)  A step = hash + stable sort + cell table + permute/predict +
neighbour lists + 13 contact sweeps + finalize, inputs resident in HBM.

Before any warm-up, the scene is advanced --settle steps (default 100)
outside every timed region: the lattice starts 0.1 m above the floor and
its bottom layers land around step 71, so the timed window (steps
settle+warmup .. settle+warmup+steps) measures the landed column (lower
layers on the floor, the upper ones still falling onto them: median speed
~1.8 m/s at steps 100..125), not the untouched free fall.

Prints one JSON line.  `value` = particle-steps/s over the timed steps.
Per-stage times come from hipEvents on the launch stream around every stage
of every 10th timed step (--stage-every); the other steps carry no events.
`roofline` covers the dominant kernel (the solver contact sweep) from those
events; each sweep is followed by a short pass over the particles it queued
for whole-wave stencil scans, timed as its own stage (`sweep_queue`).  The
sweep's algorithmic bytes per launch are
N*32 + N_b*12 (read x*, m, m_s and write x* per particle; boundary positions
once).  `cpu_baseline` runs the C++/OpenMP FP32 oracle on a bounded sample
of the same state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s at 1M particles + % HBM roofline; upsampled fine-particles/s"


def bytes_per_particle_step(n, cells, nb, S, K):
    """Algorithmic bytes of an ideally fused step (compulsory traffic)."""
    return 208 + 8 * cells / n + K * (32 + 12 * nb / n) + 24 * S


def upsample_phase(args, lib, torch, C, np):
    """Plain-upsampler throughput: 250k coarse particles (r = 0.02, 63^3
    lattice in a 3 m box) with 8 fine particles each (2,000,376), fine
    offsets U(+-0.02).  Each timed step = coarse gg_step + fine cell sort +
    upsampler pass; fine-particles/s counts the fine work (sort + pass) from
    the stage timers.  Upsampler algorithmic bytes per launch:
    48 N_f (read+write fine pos, vel) + 24 N (coarse pos, vel) + 12 N_b +
    8 C (two cell tables).

    The learned upsampler runs on the same states: trained on the GPU on 50k
    pairs from 3 snapshots after the warm-up, then timed beside the plain
    one (features + network + teacher fallback on flagged particles) and
    scored against the teacher at the end.  Its network roofline counts
    2 (96 128 + 2 128 128 + 128 24) = 96,256 flops per coarse particle
    against the 2.5 PFLOP/s dense bf16 peak.  With --scale-side > 0 the two
    upsamplers are also compared at 8M fine particles (scale_phase)."""
    from granular_gpu_amd import _lib, scenes
    from granular_gpu_amd.device import FineSet, PlainSim

    side, box, per = args.upsample_side, args.upsample_box, 8
    scene = scenes.plain_box(side, box, dt=args.dt, g=-9.8, r=0.02)
    n, nb = scene.n, scene.boundary.shape[0]
    cells = scene.cells[0] * scene.cells[1] * scene.cells[2]
    sim = PlainSim(scene, stab=3, solver=5)
    fine = scenes.fine_cloud(scene.pos, per, 0.02, seed=7)
    fs = FineSet(sim, fine)
    nf = fs.n
    print(f"# upsample scene: coarse={n} fine={nf} walls={nb} cells={cells}", file=sys.stderr,
          flush=True)

    learned = None
    if not args.no_learned:
        from granular_gpu_amd.learned import LearnedUpsampler, MlpNet, Trainer
        learned = LearnedUpsampler(MlpNet(seed=6))

    def run(k, with_learned=False):
        for _ in range(k):
            sim.step()
            fs.sort()
            fs.upsample()
            if with_learned:
                learned.infer(sim, fine_pos=lpos)

    run(args.upsample_warmup)
    train = None
    fine_truth = None
    if learned is not None:
        # training pairs: the coarse scene and its fine twin (r = 0.01, 8 per
        # coarse on the template points, scenes.paired_boxes) run side by
        # side from step 0; 3 snapshots 10 steps apart after the warm-up,
        # 50k pairs drawn from them; then scored on the pair 10 steps later
        from granular_gpu_amd.learned import collect_pairs, fine_targets
        pc, pf = scenes.paired_boxes(side, box, r=0.02, seed=1, dt=args.dt, g=-9.8)
        csim_t, fsim_t = PlainSim(pc, stab=3, solver=5), PlainSim(pf, stab=3, solver=5)
        t0 = time.perf_counter()
        xs, ts = collect_pairs(learned, csim_t, fsim_t, snapshots=3, every=10,
                               warmup=args.upsample_warmup)
        torch.cuda.synchronize()
        t_pairs = time.perf_counter() - t0
        g = torch.Generator(device="cpu").manual_seed(1)
        pick = torch.randperm(xs.shape[0], generator=g)[:50000].to(xs.device)
        kept = int(xs.shape[0])
        xs, ts = xs[pick].contiguous(), ts[pick].contiguous()
        tr = Trainer(learned.net, batch=1024)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        losses = tr.fit(xs, ts, steps=args.learned_train_steps, lr=2e-3, lr_final=1e-4, seed=0)
        torch.cuda.synchronize()
        train = {"data": f"coarse {pc.n} (r=0.02) + fine twin {pf.n} (r=0.01) plain runs, "
                         f"steps {args.upsample_warmup + 10}, {args.upsample_warmup + 20}, "
                         f"{args.upsample_warmup + 30}; targets = velocity of the fine "
                         f"particle nearest each template point minus the parent's",
                 "pairs_available": kept, "pair_generation_s": round(t_pairs, 3),
                 "pairs": int(xs.shape[0]), "steps": args.learned_train_steps, "batch": 1024,
                 "seconds": round(time.perf_counter() - t0, 3),
                 "loss_first": float(losses[:20].mean()), "loss_last": float(losses[-50:].mean())}
        # against the fine-run truth 10 steps after the last snapshot
        csim_t.step(10)
        fsim_t.step(10)
        csim_t.grid_build()
        fsim_t.grid_build()
        feat, flags, _ = learned.features(csim_t)
        truth, d2, _ = fine_targets(csim_t, fsim_t, learned.cfg)
        keep = (flags == 0) & (d2 < (0.02 ** 2)).all(dim=1)
        yt = truth[keep]
        y_net = tr.forward(feat[keep].contiguous())
        y_tch = learned.teacher(csim_t)[0][keep]

        def rms(a):
            return round(float(torch.sqrt(((a - yt) ** 2).mean()).item()), 5)
        fine_truth = {"step": args.upsample_warmup + 40, "coarse_scored": int(keep.sum().item()),
                      "velocity_rms_m_per_s": {"learned": rms(y_net),
                                               "rigid": rms(torch.zeros_like(yt)),
                                               "teacher": rms(y_tch)}}
        del csim_t, fsim_t, feat, flags, truth, d2, y_net, y_tch
        torch.cuda.empty_cache()
        lpos = torch.empty((sim.n, 24), dtype=torch.float32, device=sim.pos.device)
    torch.cuda.synchronize()
    lib.gg_profiler_enable(args.upsample_steps * 32 + 16)
    t0 = time.perf_counter()
    run(args.upsample_steps, with_learned=learned is not None)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    tot = (C.c_double * len(_lib.STAGES))()
    cnt = (C.c_int * len(_lib.STAGES))()
    _lib.check("gg_profiler_collect", lib.gg_profiler_collect(tot, cnt, len(_lib.STAGES)))
    lib.gg_profiler_enable(0)
    st = {}
    for k, name in enumerate(_lib.STAGES):
        if cnt[k]:
            st[name] = {"total_ms": round(tot[k], 4), "launches": cnt[k],
                        "avg_us": round(1000.0 * tot[k] / cnt[k], 3)}
    fp, fv = fs.host()
    assert np.isfinite(fp).all() and np.isfinite(fv).all(), "non-finite fine state"
    up_us, sort_us = st["upsample"]["avg_us"], st["upsample_sort"]["avg_us"]
    per_launch = 48 * nf + 24 * n + 12 * nb + 8 * cells
    ach = per_launch / (up_us * 1e-6) / 1e9
    pmc_up = {}
    pmc_path = ROOT / "profiles" / "pmc_c3_upsample_learned.json"
    if pmc_path.exists():
        try:
            pmc_up = json.loads(pmc_path.read_text()).get("kernels", {})
        except Exception:
            pmc_up = {}

    def traffic_of(kernel):
        # by name prefix: template arguments follow the launch configuration
        for k, v in pmc_up.items():
            if k.startswith(kernel):
                return v.get("hbm_bytes_per_launch")
        return None
    print(f"# upsample timed {args.upsample_steps} steps in {el:.3f} s", file=sys.stderr,
          flush=True)
    learned_out = None
    if learned is not None:
        from granular_gpu_amd.learned import fine_error
        _, tfp, _ = learned.teacher(sim)
        r_fine = 0.01
        e_l = fine_error(lpos, tfp) / r_fine
        e_r = fine_error(learned.rigid(sim), tfp) / r_fine
        fl = learned.features(sim)[1]
        nflag = int((fl != 0).sum().item())
        # the fallback overlaps the network (second stream): time whole calls
        # between events on the calling stream, on the final coarse state
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        learned.infer(sim, fine_pos=lpos)
        e0.record()
        for _ in range(args.upsample_steps):
            learned.infer(sim, fine_pos=lpos)
        e1.record()
        torch.cuda.synchronize()
        t_all = 1000.0 * e0.elapsed_time(e1) / args.upsample_steps
        # the network's stage bracket over back-to-back calls (brackets that
        # span the fork to the side stream also time the other stream's work,
        # so the features and fallback kernels are timed by the kernel trace)
        lib.gg_profiler_enable(args.upsample_steps * 8 + 16)
        for _ in range(args.upsample_steps):
            learned.infer(sim, fine_pos=lpos)
        torch.cuda.synchronize()
        # gg_profiler_collect adds to the arrays it is given: fresh ones
        ltot = (C.c_double * len(_lib.STAGES))()
        lcnt = (C.c_int * len(_lib.STAGES))()
        _lib.check("gg_profiler_collect", lib.gg_profiler_collect(ltot, lcnt, len(_lib.STAGES)))
        lib.gg_profiler_enable(0)
        k_mlp = _lib.STAGES.index("mlp")
        t_net = round(1000.0 * ltot[k_mlp] / lcnt[k_mlp], 3)
        per_call = lcnt[k_mlp] / args.upsample_steps  # network launches per call
        flops = 96256 * n / per_call
        ach_l = flops / (t_net * 1e-6) / 1e12
        learned_out = {
            "network": "96-128-128-128-24 ReLU, bf16 weights/activations, fp32 accumulate",
            "training": train,
            "vs_fine_run_truth": fine_truth,
            "fine_particles_per_s": round(nf / (t_all * 1e-6), 1),
            "speedup_vs_plain": round((up_us + sort_us) / t_all, 3),
            "us_per_call": round(t_all, 3),
            "network_us_per_launch": t_net, "network_launches_per_call": per_call,
            "fallback_coarse": nflag,
            "error_vs_teacher_r_fine": round(e_l, 5),
            "rigid_template_error_r_fine": round(e_r, 5),
            "roofline": {"kernel": "k_mlp_infer (fused MFMA network)", "bound": "mfma",
                         "achieved": round(ach_l, 2), "peak": 2500.0, "unit": "TFLOP/s",
                         "frac": round(ach_l / 2500.0, 4),
                         "traffic": traffic_of("void gg::k_mlp_infer<16, 1>"),
                         "algorithmic_flops_per_launch": int(flops), "avg_launch_us": t_net},
        }
        print(f"# learned: {1e-9 * nf / (t_all * 1e-6):.2f} G fine/s, error {e_l:.4f} r_fine "
              f"(rigid {e_r:.4f})", file=sys.stderr, flush=True)
        if args.scale_side > 0:
            learned_out["at_8m_fine"] = scale_phase(args, lib, torch, C, np, learned)
    return {
        "workload": f"{side}^3 coarse lattice r=0.02 (m=8) in a {box}^3 m box, {per} fine "
                    f"per coarse (offsets U(+-0.02), seed 7), {scene.cells[0]}^3 cells, "
                    f"{nb} walls, coarse 3+5 sweeps, steps {args.upsample_warmup}.."
                    f"{args.upsample_warmup + args.upsample_steps}",
        "coarse": n, "fine": nf, "steps": args.upsample_steps,
        "fine_particles_per_s": round(nf / ((up_us + sort_us) * 1e-6), 1),
        "fine_particles_per_s_pass_only": round(nf / (up_us * 1e-6), 1),
        "coarse_plus_fine_ms_per_step": round(1000.0 * el / args.upsample_steps, 4),
        "stages": {k: v for k, v in st.items() if not k.startswith("mlp")},
        "roofline": {"kernel": "k_upsample (plain upsampler pass)", "bound": "hbm",
                     "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic_of("gg::k_upsample"),
                     "algorithmic_bytes_per_launch": per_launch, "avg_launch_us": up_us},
        "learned": learned_out,
    }


def scale_phase(args, lib, torch, C, np, learned):
    """The plain and the learned upsampler at 1M coarse particles (100^3
    lattice, r = 0.02, 8 fine each: 8M fine), with the network trained on
    the 250k scene.  Plain: fine sort + pass from the stage timers over the
    timed steps; learned: whole calls between events on the calling stream."""
    from granular_gpu_amd import _lib, scenes
    from granular_gpu_amd.device import FineSet, PlainSim
    from granular_gpu_amd.learned import fine_error

    side, per = args.scale_side, 8
    box = float(np.float32(0.04 * side + 0.52))
    scene = scenes.plain_box(side, box, dt=args.dt, g=-9.8, r=0.02)
    sim = PlainSim(scene, stab=3, solver=5)
    fs = FineSet(sim, scenes.fine_cloud(scene.pos, per, 0.02, seed=7))
    n, nf = sim.n, fs.n
    print(f"# scale scene: coarse={n} fine={nf} cells={scene.cells}", file=sys.stderr,
          flush=True)
    for _ in range(args.upsample_warmup):
        sim.step()
        fs.sort()
        fs.upsample()
    steps = max(10, args.upsample_steps // 4)
    torch.cuda.synchronize()
    lib.gg_profiler_enable(steps * 32 + 16)
    for _ in range(steps):
        sim.step()
        fs.sort()
        fs.upsample()
    torch.cuda.synchronize()
    tot = (C.c_double * len(_lib.STAGES))()
    cnt = (C.c_int * len(_lib.STAGES))()
    _lib.check("gg_profiler_collect", lib.gg_profiler_collect(tot, cnt, len(_lib.STAGES)))
    lib.gg_profiler_enable(0)
    avg = {name: 1000.0 * tot[k] / cnt[k] for k, name in enumerate(_lib.STAGES) if cnt[k]}
    plain_us = avg["upsample"] + avg["upsample_sort"]
    lpos = torch.empty((n, 24), dtype=torch.float32, device=sim.pos.device)
    learned.infer(sim, fine_pos=lpos)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    # 100 calls (~25 ms): a window long enough for a stable average
    calls = 100
    e0.record()
    for _ in range(calls):
        learned.infer(sim, fine_pos=lpos)
    e1.record()
    torch.cuda.synchronize()
    t_all = 1000.0 * e0.elapsed_time(e1) / calls
    _, tfp, _ = learned.teacher(sim)
    err = fine_error(lpos, tfp) / 0.01
    print(f"# scale: plain {plain_us:.1f} us, learned {t_all:.1f} us, error {err:.4f} r_fine",
          file=sys.stderr, flush=True)
    return {
        "workload": f"{side}^3 coarse lattice r=0.02 in a {box:.2f}^3 m box, {per} fine per "
                    f"coarse, {scene.cells[0]}^3 cells, steps {args.upsample_warmup}.."
                    f"{args.upsample_warmup + steps}; network trained on the 250k scene",
        "coarse": n, "fine": nf,
        "plain_us": round(plain_us, 3),
        "plain_fine_particles_per_s": round(nf / (plain_us * 1e-6), 1),
        "learned_us_per_call": round(t_all, 3),
        "learned_fine_particles_per_s": round(nf / (t_all * 1e-6), 1),
        "speedup_vs_plain": round(plain_us / t_all, 3),
        "error_vs_teacher_r_fine": round(err, 5),
    }


def excavator_phase(args, lib, torch, C, np):
    """The excavator dig at size through the host system: 4M particles of
    r = 0.004, a 200k-particle bucket on the scripted arc, adaptive merge/split
    with the radius-level rules, 160-candidate lists with the mid-step
    refresh, and the learned upsampler writing 8 fine particles per granular
    particle from GranularSystem::step (an untrained network: the weights do
    not change the coarse run or the inference cost).  Per-stage times from
    the library's stage profiler over steps --excavator-from..--excavator-to."""
    import math
    from granular_gpu_amd import _lib, host, scenes
    from granular_gpu_amd.learned import init_master

    r = 0.004
    a, b = args.excavator_from, args.excavator_to
    dig = scenes.excavator_dig(bed=(2.0, 1.0), fill=1.11, r=r, bucket_count=200_000)
    s = dig.scene
    s.density = np.float32(3.0 / (4.0 * math.pi * r ** 3))
    sysm = host.HostSystem.from_scene(s, is_move_boundary=True, is_adaptive=1,
                                      boundary_animated=dig.animated)
    sysm.set_merge_rules()
    sysm.set_lists(gmax=160, bmax=96, skin=0.75 * r)
    sysm.set_solver(3, 5, 0.0, 0.0, _lib.GG_FLAG_LIST_REFRESH)
    W, bias = init_master(6)
    sysm.set_learned((W.view(np.uint32) >> 16).astype(np.uint16), bias, radius=r,
                     guard_boundary=2.5 * r, guard_floor=4 * r, guard_side=2.5 * r)
    st = host.animation_default()
    for key, v in dig.script.items():
        if key in ("translation_direction", "rotation_center"):
            getattr(st, key)[:] = tuple(float(x) for x in v)
        else:
            setattr(st, key, v)
    dt = float(s.dt)
    print(f"# excavator: n={s.n} boundary={s.boundary.shape[0]}, steps {a}..{b}",
          file=sys.stderr, flush=True)
    for k in range(b):
        if k == a:
            lib.gg_profiler_enable(200000)
        sysm.update_animation(st, dt)
        sysm.step(1)
        if (k + 1) % 100 == 0:
            print(f"# excavator step {k + 1}: {sysm.size('granular')} particles",
                  file=sys.stderr, flush=True)
    tot = (C.c_double * len(_lib.STAGES))()
    cnt = (C.c_int * len(_lib.STAGES))()
    _lib.check("gg_profiler_collect", lib.gg_profiler_collect(tot, cnt, len(_lib.STAGES)))
    lib.gg_profiler_enable(0)
    ft = sysm.frame_times()
    out = {"workload": f"excavator dig, {s.n} particles r = {r} at step 0, 200k-particle bucket, "
                       f"adaptive (radius levels), learned 8-per-particle output, steps {a}..{b}",
           "particles": int(sysm.size("granular")), "fine": int(sysm.size("fine")),
           "frame_ms": round(float(np.mean(ft[a:b])), 3),
           "stages_ms_per_step": {name: round(tot[k] / (b - a), 3)
                                  for k, name in enumerate(_lib.STAGES) if cnt[k]},
           "launches_per_step": {name: round(cnt[k] / (b - a), 2)
                                 for k, name in enumerate(_lib.STAGES) if cnt[k]}}
    sysm.close()
    return out


def funnel_phase(args, lib, torch, C, np):
    """The funnel pour at size: 1M particles dropped as a column onto a cone
    of 400k boundary particles, adaptive merge/split with the radius-level
    rules (masses 0.125 / 1 / 8; merge deeper than 3 r below the free surface,
    split within 2 r of it), 3 + 5 sweeps.  Steps 0..400 untimed (the column
    lands and starts draining at step ~360), then steps 400..500 timed, with
    the library's stage profiler on.  Lists of 160 granular / 96 boundary
    candidates (a particle on the cone has more than 16 boundary candidates:
    with 16 it scans its stencil every sweep, 15.9 against 10.5 ms per step) and the
    mid-step list refresh; the lists do not change the results."""
    from granular_gpu_amd import _lib, scenes
    from granular_gpu_amd.device import AdaptiveSim

    s = scenes.funnel_column()
    sim = AdaptiveSim(s, capacity=4 * s.n, max_mass=8.0, stab=3, solver=5, gmax=160, bmax=96,
                      flags=_lib.GG_FLAG_LIST_REFRESH)
    sim.params.adaptive_rules = 1
    for key, v in dict(min_mass=0.125, split_mass=0.125, merge_dist=1.75, merge_depth=3.0,
                       split_depth=2.0).items():
        setattr(sim.params, key, v)
    print(f"# funnel: n={s.n} boundary={s.boundary.shape[0]}", file=sys.stderr, flush=True)
    a, b = 400, 500
    for _ in range(a // 100):
        sim.step(100)
        torch.cuda.synchronize()
        print(f"# funnel: {sim.n} particles", file=sys.stderr, flush=True)
    n0 = int(sim.n)
    lib.gg_profiler_enable(200000)
    t0 = time.perf_counter()
    sim.step(b - a)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    tot = (C.c_double * len(_lib.STAGES))()
    cnt = (C.c_int * len(_lib.STAGES))()
    _lib.check("gg_profiler_collect", lib.gg_profiler_collect(tot, cnt, len(_lib.STAGES)))
    lib.gg_profiler_enable(0)
    stats = np.array(sim.stats, np.int64)[a:b]
    psteps = int(stats[:, 2].sum())  # particles at the start of each timed step
    return {"workload": f"funnel column, {s.n} particles over a {s.boundary.shape[0]}-particle "
                        f"cone and walls, radius-level merge/split, steps {a}..{b}",
            "particles": [n0, int(sim.n)], "ms_per_step": round(1000 * el / (b - a), 3),
            "particle_steps_per_s": round(psteps / el, 1),
            "merges": int(stats[:, 0].sum()), "splits": int(stats[:, 1].sum()),
            "stages_ms_per_step": {name: round(tot[k] / (b - a), 3)
                                   for k, name in enumerate(_lib.STAGES) if cnt[k]}}


def batch_phase(args, lib, torch, C, np):
    """Batched scenes: 64 independent plain box scenes with particle counts
    log-uniform in [4k, 256k] (seeded; 4.17M particles in all), per-scene
    friction mu_s in [0.3, 0.7], mu_k = 0.8 mu_s, 3 + 10 sweeps, dt = 1/240
    (the plain box's), g = -9.81.  Headline: the scenes stored back to back
    in one set and advanced by gg_step_scenes, one launch per stage for all
    64, captured in a hipGraph whose block sizes the tuner picked.  Beside
    it: the per-scene graph (every scene's own step chain on parallel
    branches) and the same kernels launched eagerly scene after scene."""
    from granular_gpu_amd import scenes
    from granular_gpu_amd.device import BatchSim, PlainSim, SceneBatch

    dt, g = args.batch_dt, -9.81
    sizes = scenes.batch_box_sizes(args.batch_scenes)
    scs = [scenes.batch_box(side, seed=1 + k, dt=dt, g=g) for k, (side, _, _) in enumerate(sizes)]
    fr = [(ms, mk) for _, ms, mk in sizes]
    total = sum(sc.n for sc in scs)
    cells = [sc.cells[0] * sc.cells[1] * sc.cells[2] for sc in scs]
    bytes_step = sum(sc.n * bytes_per_particle_step(sc.n, c, sc.boundary.shape[0], 3, 13)
                     for sc, c in zip(scs, cells))
    print(f"# batch: {len(scs)} scenes, {total} particles", file=sys.stderr, flush=True)

    # lists lapse inside the step once the columns break through their
    # floors at this time step: rebuilt mid-step when the sweeps' stencil
    # scans say so (gated on the device, results unchanged)
    from granular_gpu_amd._lib import GG_FLAG_LIST_REFRESH
    # one set, one launch per stage
    sb = SceneBatch(scs, fr, stab=3, solver=10, flags=GG_FLAG_LIST_REFRESH)
    sb.step(args.batch_warmup)
    sb.capture(steps_per_launch=1)
    tune1 = sb.autotune(trials=5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sb.launch(args.batch_steps)
    torch.cuda.synchronize()
    fused = (time.perf_counter() - t0) / args.batch_steps
    assert torch.isfinite(sb.pos).all() and torch.isfinite(sb.vel).all(), "non-finite batch"
    sb.close_graph()
    del sb
    torch.cuda.empty_cache()

    # each scene's own step chain in one graph, and eagerly, over the same
    # steps from fresh copies of the scenes
    def per_scene(eager_mode):
        # (without the refresh: its two launches per refresh point would be
        # paid 64 times per step here)
        sims = [PlainSim(sc, stab=3, solver=10, mu_s=ms, mu_k=mk)
                for sc, (ms, mk) in zip(scs, fr)]
        b = BatchSim(sims, branches=min(4, len(sims)))
        b.step(args.batch_warmup)
        tune = None if eager_mode else b.autotune(trials=5)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if eager_mode:
            b.step_eager(args.batch_steps)
        else:
            b.step(args.batch_steps)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / args.batch_steps
        b.close()
        del b, sims
        torch.cuda.empty_cache()
        return el, tune

    if args.batch_one_set:
        graph = eager = float("nan")
        tune2 = None
    else:
        graph, tune2 = per_scene(False)
        eager, _ = per_scene(True)
    print(f"# batch timed {args.batch_steps} launches: one set {1000 * fused:.3f} ms/step, "
          f"per-scene graph {1000 * graph:.3f}, eager {1000 * eager:.3f}", file=sys.stderr,
          flush=True)
    ach = bytes_step / fused / 1e9
    out = {
        "workload": f"{len(scs)} plain box scenes, counts log-uniform in [4k, 256k] (seed 4), "
                    f"friction mu_s U(0.3, 0.7), mu_k = 0.8 mu_s, 3 + 10 sweeps, dt = {dt:.6g}, "
                    f"mid-step list refresh on (one set), "
                    f"steps {args.batch_warmup}..{args.batch_warmup + args.batch_steps}",
        "scenes": len(scs), "particles": total,
        "particle_steps_per_s": round(total / fused, 1),
        "ms_per_step": round(1000 * fused, 4),
        "launch": "one set, one launch per stage, captured (gg_scenes_graph)",
        "autotuned": tune1,
        "per_scene_graph": {"particle_steps_per_s": round(total / graph, 1),
                            "ms_per_step": round(1000 * graph, 4), "autotuned": tune2},
        "eager_particle_steps_per_s": round(total / eager, 1),
        "eager_ms_per_step": round(1000 * eager, 4),
        "step_hbm_model": {"bytes_per_step": int(bytes_step), "achieved_GBs": round(ach, 1),
                           "frac_of_hbm_roofline": round(ach / HBM_PEAK_GBS, 4)},
    }
    if args.batch_one_set:
        for key in ("per_scene_graph", "eager_particle_steps_per_s", "eager_ms_per_step"):
            out.pop(key)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--settle", type=int, default=100,
                    help="steps run before the warm-up, outside every timing (the pile lands "
                         "around step 71)")
    ap.add_argument("--graph-steps", type=int, default=None,
                    help="steps timed through the captured step graph after the timed region "
                         "(default: --steps; 0 skips)")
    ap.add_argument("--side", type=int, default=100, help="lattice side (100 -> 1M particles)")
    ap.add_argument("--box", type=float, default=2.4)
    ap.add_argument("--dt", type=float, default=0.002)
    ap.add_argument("--cpu-baseline-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-stage-timing", action="store_true")
    ap.add_argument("--stage-every", type=int, default=10,
                    help="every n-th timed step carries per-stage hipEvents (0 = none)")
    ap.add_argument("--no-upsample", action="store_true")
    ap.add_argument("--upsample-side", type=int, default=63, help="63 -> 250,047 coarse")
    ap.add_argument("--upsample-box", type=float, default=3.0)
    ap.add_argument("--upsample-steps", type=int, default=100)
    ap.add_argument("--upsample-warmup", type=int, default=20)
    ap.add_argument("--only", choices=["upsample", "batch", "excavator", "funnel"],
                    default=None,
                    help="run one secondary phase alone and print its object (profiling; "
                         "the excavator and the funnel run only this way)")
    ap.add_argument("--excavator-from", type=int, default=500)
    ap.add_argument("--excavator-to", type=int, default=600)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--scale-side", type=int, default=100,
                    help="coarse lattice side of the 8M-fine comparison (0 skips it)")
    ap.add_argument("--learned-train-steps", type=int, default=2000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=100)
    ap.add_argument("--batch-warmup", type=int, default=20)
    ap.add_argument("--batch-one-set", action="store_true",
                    help="time only the one-set batch (no per-scene graph / eager runs; "
                         "profiling)")
    ap.add_argument("--batch-dt", type=float, default=1.0 / 240.0,
                    help="time step of the batched scenes (the plain box's 1/240)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 512 MiB device-to-device copy (calibrates HBM counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        a = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "copies": 3}), flush=True)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU slab "
                          "decomposition is designed, not built",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)

    import ctypes as C

    import numpy as np
    import torch

    from granular_gpu_amd import _lib, scenes
    from granular_gpu_amd.device import PlainSim

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}), flush=True)
        sys.exit(1)
    torch.cuda.set_device(0)
    lib = _lib.load()
    if lib.gg_device_check() != 0:
        print(json.dumps({"error": "device is not gfx950"}), flush=True)
        sys.exit(1)

    if args.only in ("upsample", "batch", "excavator", "funnel"):
        # one secondary phase alone (for profiling it in isolation)
        fn = {"upsample": upsample_phase, "batch": batch_phase,
              "excavator": excavator_phase, "funnel": funnel_phase}[args.only]
        print(json.dumps({"only": args.only, args.only: fn(args, lib, torch, C, np)}),
              flush=True)
        return

    S, I, mu_s, mu_k = 3, 10, 0.5, 0.4
    scene = scenes.plain_box(args.side, args.box, dt=args.dt, g=-9.8)
    n, nb = scene.n, scene.boundary.shape[0]
    cells = scene.cells[0] * scene.cells[1] * scene.cells[2]
    sim = PlainSim(scene, stab=S, solver=I, mu_s=mu_s, mu_k=mu_k)
    print(f"# scene {scene.name}: n={n} walls={nb} cells={cells}", file=sys.stderr, flush=True)

    sim.step(args.settle)
    torch.cuda.synchronize()
    print(f"# settled: {args.settle} steps", file=sys.stderr, flush=True)
    sim.step(args.warmup)
    torch.cuda.synchronize()
    print(f"# warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    t_first = args.settle + args.warmup

    K = S + I
    # Per-stage hipEvents (on the launch stream) bracket every stage of every
    # --stage-every-th timed step; the events cost ~10% of the steps they
    # are on, so the others run without them.
    every = 0 if args.no_stage_timing else max(0, args.stage_every)
    sampled = [k for k in range(args.steps) if every and k % every == 0]
    if sampled:
        lib.gg_profiler_enable(len(sampled) * (2 * K + 4) + 16)  # sweeps + queue passes
        lib.gg_profiler_pause(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if sampled:
        for k in range(args.steps):
            on = k % every == 0
            if on:
                lib.gg_profiler_pause(0)
            sim.step(1)
            if on:
                lib.gg_profiler_pause(1)
    else:
        sim.step(args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    print(f"# timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr, flush=True)

    stage_ms = {}
    roofline = None
    if sampled:
        tot = (C.c_double * _lib.STAGES.__len__())()
        cnt = (C.c_int * len(_lib.STAGES))()
        _lib.check("gg_profiler_collect", lib.gg_profiler_collect(tot, cnt, len(_lib.STAGES)))
        lib.gg_profiler_enable(0)
        for k, name in enumerate(_lib.STAGES):
            if cnt[k]:
                stage_ms[name] = {"total_ms": round(tot[k], 4), "launches": cnt[k],
                                  "avg_us": round(1000.0 * tot[k] / cnt[k], 3)}
        sw = stage_ms.get("sweep")
        if sw:
            per_launch_bytes = n * 32 + nb * 12
            avg_s = sw["avg_us"] * 1e-6
            achieved = per_launch_bytes / avg_s / 1e9
            traffic = None
            pmc = ROOT / "profiles" / "pmc_sweep_m1.json"
            if pmc.exists():
                try:
                    traffic = json.loads(pmc.read_text()).get("hbm_bytes_per_launch")
                except Exception:
                    traffic = None
            roofline = {"kernel": "k_sweep<true, false, 64> (solver contact sweep)", "bound": "hbm",
                        "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                        "algorithmic_bytes_per_launch": per_launch_bytes,
                        "avg_launch_us": sw["avg_us"]}

    value = n * args.steps / elapsed
    bps = bytes_per_particle_step(n, cells, nb, S, K)
    step_frac = value * bps / (HBM_PEAK_GBS * 1e9)

    # the same fixed-N step captured into a hipGraph (gg_step_graph), timed
    # right after the eager window by events on the launch stream
    graph = None
    gsteps = args.steps if args.graph_steps is None else args.graph_steps
    if gsteps > 0:
        sg = sim.graph(1)
        sg.launch(2)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        sg.launch(gsteps)
        e1.record()
        torch.cuda.synchronize()
        g_ms = e0.elapsed_time(e1) / gsteps
        sg.close()
        t_g = t_first + args.steps + 2
        graph = {"steps": f"{t_g}..{t_g + gsteps}", "ms_per_step": round(g_ms, 4),
                 "particle_steps_per_s": round(n / (g_ms * 1e-3), 1)}
        print(f"# graph: {gsteps} steps at {g_ms:.4f} ms/step", file=sys.stderr, flush=True)

    pos, vel, mass = sim.host()
    assert np.isfinite(pos).all() and np.isfinite(vel).all(), "non-finite state"

    upsample = None
    if not args.no_upsample:
        upsample = upsample_phase(args, lib, torch, C, np)
    batch = None
    if not args.no_batch:
        batch = batch_phase(args, lib, torch, C, np)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po
        olib = po.load()
        p = po.params_from(scene, S, I, mu_s, mu_k)
        bpos, bcs = po.sorted_boundary(scene.boundary, scene.cell_length, scene.cells)

        def cpu_rate(max_steps, budget_s):
            """particle-steps/s of the FP32 oracle from the current state."""
            done, cp, cv, cm = 0, pos, vel, mass
            tc0 = time.perf_counter()
            while done < max_steps:
                cp, cv, cm = po.plain_steps_f32(cp, cv, cm, bpos, bcs, p, 1)[:3]
                done += 1
                if time.perf_counter() - tc0 >= budget_s:
                    break
            return n * done / (time.perf_counter() - tc0), done

        olib.oracle_set_threads(0)
        threads = olib.oracle_get_threads()
        rate_all, steps_all = cpu_rate(8, args.cpu_baseline_seconds)
        olib.oracle_set_threads(1)
        rate_one, steps_one = cpu_rate(2, args.cpu_baseline_seconds)
        olib.oracle_set_threads(0)
        model = None
        try:
            for line in Path("/proc/cpuinfo").read_text().splitlines():
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": round(rate_all, 1), "unit": "particle-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"{steps_all} step(s) on {threads} threads and {steps_one} on 1 thread "
                         f"of the same 1M-particle state after the GPU runs, FP32 oracle "
                         f"(C++/OpenMP, full 27-cell scans)",
               "one_thread": round(rate_one, 1), "cpu_model": model,
               "host_cpus": os.cpu_count()}

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain box {args.side}^3 jittered lattice (r=0.01, seed 1) in a "
                        f"{args.box}^3 m box, {scene.cells[0]}^3 cells, {nb} wall particles, "
                        f"dt={args.dt}, g=-9.8, {S} stabilisation + {I} solver sweeps, "
                        f"friction {mu_s}/{mu_k}, FP32, landed column: steps {t_first}.."
                        f"{t_first + args.steps} (after {args.settle} settling + "
                        f"{args.warmup} warm-up steps), neighbour-list skin {sim.params.skin}",
            "particles": n, "boundary_particles": nb, "cells": cells,
            "parallelism": "single GPU",
        },
        "step_hbm_model": {"bytes_per_particle_step": round(bps, 1),
                           "frac_of_hbm_roofline": round(step_frac, 4)},
        "stages": stage_ms,
        "stage_sample": (f"{len(sampled)} of the {args.steps} timed steps (every {every}th)"
                         if stage_ms else None),
        "roofline": roofline,
        "graph": graph,
        "cpu_baseline": cpu,
        "upsample": upsample,
        "batch": batch,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
