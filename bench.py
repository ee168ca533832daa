#!/usr/bin/env python3
"""Benchmark of the SPH2S solver step on one MI355X.

Workload: the 1,000,000-particle single-scale dam break (100^3 lattice,
domain [0, 3.125]^3, h = 0.0457, FP32 SoA float4 state), whose defined run is
1000 steps.  One "step" = hash -> radix sort -> reorder + cell ranges ->
density -> force + symplectic Euler + wall clamp over all particles, one
captured hipGraph launch.

The cost of a step grows several-fold as the column collapses and compresses
(28 neighbours per particle at rest, 200 at step 600), so a window of K
consecutive steps measures its own phase of the run.  The K timed steps are
therefore spread evenly over the 1000-step trajectory: step floor(k * 1000 / K)
for k = 0..K-1, each launched from the trajectory's own state, bracketed by a
synchronize and hipEvents on the solver's stream; the steps in between advance
the state untimed.  value = N * K / (sum of the K step times), an estimate of
the whole run's rate; trajectory_1000 times the whole run in one bracket.

Prints ONE JSON line with:
  value       particle-updates/s over the K sampled steps (state resident in HBM)
  window      which steps of the trajectory were timed
  roofline    dominant kernel (force+integrate): algorithmic bytes per launch
              (52 N + 8 C) / its mean duration over the same K steps, measured
              with hipEvents around each pass on the solver's stream by a
              second solver that replays the same trajectory (graph and eager
              launches give bitwise-identical states)
  trajectory_1000  the full 1000-step run timed in one bracket, and the force
              kernel's mean over all its steps (what rocprofv3 averages)
  cpu_baseline the FP32 OpenMP CPU step (a port of the same algorithm) on the
              same scene, timed for a bounded number of steps, with the FP64
              single-threaded oracle step beside it
  adjoint     the differentiable rollout (16 scenes x 65,536 particles, 200
              steps, checkpoint every 20): forward and forward+adjoint times and
              particle-updates/s, and their ratio
  batched     64 uneven scenes (log-uniform 4,096..262,144 particles, 500
              steps) in one batched step vs the same scenes as 64 single-scene
              solvers (each on its own stream, and serialised)
  walls       the reference's own single-scale scene: 215,460 fluid particles
              in a tank of 860,390 boundary particles (boundary-particle walls)
  two_scale   the two-scale dam break (63^3 low-resolution particles, 1->8
              split / 8->1 merge region over 20% of the domain), under the
              fixed coupling and under the reference's blended coupling
Each of adjoint, batched and two_scale carries a "cpu" figure: the same work
on the host, timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3
METRIC = "particle-updates/s at 1M particles, fwd and fwd+adjoint, % HBM roofline"
TRAJECTORY_STEPS = 1000  # the defined run of the 1M-particle dam break
LEGS = ("main", "trajectory", "adjoint", "batched", "walls", "two_scale", "two_scale_blended",
        "cpu")


def algorithmic_bytes(n: int, cells: int) -> dict:
    """Compulsory traffic per pass (each array read once / written once)."""
    return {
        "hash_reset": 20 * n,
        "radix_sort": 16 * n,
        "reorder_cells": 56 * n + 8 * cells,
        "density": 16 * n + 8 * cells,
        "force_integrate": 52 * n + 8 * cells,
        "step": 160 * n + 24 * cells,
    }


def load_traffic(workload: str):
    """Measured HBM bytes per force launch from the committed rocprofv3 summary."""
    path = os.path.join(ROOT, "profiles", f"traffic_force_integrate_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def host_cores() -> dict:
    """Host CPU counts: what the OpenMP CPU step uses, the process's affinity
    (nproc) and the machine's logical CPUs."""
    import oracle as O
    return {"openmp_threads": O.CpuSolver.threads(), "nproc": len(os.sched_getaffinity(0)),
            "logical_cpus": os.cpu_count()}


def cpu_baseline(scene, budget_s: float, max_steps: int, fp64_budget_s: float):
    """FP32 OpenMP step of the oracle library (a port of the same algorithm)
    on the same scene and state, then the FP64 single-threaded oracle step."""
    import oracle as O  # the CPU port of the step; test/bench infrastructure only
    cs = O.CpuSolver(O.params_from_config(scene.config), scene.positions, scene.velocities)
    cs.step(scene.dt, 1)  # warm-up
    done, t0 = 0, time.perf_counter()
    while done < max_steps:
        cs.step(scene.dt, 1)
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    out = {
        "value": scene.n * done / dt,
        "unit": "particle-updates/s",
        "cores": O.CpuSolver.threads(),
        "kind": "port",
        "sample": f"{scene.name}: steps 2..{done + 1} as {done} FP32 OpenMP steps of {scene.n} "
                  f"particles after 1 warm-up step ({dt:.1f} s)",
        "host": host_cores(),
    }
    out["fp64_single_thread"] = cpu_fp64_step(scene, fp64_budget_s)
    return out


def cpu_fp64_step(scene, budget_s: float) -> dict:
    """The FP64 oracle step, single-threaded, from the scene's initial state:
    one step to size the sample, then as many steps as fit the budget."""
    import oracle as O
    p = O.params_from_config(scene.config)
    t0 = time.perf_counter()
    O.rollout(p, scene.positions, scene.velocities, scene.dt, 1)
    t1 = time.perf_counter() - t0
    k = max(1, int(budget_s / max(t1, 1e-3)))
    t0 = time.perf_counter()
    O.rollout(p, scene.positions, scene.velocities, scene.dt, k)
    dt = time.perf_counter() - t0
    return {"value": scene.n * k / dt, "unit": "particle-updates/s", "cores": 1,
            "sample": f"{scene.name}: steps 1..{k} of the FP64 oracle, one thread ({dt:.1f} s)"}


def cpu_rollout(budget_s: float) -> dict:
    """FP64 oracle forward + adjoint of the rollout, one thread, on scene 0 of
    the 16-scene rollout (65,536 particles), for as many steps as fit."""
    import numpy as np
    import oracle as O
    from sph2s_final_amd import scenes
    bs = scenes.rollout_16x64k()
    s0 = bs.scenes[0]
    p = O.params_from_config(s0.config)
    x0, v0 = s0.positions.astype(np.float64), s0.velocities.astype(np.float64)
    t0 = time.perf_counter()
    O.rollout_loss_grad(p, x0, v0, bs.dt, 1, x0)
    t1 = time.perf_counter() - t0
    k = max(1, min(200, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    O.rollout_loss_grad(p, x0, v0, bs.dt, k, x0)
    dt = time.perf_counter() - t0
    return {"fwd_adjoint_pu_s": s0.n * k / dt, "cores": 1, "kind": "port",
            "sample": f"scene 0 ({s0.n} particles), {k} steps forward + adjoint, FP64 oracle, "
                      f"one thread ({dt:.1f} s)"}


def cpu_batched(bs, steps: int) -> dict:
    """The FP32 OpenMP CPU step over every scene of the batch, a few steps."""
    import oracle as O
    solvers = [O.CpuSolver(O.params_from_config(s.config), s.positions, s.velocities)
               for s in bs.scenes]
    t0 = time.perf_counter()
    for _ in range(steps):
        for c in solvers:
            c.step(bs.dt, 1)
    dt = time.perf_counter() - t0
    return {"pu_s": bs.n * steps / dt, "cores": O.CpuSolver.threads(), "kind": "port",
            "sample": f"all {len(bs.scenes)} scenes ({bs.n} particles), steps 1..{steps}, FP32 "
                      f"OpenMP ({dt:.1f} s)"}


def cpu_two_scale(sc, steps: int) -> dict:
    """The FP64 two-scale oracle step, one thread, from the scene's split state."""
    import oracle as O
    from sph2s_final_amd import Grid
    gh = Grid.MakeGrid((0.0, 0.0, 0.0), [sc.config.domain_hi[d] for d in range(3)],
                       0.5 * sc.config.h)
    o = O.OracleTwoScale(O.params_from_config(sc.config), gh, sc.region_lo, sc.region_hi, sc.n)
    o.set_state(sc.positions, sc.velocities)
    o.split_merge()
    active, t0 = 0, time.perf_counter()
    for _ in range(steps):
        c = o.counts()
        active += c[0] + c[1]
        o.step(sc.dt)
    dt = time.perf_counter() - t0
    return {"pu_s": active / dt, "cores": 1, "kind": "port",
            "sample": f"{sc.name}: steps 1..{steps} from the split state, FP64 oracle, one "
                      f"thread ({dt:.1f} s, {active // steps} active particles per step)"}


def bench_trajectory(torch, scene, alg) -> dict:
    """The whole 1000-step run: graph-replayed steps timed in one bracket,
    then the same trajectory replayed eagerly with per-pass hipEvents (the
    force kernel's mean over every step, as rocprofv3 averages it)."""
    from sph2s_final_amd import ParticleData, Solver
    from sph2s_final_amd import _capi as C
    n = scene.n
    s = Solver(ParticleData(n, scene.positions, scene.velocities), None, scene.config)
    stream = torch.cuda.ExternalStream(s.views().stream)
    s.synchronize()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    s.Advance(scene.dt, TRAJECTORY_STEPS)
    b.record(stream)
    s.synchronize()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / TRAJECTORY_STEPS
    s.close()
    r = Solver(ParticleData(n, scene.positions, scene.velocities), None, scene.config)
    per_pass = (C.ctypes.c_float * len(C.PASS_NAMES))()
    C.check(C.lib().sph_profile_passes(r._h, scene.dt, TRAJECTORY_STEPS, per_pass))
    r.close()
    pass_ms = {name: per_pass[i] / TRAJECTORY_STEPS for i, name in enumerate(C.PASS_NAMES)}
    f_gbs = alg["force_integrate"] / (pass_ms["force_integrate"] / 1e3) / 1e9
    return {"steps": TRAJECTORY_STEPS, "ms_per_step": ms, "pu_s": n / (ms / 1e3),
            "pass_ms": pass_ms, "force_integrate_gbs": f_gbs,
            "force_integrate_frac": f_gbs / HBM_PEAK_GBS,
            "step_hbm_frac": alg["step"] / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}


def _timed(fn, torch):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def bench_adjoint(torch, steps: int, every: int, record_all: bool = True,
                  cpu_budget_s: float = 0.0) -> dict:
    """Forward vs forward+adjoint of the differentiable batched rollout."""
    import numpy as np
    from sph2s_final_amd import scenes
    from sph2s_final_amd.autograd import Rollout
    bs = scenes.rollout_16x64k()
    pos = torch.from_numpy(np.concatenate([s.positions for s in bs.scenes])).cuda()
    vel = torch.from_numpy(np.concatenate([s.velocities for s in bs.scenes])).cuda()
    hidden_k, hidden_mu = torch.from_numpy(bs.stiffness), torch.from_numpy(bs.viscosity)
    guess_k = torch.full((len(bs.scenes),), scenes.STIFFNESS)
    guess_mu = torch.full((len(bs.scenes),), scenes.VISCOSITY)
    with Rollout([s.config for s in bs.scenes], bs.counts, bs.dt, steps, every) as r:
        target = r.forward(pos, vel, hidden_k, hidden_mu)
        x = r.forward(pos, vel, guess_k, guess_mu)   # warm-up (graph capture)
        r.backward((2.0 * (x - target) / x.numel()).contiguous())
        res = {}
        for _ in range(2):
            start, mid, stop = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            start.record()
            x = r.forward(pos, vel, guess_k, guess_mu)
            mid.record()
            seed = (2.0 * (x - target) / x.numel()).contiguous()
            gp, gv, gk, gm = r.backward(seed)
            stop.record()
            torch.cuda.synchronize()
            fwd, tot = start.elapsed_time(mid), start.elapsed_time(stop)
            if not res or tot < res["fwd_adjoint_ms"]:
                res = {"fwd_ms": fwd, "fwd_adjoint_ms": tot}
        n = bs.n
        res.update({
            "workload": bs.name, "particles": n, "scenes": len(bs.scenes), "steps": steps,
            "checkpoint_every": every, "device_bytes": r.device_bytes,
            "ratio": res["fwd_adjoint_ms"] / res["fwd_ms"],
            "fwd_pu_s": n * steps / (res["fwd_ms"] / 1e3),
            "fwd_adjoint_pu_s": n * steps / (res["fwd_adjoint_ms"] / 1e3),
            "loss": float(((x - target) ** 2).mean()),
            "grads_finite": bool(np.isfinite(gk).all() and np.isfinite(gm).all()
                                 and torch.isfinite(gv).all().item()),
            "device_errors": r.device_errors(),
        })
    if record_all and every < steps:
        # The same rollout keeping every step's record (no recomputation):
        # about 76 MB per step at this size, 15 GB for 200 steps.
        res["record_every_step"] = bench_adjoint(torch, steps, steps, record_all=False)
    if record_all and cpu_budget_s > 0:
        res["cpu"] = cpu_rollout(cpu_budget_s)
    return res


def bench_batched(torch, steps: int, cpu_steps: int = 0) -> dict:
    """64 uneven scenes: one batched step vs 64 single-scene solvers."""
    from sph2s_final_amd import BatchSolver, ParticleData, Solver, scenes
    bs = scenes.batched_64()
    n, dt = bs.n, bs.dt
    from sph2s_final_amd import autotune
    b = BatchSolver([s.config for s in bs.scenes], [s.positions for s in bs.scenes],
                    [s.velocities for s in bs.scenes])
    b.Advance(dt, 1)
    b.synchronize()
    word, tune_ms = b.autotune(dt, 5)
    b.Advance(dt, 1)
    tb = _timed(lambda: (b.Advance(dt, steps), b.synchronize()), torch)
    errs = b.device_errors()
    b.close()
    singles = [Solver(ParticleData(s.n, s.positions, s.velocities), None, s.config)
               for s in bs.scenes]
    for s in singles:
        s.Advance(dt, 1)
    for s in singles:
        s.synchronize()

    def streams():
        for _ in range(steps):
            for s in singles:
                s.Advance(dt, 1)
        for s in singles:
            s.synchronize()

    def serial():
        for _ in range(steps):
            for s in singles:
                s.Advance(dt, 1)
                s.synchronize()

    ts = _timed(streams, torch)
    tser = _timed(serial, torch)
    for s in singles:
        s.close()
    return {"workload": bs.name, "scenes": len(bs.scenes), "particles": n,
            "steps": steps, "pu_s": n * steps / tb, "ms_per_step": tb / steps * 1e3,
            "one_per_launch_streams_pu_s": n * steps / ts,
            "one_per_launch_serial_pu_s": n * steps / tser,
            "speedup_vs_streams": ts / tb, "speedup_vs_serial": tser / tb,
            "device_errors": errs,
            "tune": {"word": word, "launch": autotune.describe(word),
                     "density_ms_default_tuned": tune_ms[:2],
                     "force_ms_default_tuned": tune_ms[2:]},
            "note": "one-per-launch: each scene its own solver and graph; 'streams' "
                    "launches all 64 per step before waiting, 'serial' waits after each",
            **({"cpu": cpu_batched(bs, cpu_steps)} if cpu_steps > 0 else {})}


def bench_walls(torch, steps: int) -> dict:
    """The reference's own single-scale scene: 215,460 fluid particles in a
    tank of 860,390 boundary particles with a pillar (boundary-particle walls,
    tension 0.8, no wall clamp)."""
    from sph2s_final_amd import FIELD_POSITION, ParticleData, Solver, scenes
    sc = scenes.walled_dam_break_215k()
    s = Solver(ParticleData(sc.n, sc.positions), ParticleData(len(sc.walls), sc.walls), sc.config)
    s.Advance(sc.dt, 20)
    s.synchronize()
    el = _timed(lambda: (s.Advance(sc.dt, steps), s.synchronize()), torch)
    pass_ms = s.profile_passes(sc.dt, 20) / 20.0
    x = s.download(FIELD_POSITION)
    errs = s.device_errors()
    s.close()
    import numpy as np
    return {"workload": sc.name, "particles": sc.n, "walls": int(len(sc.walls)), "steps": steps,
            "ms_per_step": el / steps * 1e3, "pu_s": sc.n * steps / el,
            "pass_ms": {k: float(v) for k, v in zip(
                ["hash_reset", "radix_sort", "reorder_cells", "density", "walls_force_integrate"],
                pass_ms)},
            "state_finite": bool(np.all(np.isfinite(x))), "device_errors": errs}


def bench_two_scale(torch, steps: int, blended: bool = False, cpu_steps: int = 0) -> dict:
    """Two-scale dam break; blended = the reference's lambda-blended coupling
    (fade timeline with increment 0.0125, velocity cap 0.8 h / dt, r/sqrt(3)
    split offsets) instead of the fixed coupling."""
    from sph2s_final_amd import (FIELD_BLEND, SPH2_COUPLING_BLENDED, SPH2_SPLIT_REFERENCE,
                                 TwoScaleSolver, scenes)
    sc = scenes.two_scale_250k()
    t = TwoScaleSolver(sc.config, sc.region_lo, sc.region_hi, sc.positions, sc.velocities,
                       max_low=sc.n, max_high=8 * sc.n)
    if blended:
        t.set_options(coupling=SPH2_COUPLING_BLENDED, split_offsets=SPH2_SPLIT_REFERENCE,
                      blend_increment=0.0125, velocity_cap=0.8)
    t.Advance(sc.dt, 1)
    t.synchronize()
    c0 = t.counts()
    el = _timed(lambda: (t.Advance(sc.dt, steps), t.synchronize()), torch)
    c1 = t.counts()
    errs = t.device_errors()
    out = {"workload": sc.name + ("_blended" if blended else ""), "steps": steps,
           "ms_per_step": el / steps * 1e3}
    active = 0.5 * ((c0[0] + c0[1]) + (c1[2] + c1[3]))  # mean particles per step (end points)
    out.update({"pu_s": active * steps / el, "counts_start": list(c0[:2]),
                "counts_end": list(c1[:2]), "device_errors": errs,
                "note": "pu_s = mean of the start and end active counts (low + high) x steps / t"})
    if blended:
        import numpy as np
        m0 = float(np.float32(sc.config.mass))
        eff = m0 * t.download(0, FIELD_BLEND).astype(np.float64).sum() + \
            m0 / 8 * t.download(1, FIELD_BLEND).astype(np.float64).sum()
        out["effective_mass_ratio"] = eff / (m0 * sc.n)
    t.close()
    if cpu_steps > 0:
        out["cpu"] = cpu_two_scale(sc, cpu_steps)
    return out


def bench_main(torch, args) -> dict:
    """The headline: K steps of the 1M dam break spread over its 1000-step run."""
    from sph2s_final_amd import ParticleData, Solver, scenes
    from sph2s_final_amd import _capi as C
    scene = scenes.dam_break_1m()
    n = scene.n
    run_len = max(TRAJECTORY_STEPS, args.steps)
    sample = [k * run_len // args.steps for k in range(args.steps)]

    # Warm-up: W graph steps on a solver of its own (capture, clocks, caches).
    warm = Solver(ParticleData(n, scene.positions, scene.velocities), None, scene.config)
    warm.Advance(scene.dt, max(args.warmup, 1))
    warm.synchronize()
    warm.close()

    # The K timed steps, spread over the trajectory.
    solver = Solver(ParticleData(n, scene.positions, scene.velocities), None, scene.config)
    views = solver.views()
    cells = int(views.num_cells)
    stream = torch.cuda.ExternalStream(views.stream)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    step_ms, at = [], 0
    for s_k in sample:
        if s_k > at:
            solver.Advance(scene.dt, s_k - at)   # untimed: the trajectory up to step s_k
        solver.synchronize()
        torch.cuda.synchronize()
        start.record(stream)
        solver.Advance(scene.dt, 1)
        stop.record(stream)
        solver.synchronize()
        torch.cuda.synchronize()
        step_ms.append(start.elapsed_time(stop))
        at = s_k + 1
    ms = sum(step_ms)
    ms_per_step = ms / args.steps
    value = n * args.steps / (ms / 1e3)
    pos = solver.positions()
    finite = bool(__import__("numpy").isfinite(pos).all())
    dev_errors = solver.device_errors()
    solver.close()

    # Per-pass durations and pair counts at the same K steps: a second solver
    # follows the same trajectory and runs each sampled step eagerly with
    # hipEvents bracketing every pass on its stream (graph and eager steps are
    # bitwise identical), then counts that step's candidate and in-range pairs.
    replay = Solver(ParticleData(n, scene.positions, scene.velocities), None, scene.config)
    per_pass = (C.ctypes.c_float * len(C.PASS_NAMES))()
    cand = nbr = 0
    at = 0
    for s_k in sample:
        if s_k > at:
            replay.Advance(scene.dt, s_k - at)
        C.check(C.lib().sph_profile_passes(replay._h, scene.dt, 1, per_pass))
        c, b = replay.count_pairs()
        cand += c
        nbr += b
        at = s_k + 1
    replay.close()
    pass_ms = {name: per_pass[i] / args.steps for i, name in enumerate(C.PASS_NAMES)}
    cand /= args.steps
    nbr /= args.steps
    flop_step = 18.0 * cand + 50.0 * nbr
    valu_tflops = flop_step / (ms_per_step / 1e3) / 1e12

    alg = algorithmic_bytes(n, cells)
    force_s = pass_ms["force_integrate"] / 1e3
    achieved = alg["force_integrate"] / force_s / 1e9
    traffic = load_traffic(scene.name)
    step_gbs = alg["step"] / (ms_per_step / 1e3) / 1e9

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": scene.name, "particles": n, "cells": cells,
                   "domain": [scene.config.domain_hi[d] for d in range(3)],
                   "h": round(float(scene.config.h), 6), "dt": scene.dt, "graph": True,
                   "path": "forward"},
        "window": {"trajectory_steps": run_len, "timed_steps": args.steps,
                   "spacing": run_len / args.steps, "first": sample[0], "last": sample[-1],
                   "t_first_s": sample[0] * scene.dt, "t_last_s": sample[-1] * scene.dt,
                   "step_ms_min": min(step_ms), "step_ms_max": max(step_ms),
                   "note": "step floor(k * trajectory_steps / timed_steps), k = 0..timed_steps-1, "
                           "each timed alone; the steps between advance the state untimed"},
        "roofline": {"bound": "hbm", "kernel": "force_integrate", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "bytes_per_launch": alg["force_integrate"],
                     "launch_ms": pass_ms["force_integrate"]},
        "step_hbm": {"alg_bytes": alg["step"], "achieved_gbs": step_gbs,
                     "frac": step_gbs / HBM_PEAK_GBS},
        "pass_ms": pass_ms,
        "pass_hbm": {k: {"alg_bytes": alg[k], "gbs": alg[k] / (v / 1e3) / 1e9,
                         "frac": alg[k] / (v / 1e3) / 1e9 / HBM_PEAK_GBS}
                     for k, v in pass_ms.items() if k in alg and v > 0},
        "valu_model": {"candidates_per_particle": cand / n, "neighbours_per_particle": nbr / n,
                       "flop_per_step": flop_step, "achieved_tflops": valu_tflops,
                       "peak_tflops": FP32_VALU_PEAK_TFLOPS,
                       "frac": valu_tflops / FP32_VALU_PEAK_TFLOPS,
                       "frac_unpacked": valu_tflops / (FP32_VALU_PEAK_TFLOPS / 2),
                       "samples": args.steps,
                       "note": "18 flop per candidate pair (distance test, both passes) + 50 per "
                               "in-range pair, pairs counted at every timed step; "
                               "frac_unpacked against half the peak (packed FP32 is not used)"},
        "state_finite": finite,
        "device_errors": dev_errors,
    }
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-trajectory", action="store_true",
                    help="skip the full 1000-step run timed in one bracket")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline")
    ap.add_argument("--cpu-fp64-budget", type=float, default=8.0,
                    help="seconds of the FP64 single-threaded oracle step")
    ap.add_argument("--cpu-max-steps", type=int, default=400)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--batched-steps", type=int, default=500)
    ap.add_argument("--two-scale-steps", type=int, default=500)
    ap.add_argument("--walls-steps", type=int, default=500)
    ap.add_argument("--only-main", action="store_true",
                    help="skip the adjoint, batched and two-scale measurements")
    ap.add_argument("--legs", default=",".join(LEGS),
                    help="comma-separated measurements to run, of " + ",".join(LEGS))
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run plain 256 MiB device copies (scales rocprofv3 byte counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        x = torch.empty(64 << 20, dtype=torch.float32, device="cuda").uniform_()
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": x.numel() * 4, "copies": 5}), flush=True)
        return 0

    legs = set(v for v in args.legs.split(",") if v)
    unknown = legs - set(LEGS)
    if unknown:
        print(f"bench.py: unknown legs {sorted(unknown)}", file=sys.stderr)
        return 2
    if args.only_main:
        legs &= {"main", "trajectory", "cpu"}
    if args.no_trajectory:
        legs.discard("trajectory")
    if args.no_cpu_baseline:
        legs.discard("cpu")

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: the solver runs on exactly one GPU; multi-GPU runs are not supported "
              "(scenes of a batch would be bin-packed onto GPUs, one large scene split into "
              "x-slabs with one-cell halos; neither is built)", file=sys.stderr)
        return 2

    import torch
    from sph2s_final_amd import scenes

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1

    scene = scenes.dam_break_1m()
    if "main" in legs:
        out = bench_main(torch, args)
    else:  # a partial run (profiling one leg): not the benchmark line
        out = {"metric": METRIC, "value": None, "legs": sorted(legs)}
    if "trajectory" in legs:
        cells = int(scene.config.grid.dims[0]) * int(scene.config.grid.dims[1]) * \
            int(scene.config.grid.dims[2])
        out["trajectory_1000"] = bench_trajectory(torch, scene, algorithmic_bytes(scene.n, cells))
    cpu = "cpu" in legs
    if "adjoint" in legs:
        out["adjoint"] = bench_adjoint(torch, args.adjoint_steps, args.checkpoint_every,
                                       cpu_budget_s=4.0 if cpu else 0.0)
        out["value_fwd_adjoint"] = out["adjoint"]["fwd_adjoint_pu_s"]
    if "batched" in legs:
        out["batched"] = bench_batched(torch, args.batched_steps, cpu_steps=3 if cpu else 0)
    if "walls" in legs:
        out["walls"] = bench_walls(torch, args.walls_steps)
    if "two_scale" in legs:
        out["two_scale"] = bench_two_scale(torch, args.two_scale_steps,
                                           cpu_steps=3 if cpu else 0)
    if "two_scale_blended" in legs:
        out["two_scale_blended"] = bench_two_scale(torch, args.two_scale_steps, blended=True)
    if cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.cpu_budget, args.cpu_max_steps,
                                           args.cpu_fp64_budget)
        if out.get("value"):
            out["cpu_baseline"]["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
