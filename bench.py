"""Benchmark: PBF dam break, 1,000,000 particles, 2 solver iterations, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (the 1M dam-break configuration): 100^3 lattice at spacing 2r
(r = 0.005 m, seeded jitter +-0.1 r) in a 2x1x1 m box, h = 2.5 r, rho0 = 1000,
g = -9.81, dt = 1/120, 2 solver iterations, XSPH c = 0.01, hashed cell grid with a
2^22-slot table, FP32, 500 steps.  A step is one full PBF step (predict, grid build,
solve, velocity, XSPH, boundary); state stays in HBM.

The dam break changes character along its 500 steps (about 30 lattice steps with ~6
neighbours a particle, then a collapsed fluid with ~300), so the timed steps sample
the whole trajectory:
- W warm-up steps (graph replay);
- an untimed graph-replay pass over the configured trajectory, steps W .. W+500,
  keeping device snapshots of the state at K stratified steps W + floor((k + 1/2) 500 / K);
- the autotuner, timing each setting from 4 of those states spread over the trajectory;
- an untimed pre-pass runs the trajectory again from the end of the warm-up, eager with
  a HIP event after every launch (its mean is reported as `trajectory`);
- the timed region: K steps, one from each snapshot (restored outside the timed
  intervals), each eager with per-launch events and bracketed by synchronisations;
  `value` is particle-steps/s over them and the per-kernel durations (the roofline)
  are those of exactly these launches;
- the same K states stepped by graph replay, and by the drop-in launchers.

Prints ONE JSON line with the roofline of the dominant kernel, the batched ensemble,
the 8M splash (sampled over its 300 steps the same way), the adjoint rollout and the CPU
oracle timed on this host on a bounded sample of the same window (plus the
32k-particle 200-step CPU configuration in full).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "cuda-pbf-fluid-sim_amd"))
sys.path.insert(0, REPO)

from pbfx import params as P  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
# VALU issue rate of the whole chip in wave64 instructions per second, from the
# published 157.3 TFLOP/s FP32 vector peak: that figure counts packed FP32 FMAs
# (2 lanes x 2 flops per lane-op), so one wave64 instruction a SIMD every 4 cycles is
# 157.3e12 / (64 lanes x 2 x 2).
VALU_PEAK = 157.3e12 / (64 * 2 * 2)

# Algorithmic bytes per particle of each launch label (FP32 xyz = 12 B, each datum
# counted once per kernel): predict+clamp 36 + hash 20; sort 16 per pass + 4; the
# candidate mask's build 16 (its mask words depend on the state, so the measured
# traffic counts them); lambda 16; delta+apply 28 (+36 for the fused velocity on the
# last iteration); XSPH + clamp 36.
BYTES_PER_PARTICLE = {
    "predict_hash": 56.0,
    "mask_build": 16.0,
    "lambda": 16.0,
    "delta": 28.0,
    "delta_last": 64.0,
    "xsph_scatter": 36.0,
}

# HIP kernels of the launch labels (delta and delta_last are the first and last
# launches of one kernel).
KERNELS = {
    "k_predict_hash": ["predict_hash"],
    "k_mask_build": ["mask_build"],
    "k_lambda_sorted": ["lambda"],
    "k_delta_sorted": ["delta", "delta_last"],
    "k_xsph_scatter": ["xsph_scatter"],
}

TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic_dam_break_1m.json")
INSTR_FILE = os.path.join(REPO, "profiles", "instructions_dam_break_1m.json")

# Launch knobs of the neighbour walk, which change what its kernels execute and read;
# profiled counts hold only for the setting they were measured at.  (The sort's knobs
# do not touch these kernels.)
WALK_KNOBS = ("candidate_mask", "stream_ratio", "chunk_bits", "walk_block", "lane_order")

# The dam break's tuning space: the walk settings its counter summaries cover (the
# candidate mask on or off, 128 or 256 lanes a block), the rest of the walk fixed at
# what every measurement of the collapsed fluid chose (per-lane walk, chunk masks, lane
# order), and the sort's knobs, which do not touch the profiled kernels.
BENCH_FIXED = {"stream_ratio": 1, "chunk_bits": True, "lane_order": True}
BENCH_SPACE = {"candidate_mask": (True, False), "stream_ratio": (1,), "chunk_bits": (True,),
               "sort_items": (0, 4, 8, 16), "sort_digit_bits": (8, 6, 4), "walk_block": (256, 128),
               "lane_order": (True,)}


def sample_offsets(total, k):
    """k stratified steps of a total-step trajectory: the midpoints of k equal strata."""
    return [((2 * j + 1) * total) // (2 * k) for j in range(k)]


def _profiled(path, kernel, steps, knobs, field):
    """Per-launch mean of a profiled figure (`field`: "valu" or "bytes") of `kernel` over the
    dam break's `steps` (absolute step numbers), from a counter summary kept next to the
    other profiles, for the same walk knobs; or (None, reason).  The summary keeps the
    figure of each step the bench samples by default and the mean over the trajectory;
    other steps get that mean (said in the note).  The simulation is deterministic, so a
    step's counts are the same in every run."""
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, "no counter summary"
    for run in rec.get("runs", []):
        prof = run.get("tuning") or {}
        if not all(prof.get(n) == knobs.get(n) for n in WALK_KNOBS):
            continue
        ps = (run.get("per_step") or {}).get(kernel)
        src = f"{os.path.relpath(path, REPO)} ({rec.get('command', '').replace('<TUNING>', json.dumps(prof))})"
        if not ps:
            return None, f"{kernel} not in {os.path.relpath(path, REPO)}"
        per = ps.get("steps") or {}
        if per and all(str(t) in per for t in steps):
            return sum(per[str(t)] for t in steps) / (len(steps) * ps["launches_per_step"]), src
        if "trajectory_mean" in ps:
            return ps["trajectory_mean"], f"{src}; mean over the profiled trajectory {ps['trajectory']} " \
                                          "(other steps timed)"
        return None, f"{os.path.relpath(path, REPO)} holds no figures for the timed steps"
    return None, "nothing profiled at this run's launch knobs"


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def sort_bytes(key_bits):
    passes = max(1, -(-key_bits // 8))
    return 16.0 * passes + 4.0


class Window:
    """Per-label launch times and wall time of a set of steps."""

    def __init__(self):
        self.ms, self.cnt, self.wall_s, self.steps = {}, {}, 0.0, 0

    def add(self, launches, wall_s):
        for label, ms in launches:
            self.ms[label] = self.ms.get(label, 0.0) + ms
            self.cnt[label] = self.cnt.get(label, 0) + 1
        self.wall_s += wall_s
        self.steps += 1

    def kernels(self, n, bpp):
        out = {}
        for name, labels in KERNELS.items():
            launches = sum(self.cnt.get(l, 0) for l in labels)
            if not launches:
                continue
            ms = sum(self.ms.get(l, 0.0) for l in labels)
            nbytes = sum(bpp[l] * n * self.cnt.get(l, 0) for l in labels)
            out[name] = {"labels": labels, "launches_per_step": launches / self.steps,
                         "avg_launch_ms": round(ms / launches, 5), "ms_per_step": round(ms / self.steps, 5),
                         "bytes_per_launch": nbytes / launches, "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                         "frac": round(nbytes / (ms * 1e-3) / HBM_PEAK, 5)}
        return out


def timed_step(sim, state):
    """Restore `state` (device tensors), then one eager step with per-launch events,
    bracketed by synchronisations: ([(label, ms), ...], wall seconds)."""
    sim.set_state_device(*state)
    sim.synchronize()
    t0 = time.perf_counter()
    launches = sim.profile_launches()
    sim.synchronize()
    return launches, time.perf_counter() - t0


def graph_step_s(sim, state):
    """Restore `state`, then one graph-replayed step between synchronisations (seconds)."""
    sim.set_state_device(*state)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(1)
    sim.synchronize()
    return time.perf_counter() - t0


def bench_drop_in(scene, states, labels):
    """The same states through the drop-in launchers (LaunchLambda / LaunchDeltaApply /
    LaunchXSPH over the reference's layout, dense cell ranges), as the reference's own
    phase pipeline would call them after swapping in this library: one graph-replayed
    step from each state."""
    from pbfx.sim import Simulator
    sim = Simulator.from_scene(scene, grid="dense", path="reference")
    try:
        sim.set_state_device(*states[0])
        sim.step(1)  # captures the graphs
        sim.synchronize()
        dt_s = sum(graph_step_s(sim, st) for st in states)
        n = sim.n
    finally:
        sim.close()
    print(f"[bench] drop-in launchers: {len(states)} sampled steps in {dt_s:.4f} s", file=sys.stderr, flush=True)
    return {"particle_steps_per_s": round(n * len(states) / dt_s, 1),
            "ms_per_step": round(dt_s * 1e3 / len(states), 5), "steps": labels, "grid": "dense",
            "path": "reference layout, drop-in launchers, graph replay"}


def bench_batched(count, samples, trajectory, tune_steps=0, grid="dense"):
    """The batched ensemble (one launch per phase for all scenes, graph replay) against the
    same scenes as separate plain simulators, each replaying its own captured step, on
    `samples` stratified states of the configured `trajectory` (200 steps):
    - scene by scene: one scene after another, one synchronisation per scene (the plain
      path run over the scenes);
    - concurrent streams: every scene's step enqueued on its own stream before a single
      synchronisation (the strongest unbatched arrangement).
    The pre-pass that reaches the sampled states is graph replay; its mean is reported."""
    import torch
    from pbfx.ensemble import Ensemble, ensemble_scenes, solo
    scenes = ensemble_scenes(count, seed=2024)
    total = sum(sc.num_particles for sc in scenes)
    ens = Ensemble(scenes, grid=grid)
    offs = sample_offsets(trajectory, samples)
    snaps, traj_s = [], 0.0
    done = 0
    for o in offs + [trajectory]:
        t0 = time.perf_counter()
        ens.step(o - done)
        ens.synchronize()
        traj_s += time.perf_counter() - t0
        done = o
        if o < trajectory:
            snaps.append(Simulator_state(ens))
    best = None
    if tune_steps:
        from pbfx import tune
        ens.set_state_device(*snaps[len(snaps) // 2])
        best, _, _ = tune.autotune(ens, steps=tune_steps, key=tune.workload_key(ens, f"ensemble{count}"))
    ens.set_state_device(*snaps[0])
    ens.step(1)  # (re)captures the graphs
    ens.synchronize()
    t_ens = sum(graph_step_s(ens, st) for st in snaps)
    firsts = [ens.scene_info(k)[0] for k in range(len(scenes))]
    ens.close()
    print(f"[bench] ensemble {count} scenes / {total} particles: {len(snaps)} sampled steps in {t_ens:.4f} s; "
          f"{trajectory}-step trajectory {traj_s:.3f} s", file=sys.stderr, flush=True)
    sims = [solo(sc, grid=grid) for sc in scenes]
    for sim in sims:
        if best:
            sim.set_tuning(**best)

    def load(st):
        X, V = st
        for sim, a, sc in zip(sims, firsts, scenes):
            sim.set_state_device(X[a:a + sc.num_particles].contiguous(), V[a:a + sc.num_particles].contiguous())
        torch.cuda.synchronize()

    load(snaps[0])
    for sim in sims:
        sim.step(1)  # captures each simulator's graphs
        sim.synchronize()
    t_seq = t_streams = 0.0
    for st in snaps:
        load(st)
        t0 = time.perf_counter()
        for sim in sims:
            sim.step(1)
            sim.synchronize()
        t_seq += time.perf_counter() - t0
        load(st)
        t0 = time.perf_counter()
        for sim in sims:
            sim.step(1)
        for sim in sims:
            sim.synchronize()
        t_streams += time.perf_counter() - t0
    for sim in sims:
        sim.close()
    print(f"[bench] scene by scene: {t_seq:.4f} s; concurrent streams: {t_streams:.4f} s", file=sys.stderr,
          flush=True)
    n = len(snaps)
    return {"scenes": count, "particles": total, "grid": grid,
            "timed_steps": [o for o in offs], "trajectory_steps": trajectory,
            "particle_steps_per_s": round(total * n / t_ens, 1),
            "trajectory_particle_steps_per_s": round(total * trajectory / traj_s, 1),
            "trajectory_knobs": "the defaults (the pass that took the snapshots, before tuning)",
            "scene_by_scene_particle_steps_per_s": round(total * n / t_seq, 1),
            "speedup_vs_scene_by_scene": round(t_seq / t_ens, 3),
            "concurrent_streams_particle_steps_per_s": round(total * n / t_streams, 1),
            "speedup_vs_concurrent_streams": round(t_streams / t_ens, 3), "tuning": best}


def bench_splash(samples, warmup, tune_steps=0, grid="hashed"):
    """The 8M splash (two 100x100x400 columns, r = 0.004 m, 4 iterations, dt = 1/240,
    2^24-slot hashed grid) over its configured 300 steps: a graph-replay pass that keeps
    `samples` stratified snapshots, optionally the tuner over two of them, the trajectory
    again from the end of the warm-up with the chosen knobs (its mean is the trajectory
    figure), then one graph-replayed step from each snapshot."""
    from pbfx.sim import Simulator
    scene = P.splash_8m(grid=grid)
    T = scene.steps
    offs = sample_offsets(T, samples)
    sim = Simulator.from_scene(scene, grid=grid)

    def trajectory(keep):
        snaps, done, t = [], 0, 0.0
        for o in offs + [T]:
            t0 = time.perf_counter()
            sim.step(o - done)
            sim.synchronize()
            t += time.perf_counter() - t0
            done = o
            if keep and o < T:
                snaps.append(sim.state_device())
        return snaps, t

    try:
        n = sim.n
        sim.step(warmup)
        sim.synchronize()
        warm = sim.state_device()
        snaps, _ = trajectory(True)
        best = None
        if tune_steps:
            from pbfx import tune
            pick = [snaps[len(snaps) // 4], snaps[(3 * len(snaps)) // 4]]
            best, _, _ = tune.autotune(sim, steps=tune_steps, key=tune.workload_key(sim, "splash_8m"), states=pick)
        sim.set_state_device(*warm)
        sim.step(1)  # (re)captures the graphs for the chosen knobs
        sim.set_state_device(*warm)
        sim.synchronize()
        del warm
        _, traj_s = trajectory(False)
        t = sum(graph_step_s(sim, st) for st in snaps)
        overflow = sim.hash_overflow()
    finally:
        sim.close()
    print(f"[bench] splash {n} particles: {len(snaps)} sampled steps in {t:.4f} s; {T}-step trajectory "
          f"{traj_s:.3f} s", file=sys.stderr, flush=True)
    return {"particles": n, "grid": f"{grid} (2^{scene.table_log2} slots)" if grid == "hashed" else grid,
            "solver_iters": scene.solver_iters, "trajectory_steps": f"{warmup}..{warmup + T}",
            "timed_steps": [warmup + o for o in offs], "particle_steps_per_s": round(n * len(snaps) / t, 1),
            "ms_per_step": round(t * 1e3 / len(snaps), 4),
            "trajectory_particle_steps_per_s": round(n * T / traj_s, 1),
            "trajectory_ms_per_step": round(traj_s * 1e3 / T, 4), "tuning": best, "hash_overflow": overflow,
            "timed": "one graph-replayed step from each stratified state (graphs captured beforehand); the "
                     "trajectory figure is a graph-replay pass over the whole run with the same knobs"}


def Simulator_state(sim):
    """Device copies of a simulator's (or ensemble's) state."""
    from pbfx.sim import Simulator
    return Simulator.state_device(sim)


# The walk knobs a rollout tunes for itself: its scenes are not the dam break the main
# line tunes on, and the candidate mask pays off only at some densities.
ROLLOUT_SPACE = {"candidate_mask": (True, False), "walk_block": (256, 128), "stream_ratio": (1, 2)}


def tune_rollout(sim, tune_steps):
    """Tune the walk knobs on the simulator's current state (restored afterwards); returns
    the chosen knobs, or None when tune_steps is 0."""
    if not tune_steps:
        return None
    from pbfx import tune
    best, _, _ = tune.autotune(sim, steps=tune_steps, space=ROLLOUT_SPACE, key=None)
    return {k: best[k] for k in ROLLOUT_SPACE}


def bench_adjoint_batched(steps, checkpoint, rollouts=64, knobs=None, tune_steps=0):
    """64 differentiable rollouts (each the 64x32x32 configuration, own jitter seed and
    target velocity field) advanced and differentiated together as one ensemble."""
    import torch
    from pbfx.diff import EnsembleRollout, differentiable_scene, occupancy_grid, seeded_velocity
    scenes = [differentiable_scene(seed=7 + k) for k in range(rollouts)]
    roll = EnsembleRollout(scenes, steps, checkpoint=checkpoint)
    try:
        if knobs:
            roll.sim.set_tuning(**knobs)
        X0, _ = roll.ensemble.state()
        n = len(X0)
        h = float(roll.params.kernel.h)
        x0 = torch.from_numpy(X0).cuda()
        vt = torch.from_numpy(np.concatenate([seeded_velocity(sc.num_particles, 1000 + k, 0.5)
                                              for k, sc in enumerate(scenes)])).cuda()
        XT, _ = roll.forward(x0, vt)
        chosen = tune_rollout(roll.sim, tune_steps) or knobs
        spans = [(f, f + sc.num_particles) for f, sc in zip(roll.firsts, scenes)]
        targets = [occupancy_grid(XT[a:b, :3].double(), sc.box_min, sc.box_max, cells=32, sigma=h)
                   for (a, b), sc in zip(spans, scenes)]
        v0 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        roll.forward(x0, v0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        XT, _ = roll.forward(x0, v0)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        xT = XT[:, :3].double().requires_grad_(True)
        loss = sum(((occupancy_grid(xT[a:b], sc.box_min, sc.box_max, cells=32, sigma=h) - tg) ** 2).sum()
                   for (a, b), sc, tg in zip(spans, scenes, targets))
        loss.backward()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gX0, gV0, gc = roll.backward(xT.grad.float())
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        finite = bool(torch.isfinite(gV0).all().item() and torch.isfinite(gX0).all().item()
                      and np.all(np.isfinite(gc)))
    finally:
        roll.close()
    print(f"[bench] {rollouts} batched rollouts, {n} particles: forward {t_fwd:.4f} s, backward {t_bwd:.4f} s",
          file=sys.stderr, flush=True)
    return {"rollouts": rollouts, "particles": n, "knobs": chosen or "defaults",
            "forward_ms_per_step": round(t_fwd * 1e3 / steps, 4),
            "backward_over_forward": round(t_bwd / t_fwd, 3),
            "rollout_steps_per_s": round(rollouts * steps / (t_fwd + t_bwd), 1), "grad_finite": finite}


def bench_adjoint(steps=60, checkpoint=10, rollouts=64, knobs=None, tune_steps=0):
    """The differentiable rollout: 65,536 particles (64x32x32 lattice), 60 steps, checkpoint every
    10.  Loss: squared L2 distance between the 32^3 Gaussian occupancy grid (sigma = h) of X_T and
    that of a target rollout from a seeded perturbed initial velocity.  Times the checkpointed
    forward, then the reverse pass split into re-recording the segments and the reverse kernels
    (renormalised every step; the gradient is returned in float64)."""
    import torch
    from pbfx.diff import Rollout, differentiable_scene, initial_state, occupancy_grid, seeded_velocity
    scene = differentiable_scene()
    sp = P.build_sim_params(scene)
    n = sp.numParticles
    h = float(sp.kernel.h)
    X0, _ = initial_state(scene)
    x0 = torch.from_numpy(X0).cuda()
    roll = Rollout(sp, steps, checkpoint=checkpoint)
    try:
        if knobs:
            roll.sim.set_tuning(**knobs)
        vt = torch.from_numpy(seeded_velocity(n, 77, 0.5)).cuda()
        XT, _ = roll.forward(x0, vt, sp.xsph_c)
        chosen = tune_rollout(roll.sim, tune_steps) or knobs
        target = occupancy_grid(XT[:, :3].double(), scene.box_min, scene.box_max, cells=32, sigma=h)
        v0 = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
        roll.forward(x0, v0, sp.xsph_c)  # warm-up of the same rollout
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        XT, _ = roll.forward(x0, v0, sp.xsph_c)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        xT = XT[:, :3].double().requires_grad_(True)
        loss = ((occupancy_grid(xT, scene.box_min, scene.box_max, cells=32, sigma=h) - target) ** 2).sum()
        loss.backward()
        roll.timed = True
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gX0, gV0, gc = roll.backward(xT.grad.float())
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        tm = dict(roll.timing)
        gnorm = float(gV0[:, :3].norm())
        finite = bool(torch.isfinite(gV0).all().item() and torch.isfinite(gX0).all().item() and math.isfinite(gc))
        log2 = roll.log2_scale
    finally:
        roll.close()
    print(f"[bench] adjoint rollout {n} particles x {steps} steps: forward {t_fwd:.4f} s, backward {t_bwd:.4f} s",
          file=sys.stderr, flush=True)
    batched = bench_adjoint_batched(steps, checkpoint, rollouts, knobs, tune_steps) if rollouts > 1 else None
    return {"particles": n, "steps": steps, "checkpoint": checkpoint, "loss": float(loss.detach()), "batched": batched,
            "knobs": chosen or "defaults",
            "forward_ms_per_step": round(t_fwd * 1e3 / steps, 4),
            "record_ms_per_step": round(tm["record_s"] * 1e3 / steps, 4),
            "reverse_ms_per_step": round(tm["reverse_s"] * 1e3 / steps, 4),
            "backward_over_forward": round(t_bwd / t_fwd, 3),
            "reverse_kernels_over_forward": round(tm["reverse_s"] / t_fwd, 3),
            "grad_v0_norm": gnorm, "grad_c": gc, "grad_finite": finite, "grad_log2_scale": log2,
            "note": "forward = eager checkpointed rollout; backward = re-record every segment + reverse kernels, "
                    "renormalised every step by a power of two (the dynamics amplify the gradient about 8x a "
                    "step; grad_log2_scale is the exponent taken out), result in float64"}


def cpu_baseline(scene, sp, states, labels, cfg1_gpu):
    """The CPU oracle on this host: one step from each of `states` (host arrays of the
    same sampled window as the GPU's timed steps), median of 3 runs each, on every host
    thread; one of them on one thread; and the 32,768-particle dam break (dense grid,
    200 steps from the seeded lattice) in full, median of 3."""
    from oracle import oracle as O
    cores = O.num_threads()
    n = sp.numParticles
    meds = []
    for X, V in states:
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            O.run(X, V, sp, 1, compact=True)
            runs.append(time.perf_counter() - t0)
        meds.append(sorted(runs)[1])
    O.set_num_threads(1)
    X, V = states[len(states) // 2]
    t0 = time.perf_counter()
    O.run(X, V, sp, 1, compact=True)
    one_s = time.perf_counter() - t0
    O.set_num_threads(cores)
    s32 = P.dam_break_32k()
    sp32 = P.build_sim_params(s32)
    X32 = P.seed_scene_reference(s32, O.seed_lattice)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        O.run(X32, np.zeros_like(X32), sp32, s32.steps)
        runs.append(time.perf_counter() - t0)
    cfg1_s = sorted(runs)[1]
    n32 = sp32.numParticles
    print(f"[bench] cpu: sampled steps {[round(m, 3) for m in meds]} s, one thread {one_s:.2f} s, "
          f"32k x 200 {[round(r, 2) for r in runs]} s", file=sys.stderr, flush=True)
    return {"value": round(n * len(meds) / sum(meds), 1), "unit": "particle-steps/s", "cores": cores,
            "kind": "port",
            "sample": f"dam_break_1m, compact grid: one step from each of the states before steps {labels} "
                      f"(the GPU's timed window), median of 3 runs each, OpenMP over particles, "
                      f"{sum(meds):.2f} s",
            "single_thread_value": round(n / one_s, 1),
            "single_thread_sample": f"the step from step {labels[len(labels) // 2]} on one thread, {one_s:.2f} s",
            "dam_break_32k": {"particles": n32, "steps": s32.steps, "grid": "dense",
                              "cpu_s_median_of_3": round(cfg1_s, 3),
                              "cpu_particle_steps_per_s": round(n32 * s32.steps / cfg1_s, 1),
                              "gpu_particle_steps_per_s": cfg1_gpu,
                              "gpu_over_cpu": round(cfg1_gpu / (n32 * s32.steps / cfg1_s), 1)},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def gpu_dam_break_32k():
    """The 32,768-particle dam break on the GPU: its 200 steps, graph replay (dense grid)."""
    from pbfx.sim import Simulator
    scene = P.dam_break_32k()
    sim = Simulator.from_scene(scene)
    try:
        X0, V0 = sim.state_device()
        sim.step(2)  # captures the graphs of both ping-pong parities
        sim.set_state_device(X0, V0)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(scene.steps)
        sim.synchronize()
        dt_s = time.perf_counter() - t0
    finally:
        sim.close()
    return round(scene.num_particles * scene.steps / dt_s, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps, sampled along the trajectory")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--trajectory", type=int, default=None,
                    help="steps of the configured run the timed steps sample (default: the configuration's 500)")
    ap.add_argument("--tune-states", type=int, default=4,
                    help="the autotuner times each setting from this many states evenly spread over the "
                         "trajectory (taken by a graph-replay pass over it), then the run restarts from the end "
                         "of the warm-up")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch knobs")
    ap.add_argument("--tune-steps", type=int, default=5)
    ap.add_argument("--tuning", default=None,
                    help='fixed launch knobs as JSON (e.g. {"candidate_mask": true, "stream_ratio": 1}); '
                         'skips the autotuner, as profiler runs do to time one setting')
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--grid", default="hashed", choices=["hashed", "compact", "dense"])
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true", help="skip timing the drop-in launchers' path")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-samples", type=int, default=5)
    ap.add_argument("--no-splash", action="store_true", help="skip the 8M splash section")
    ap.add_argument("--splash-samples", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that "
                         "scales the HBM byte counters of a profiler run)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported in this build "
              "(the multi-GPU slab decomposition is a design, not an implementation)", file=sys.stderr)
        return 2

    from pbfx.sim import Simulator

    if a.copy_calibration_mib > 0:
        import torch
        src = torch.ones(a.copy_calibration_mib << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    scene = P.dam_break_1m()
    sp = P.build_sim_params(scene)
    N = sp.numParticles
    T = a.trajectory or scene.steps
    K = max(1, min(a.steps, T))
    W = a.warmup
    offs = sample_offsets(T, K)
    sim = Simulator.from_scene(scene, grid=a.grid)
    print(f"[bench] seeded {N} particles, grid={a.grid}, warmup {W}, trajectory {T}, {K} sampled steps",
          file=sys.stderr, flush=True)
    sim.step(W)
    sim.synchronize()
    # Untimed graph-replay pass over the configured trajectory, keeping a device snapshot
    # of the state before each sampled step (the knobs never change results, so these
    # are the states of any later run).
    warm = sim.state_device()
    snaps = []
    done = 0
    for o in offs:
        sim.step(o - done)
        done = o
        snaps.append(sim.state_device())
    tuning = None
    if a.tuning:
        sim.set_tuning(**json.loads(a.tuning))
        tuning = {"chosen": sim.tuning(), "fixed": True}
    elif not a.no_tune:
        from pbfx import tune
        pick = sorted({(2 * j + 1) * len(snaps) // (2 * a.tune_states) for j in range(a.tune_states)})
        sim.set_tuning(**BENCH_FIXED)
        best, best_ms, table = tune.autotune(sim, steps=a.tune_steps, key=tune.workload_key(sim, "dam_break_1m"),
                                             space=BENCH_SPACE,
                                             log=lambda m: print(m, file=sys.stderr, flush=True),
                                             states=[snaps[i] for i in pick])
        tuning = {"chosen": best, "ms_per_step_at_tuning": round(best_ms, 4), "settings_tried": len(table),
                  "steps_per_setting": a.tune_steps,
                  "states": f"{a.tune_steps} steps from each of the states before steps {[W + offs[i] for i in pick]}"}
    sim.set_state_device(*warm)
    del warm

    # Untimed pre-pass: the configured trajectory again, eager with per-launch events (the
    # trajectory's mean step and per-kernel times).
    traj = Window()
    for t in range(T):
        t0 = time.perf_counter()
        launches = sim.profile_launches()
        traj.add(launches, time.perf_counter() - t0)
    print(f"[bench] trajectory steps {W}..{W + T}: {traj.wall_s:.3f} s", file=sys.stderr, flush=True)
    steps_abs = [W + o for o in offs]

    # Timed region: one eager step from each snapshot.
    win = Window()
    for st in snaps:
        launches, dt = timed_step(sim, st)
        win.add(launches, dt)
    dt_s = win.wall_s
    value = N * K / dt_s
    print(f"[bench] timed {K} sampled steps in {dt_s:.4f} s", file=sys.stderr, flush=True)
    # The same states stepped by graph replay (graphs captured before the timing).
    sim.set_state_device(*snaps[0])
    sim.step(1)
    sim.synchronize()
    g_s = sum(graph_step_s(sim, st) for st in snaps)
    overflow = sim.hash_overflow()
    knobs = sim.tuning()
    X, V = sim.state()
    sim.close()

    key_bits = max(1, (sp.grid.num_cells - 1).bit_length())
    bpp = dict(BYTES_PER_PARTICLE)
    bpp["sort"] = sort_bytes(key_bits)
    kernels = win.kernels(N, bpp)
    traj_kernels = traj.kernels(N, bpp)
    for name, kr in kernels.items():
        valu, _ = _profiled(INSTR_FILE, name, steps_abs, knobs, "valu")
        if valu is not None:
            kr["valu_issue_frac"] = round(valu / (kr["avg_launch_ms"] * 1e-3) / VALU_PEAK, 4)
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dominant]
    achieved = dk["bytes_per_launch"] / (dk["avg_launch_ms"] * 1e-3)
    traffic, traffic_note = _profiled(TRAFFIC_FILE, dominant, steps_abs, knobs, "bytes")
    valu, valu_note = _profiled(INSTR_FILE, dominant, steps_abs, knobs, "valu")
    valu_roofline = {"bound": "valu", "kernel": dominant, "achieved": None, "frac": None, "note": valu_note}
    if valu is not None:
        rate = valu / (dk["avg_launch_ms"] * 1e-3)
        valu_roofline = {"bound": "valu", "kernel": dominant, "achieved": round(rate, 1), "peak": round(VALU_PEAK, 1),
                         "unit": "wave64 VALU instructions/s", "frac": round(rate / VALU_PEAK, 4),
                         "instructions_per_launch": valu, "avg_launch_ms": dk["avg_launch_ms"],
                         "source": valu_note + "; launch time measured live"}

    out = {
        "metric": "particle-steps/s @1M (2 iters); HBM GB/s vs peak; batched & adjoint/fwd ratio",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(dt_s * 1e3 / K, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": "dam_break_1m", "particles": N, "lattice": [100, 100, 100],
                   "box_m": [2.0, 1.0, 1.0], "r_m": 0.005, "h_over_r": 2.5, "solver_iters": sp.solverIters,
                   "dt": sp.dt, "xsph_c": sp.xsph_c, "grid": a.grid, "table_log2": scene.table_log2,
                   "cells": list(sp.grid.dim), "trajectory_steps": f"{W}..{W + T}",
                   "timed_steps": steps_abs,
                   "timed": f"the step from each of the {K} stratified states of the {T}-step trajectory, eager "
                            "with a HIP event after each launch, synchronised on both sides"},
        "trajectory": {"steps": f"{W}..{W + T}", "ms_per_step": round(traj.wall_s * 1e3 / T, 5),
                       "particle_steps_per_s": round(N * T / traj.wall_s, 1),
                       "value_over_trajectory": round(value / (N * T / traj.wall_s), 4),
                       "kernels": {k: {"avg_launch_ms": v["avg_launch_ms"], "ms_per_step": v["ms_per_step"]}
                                   for k, v in traj_kernels.items()}},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5), "traffic": traffic,
                     "traffic_note": traffic_note, "bytes_per_launch": dk["bytes_per_launch"],
                     "avg_launch_ms": dk["avg_launch_ms"]},
        "valu_roofline": valu_roofline,
        "graph_replay_particle_steps_per_s": round(N * K / g_s, 1),
        "graph_replay_steps": "the same sampled states, one graph launch each (graphs captured beforehand)",
        "tuning": tuning,
        "phase_ms_per_step": {k: round(v / win.steps, 5) for k, v in win.ms.items()},
        "kernels": kernels,
        "hash_overflow": overflow,
        "finite": bool(np.isfinite(X).all() and np.isfinite(V).all()),
    }

    if not a.no_drop_in:
        out["drop_in"] = bench_drop_in(scene, snaps, steps_abs)

    host_states = None
    if not a.no_cpu_baseline:
        pick = sorted({len(snaps) // 6, len(snaps) // 2, (5 * len(snaps)) // 6})
        host_states = [(snaps[i][0].cpu().numpy(), snaps[i][1].cpu().numpy()) for i in pick]
        cpu_labels = [steps_abs[i] for i in pick]
    del snaps

    if not a.no_batched:
        out["batched"] = bench_batched(a.batch_scenes, a.batch_samples, 200, 0 if a.no_tune else a.tune_steps)

    if not a.no_splash:
        out["splash_8m"] = bench_splash(a.splash_samples, W, 0 if (a.no_tune or a.tuning) else a.tune_steps)

    if not a.no_adjoint:
        # the dam break's tuned walk knobs as the start; each rollout then tunes the walk
        # knobs on its own state
        walk = None
        if tuning and "chosen" in tuning:
            walk = {k: tuning["chosen"][k] for k in WALK_KNOBS if k in tuning["chosen"]}
        out["adjoint"] = bench_adjoint(knobs=walk, tune_steps=0 if (a.no_tune or a.tuning) else a.tune_steps)

    if host_states is not None:
        out["cpu_baseline"] = cpu_baseline(scene, sp, host_states, cpu_labels, gpu_dam_break_32k())
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
