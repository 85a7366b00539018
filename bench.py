"""Benchmark: PBF dam break, 1,000,000 particles, 2 solver iterations, one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload (the 1M dam-break configuration): 100^3 lattice at spacing 2r
(r = 0.005 m, seeded jitter +-0.1 r) in a 2x1x1 m box, h = 2.5 r, rho0 = 1000,
g = -9.81, dt = 1/120, 2 solver iterations, XSPH c = 0.01, hashed cell grid with a
2^22-slot table, FP32.  A step is one full PBF step (predict, grid build, solve,
velocity, XSPH, boundary); state stays in HBM.  By default the timed steps are the
dam break's steps 2..502 (its 500 configured steps after a 2-step warm-up): about
30 lattice steps, then the collapsed fluid.

Prints ONE JSON line: particle-steps/s over the timed steps (eager, with a HIP
event after every launch on the simulator's stream, so per-kernel durations are
those of the timed launches), the same step replayed from the captured graph, the
HBM roofline of the dominant kernel, the batched ensemble, the adjoint rollout and
the CPU oracle timed on this host on a bounded sample.
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
# counted once per kernel): predict+clamp 36 + hash 20; sort 16 per pass + 4;
# lambda 16; delta+apply 28 (+36 for the fused velocity on the last iteration);
# XSPH + clamp 36.
BYTES_PER_PARTICLE = {
    "predict_hash": 56.0,
    "lambda": 16.0,
    "delta": 28.0,
    "delta_last": 64.0,
    "xsph_scatter": 36.0,
}


# HIP kernels of the launch labels (delta and delta_last are the first and last
# launches of one kernel).
KERNELS = {
    "k_predict_hash": ["predict_hash"],
    "k_lambda_sorted": ["lambda"],
    "k_delta_sorted": ["delta", "delta_last"],
    "k_xsph_scatter": ["xsph_scatter"],
}

TRAFFIC_FILE = os.path.join(REPO, "profiles", "hbm_traffic_dam_break_1m.json")
INSTR_FILE = os.path.join(REPO, "profiles", "instructions_dam_break_1m.json")


# Launch knobs of the neighbour walk, which change what its kernels execute and read;
# profiled counts hold only for the setting they were measured at.  (The sort's knobs
# do not touch these kernels.)
WALK_KNOBS = ("candidate_mask", "stream_ratio", "chunk_bits", "walk_block")


def measured_traffic(kernel, knobs):
    """HBM bytes per launch of `kernel` from the profiler counter summary kept next to
    the other profiles (FETCH_SIZE + WRITE_SIZE, each scaled by a known-size copy) for
    the same walk knobs, or (None, reason)."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, "no counter summary"
    for run in rec.get("runs", []):
        prof = run.get("tuning") or {}
        if all(prof.get(n) == knobs.get(n) for n in WALK_KNOBS):
            k = run.get("kernels", {}).get(kernel)
            if not k:
                return None, "kernel not in the counter summary"
            cmd = rec.get("command", "").replace("<TUNING>", json.dumps(prof))
            return k["bytes_per_launch"], f"{os.path.relpath(TRAFFIC_FILE, REPO)} ({cmd})"
    return None, "no traffic profiled for this run's launch knobs"


def measured_valu(kernel, warmup, steps, knobs):
    """VALU wave-instructions per launch of `kernel` from the profiler's SQ_INSTS_VALU
    summary kept next to the other profiles, or (None, reason).  The counts hold only
    for the same timed steps and the same walk knobs as the profiled run."""
    try:
        with open(INSTR_FILE) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None, "no counter summary"
    same_window = (rec.get("warmup"), rec.get("steps")) == (warmup, steps)
    for run in rec.get("runs", []):
        prof = run.get("tuning") or {}
        if not all(prof.get(n) == knobs.get(n) for n in WALK_KNOBS):
            continue
        cmd = rec.get("command", "").replace("<TUNING>", json.dumps(prof))
        src = f"{os.path.relpath(INSTR_FILE, REPO)} ({cmd})"
        k = run.get("kernels", {}).get(kernel)
        if same_window and k:
            return k["valu_per_launch"], src
        # the simulation is deterministic, so another window of the same run has the
        # per-step counts the profiled run recorded for it
        ps = (run.get("per_step_valu") or {}).get(kernel)
        if ps and ps["first_step"] <= warmup and warmup + steps <= ps["first_step"] + len(ps["valu"]):
            win = ps["valu"][warmup - ps["first_step"]:warmup - ps["first_step"] + steps]
            return sum(win) / (steps * ps["launches_per_step"]), f"{src}, per-step counts of steps " \
                f"{warmup}..{warmup + steps}"
        return None, (f"instruction counts were profiled for steps {rec.get('warmup')}.."
                      f"{(rec.get('warmup') or 0) + (rec.get('steps') or 0)}; this run times other steps")
    return None, "no instruction counts profiled for this run's launch knobs"


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


def bench_drop_in(scene, warmup, steps):
    """The same workload through the drop-in launchers (LaunchLambda / LaunchDeltaApply /
    LaunchXSPH over the reference's layout, dense cell ranges), as the reference's own
    phase pipeline would call them after swapping in this library: graph replay of the
    same steps as the timed region."""
    from pbfx.sim import Simulator
    sim = Simulator.from_scene(scene, grid="dense", path="reference")
    try:
        sim.step(warmup)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.synchronize()
        dt_s = time.perf_counter() - t0
        n = sim.n
    finally:
        sim.close()
    print(f"[bench] drop-in launchers: {steps} steps in {dt_s:.4f} s", file=sys.stderr, flush=True)
    return {"particle_steps_per_s": round(n * steps / dt_s, 1), "ms_per_step": round(dt_s * 1e3 / steps, 5),
            "steps": f"{warmup}..{warmup + steps}", "grid": "dense", "path": "reference layout, drop-in launchers"}


def bench_batched(count, warmup, steps, tune_steps=0, grid="dense"):
    """The batched ensemble (one launch per phase for all scenes, graph replay) against the
    same scenes as separate plain simulators, each replaying its own captured step:
    - scene by scene: one scene after another, `steps` steps each, one synchronisation per
      scene (the plain path run over the scenes);
    - concurrent streams: every scene's steps enqueued on its own stream before a single
      synchronisation (the strongest unbatched arrangement)."""
    from pbfx.ensemble import Ensemble, ensemble_scenes, solo
    scenes = ensemble_scenes(count, seed=2024)
    total = sum(sc.num_particles for sc in scenes)
    ens = Ensemble(scenes, grid=grid)
    ens.step(warmup)
    ens.synchronize()
    best = None
    if tune_steps:
        from pbfx import tune
        best, _, _ = tune.autotune(ens, steps=tune_steps, key=tune.workload_key(ens, f"ensemble{count}"))
    t0 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    t_ens = time.perf_counter() - t0
    ens.close()
    print(f"[bench] ensemble {count} scenes / {total} particles: {steps} steps in {t_ens:.4f} s",
          file=sys.stderr, flush=True)
    sims = [solo(sc, grid=grid) for sc in scenes]
    for sim in sims:
        if best:
            sim.set_tuning(**best)
        sim.step(warmup)
    for sim in sims:
        sim.synchronize()
    snaps = [sim.state() for sim in sims]
    t0 = time.perf_counter()
    for sim in sims:
        sim.step(steps)
        sim.synchronize()
    t_seq = time.perf_counter() - t0
    for sim, (X, V) in zip(sims, snaps):
        sim.set_state(X, V)
    for sim in sims:
        sim.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for sim in sims:
            sim.step(1)
    for sim in sims:
        sim.synchronize()
    t_streams = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    print(f"[bench] scene by scene: {t_seq:.4f} s; concurrent streams: {t_streams:.4f} s", file=sys.stderr,
          flush=True)
    return {"scenes": count, "particles": total, "steps": steps, "after_warmup_steps": warmup, "grid": grid,
            "particle_steps_per_s": round(total * steps / t_ens, 1),
            "scene_by_scene_particle_steps_per_s": round(total * steps / t_seq, 1),
            "speedup_vs_scene_by_scene": round(t_seq / t_ens, 3),
            "concurrent_streams_particle_steps_per_s": round(total * steps / t_streams, 1),
            "speedup_vs_concurrent_streams": round(t_streams / t_ens, 3), "tuning": best}


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
    target velocity field) advanced and differentiated together as one ensemble.
    knobs: launch knobs to apply first (the neighbour walk's), or None for the defaults;
    tune_steps: steps per setting when the walk knobs are then tuned on the ensemble's
    state after the target rollout (0: keep them)."""
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
        roll.backward(xT.grad.float())
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
    finally:
        roll.close()
    print(f"[bench] {rollouts} batched rollouts, {n} particles: forward {t_fwd:.4f} s, backward {t_bwd:.4f} s",
          file=sys.stderr, flush=True)
    return {"rollouts": rollouts, "particles": n, "knobs": chosen or "defaults",
            "forward_ms_per_step": round(t_fwd * 1e3 / steps, 4),
            "backward_over_forward": round(t_bwd / t_fwd, 3),
            "rollout_steps_per_s": round(rollouts * steps / (t_fwd + t_bwd), 1)}


def bench_adjoint(steps=60, checkpoint=10, rollouts=64, knobs=None, tune_steps=0):
    """The differentiable rollout: 65,536 particles (64x32x32 lattice), 60 steps, checkpoint every
    10.  Loss: squared L2 distance between the 32^3 Gaussian occupancy grid (sigma = h) of X_T and
    that of a target rollout from a seeded perturbed initial velocity.  Times the checkpointed
    forward, then the reverse pass split into re-recording the segments and the reverse kernels.
    knobs: the neighbour walk's launch knobs to apply to forward and reverse, or None;
    tune_steps: steps per setting when the walk knobs are then tuned on the rollout's own
    state after the target rollout (0: keep them)."""
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
        _, gV0, gc = roll.backward(xT.grad.float())
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0
        tm = dict(roll.timing)
        gnorm = float(gV0[:, :3].double().norm())
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
            "grad_v0_norm": gnorm, "grad_c": gc, "grad_finite": bool(math.isfinite(gnorm) and math.isfinite(gc)),
            "note": "forward = eager checkpointed rollout; backward = re-record every segment + reverse kernels. "
                    "The dynamics amplify perturbations ~8x per step (float64 reference), so the 60-step "
                    "gradient exceeds the float32 range; the timing is what this section measures"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tune-at", type=int, default=60,
                    help="step at which the autotuner times the launch knobs (the dam break has collapsed "
                         "by then); the state is then restored to the end of the warm-up")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch knobs")
    ap.add_argument("--tune-steps", type=int, default=5)
    ap.add_argument("--tuning", default=None,
                    help='fixed launch knobs as JSON (e.g. {"candidate_mask": true, "stream_ratio": 1}); '
                         'skips the autotuner, as profiler runs do to time one setting')
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--grid", default="hashed", choices=["hashed", "compact", "dense"])
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-drop-in", action="store_true", help="skip timing the drop-in launchers' path")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=30)
    ap.add_argument("--batch-warmup", type=int, default=20)
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
    sim = Simulator.from_scene(scene, grid=a.grid)
    print(f"[bench] seeded {N} particles, grid={a.grid}, warmup {a.warmup}", file=sys.stderr, flush=True)
    sim.step(a.warmup)
    sim.synchronize()
    tuning = None
    if a.tuning:
        sim.set_tuning(**json.loads(a.tuning))
        tuning = {"chosen": sim.tuning(), "fixed": True}
    elif not a.no_tune:
        from pbfx import tune
        X0, V0 = sim.state()
        ahead = max(0, a.tune_at - a.warmup)
        sim.step(ahead)
        best, best_ms, table = tune.autotune(sim, steps=a.tune_steps, key=tune.workload_key(sim, "dam_break_1m"),
                                             log=lambda m: print(m, file=sys.stderr, flush=True))
        sim.set_state(X0, V0)  # back to the end of the warm-up: the knobs never change results
        tuning = {"chosen": best, "ms_per_step_at_tuning": round(best_ms, 4), "settings_tried": len(table),
                  "steps_per_setting": a.tune_steps, "state": f"step {a.warmup + ahead}, then restored to "
                                                                f"step {a.warmup}"}
    # Timed region: K eager steps on the simulator's stream with a HIP event after every
    # launch, so the per-kernel durations are those of exactly the timed launches.
    agg, cnt = {}, {}
    t0 = time.perf_counter()
    for _ in range(a.steps):
        for label, ms in sim.profile_launches():
            agg[label] = agg.get(label, 0.0) + ms
            cnt[label] = cnt.get(label, 0) + 1
    sim.synchronize()
    dt_s = time.perf_counter() - t0
    value = N * a.steps / dt_s
    print(f"[bench] timed {a.steps} steps in {dt_s:.4f} s", file=sys.stderr, flush=True)
    # the same step replayed from the captured graph, for comparison
    gsteps = max(1, a.steps // 5)
    t0 = time.perf_counter()
    sim.step(gsteps)
    sim.synchronize()
    graph_value = N * gsteps / (time.perf_counter() - t0)
    per_step = {k: v / a.steps for k, v in agg.items()}
    overflow = sim.hash_overflow()
    sim_knobs = sim.tuning()
    X, V = sim.state()
    sim.close()

    key_bits = max(1, (sp.grid.num_cells - 1).bit_length())
    bpp = dict(BYTES_PER_PARTICLE)
    bpp["sort"] = sort_bytes(key_bits)
    kernels = {}
    for name, labels in KERNELS.items():
        launches = sum(cnt.get(l, 0) for l in labels)
        if not launches:
            continue
        ms = sum(agg.get(l, 0.0) for l in labels)
        nbytes = sum(bpp[l] * N * cnt.get(l, 0) for l in labels)
        kernels[name] = {"labels": labels, "launches_per_step": launches / a.steps,
                         "avg_launch_ms": round(ms / launches, 5), "ms_per_step": round(ms / a.steps, 5),
                         "bytes_per_launch": nbytes / launches, "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1),
                         "frac": round(nbytes / (ms * 1e-3) / HBM_PEAK, 5)}
    knobs = sim_knobs
    for name, kr in kernels.items():
        valu, _ = measured_valu(name, a.warmup, a.steps, knobs)
        if valu is not None:
            kr["valu_issue_frac"] = round(valu / (kr["avg_launch_ms"] * 1e-3) / VALU_PEAK, 4)
    dominant = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    dk = kernels[dominant]
    achieved = dk["bytes_per_launch"] / (dk["avg_launch_ms"] * 1e-3)
    traffic, traffic_note = measured_traffic(dominant, knobs)
    valu, valu_note = measured_valu(dominant, a.warmup, a.steps, knobs)
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
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt_s * 1e3 / a.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered lattice)",
        "config": {"workload": "dam_break_1m", "particles": N, "lattice": [100, 100, 100],
                   "box_m": [2.0, 1.0, 1.0], "r_m": 0.005, "h_over_r": 2.5, "solver_iters": sp.solverIters,
                   "dt": sp.dt, "xsph_c": sp.xsph_c, "grid": a.grid, "table_log2": scene.table_log2,
                   "cells": list(sp.grid.dim),
                   "timed_steps": "eager launches with a HIP event after each (graph replay reported beside)"},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5), "traffic": traffic,
                     "traffic_note": traffic_note, "bytes_per_launch": dk["bytes_per_launch"],
                     "avg_launch_ms": dk["avg_launch_ms"]},
        "valu_roofline": valu_roofline,
        "graph_replay_particle_steps_per_s": round(graph_value, 1),
        "graph_replay_steps": f"{a.warmup + a.steps}..{a.warmup + a.steps + gsteps}",
        "tuning": tuning,
        "phase_ms_per_step": {k: round(v, 5) for k, v in per_step.items()},
        "kernels": kernels,
        "hash_overflow": overflow,
        "finite": bool((X == X).all() and (V == V).all()),
    }

    if not a.no_drop_in:
        out["drop_in"] = bench_drop_in(scene, a.warmup, a.steps)

    if not a.no_batched:
        out["batched"] = bench_batched(a.batch_scenes, a.batch_warmup, a.batch_steps,
                                       0 if a.no_tune else a.tune_steps)

    if not a.no_adjoint:
        # the dam break's tuned walk knobs as the start; each rollout then tunes the walk
        # knobs on its own state
        walk = None
        if tuning and "chosen" in tuning:
            walk = {k: tuning["chosen"][k] for k in ("candidate_mask", "stream_ratio", "chunk_bits", "walk_block")
                    if k in tuning["chosen"]}
        out["adjoint"] = bench_adjoint(knobs=walk, tune_steps=0 if (a.no_tune or a.tuning) else a.tune_steps)

    if not a.no_cpu_baseline:
        from oracle import oracle as O
        Xc = P.seed_scene_reference(scene, O.seed_lattice)
        Vc = np.zeros_like(Xc)
        Xc, Vc = O.run(Xc, Vc, sp, 1, compact=True)  # warm-up step
        cores = O.num_threads()
        t0 = time.perf_counter()
        O.run(Xc, Vc, sp, a.cpu_steps, compact=True)
        cpu_s = time.perf_counter() - t0
        O.set_num_threads(1)
        t0 = time.perf_counter()
        O.run(Xc, Vc, sp, 1, compact=True)
        cpu1_s = time.perf_counter() - t0
        O.set_num_threads(cores)
        out["cpu_baseline"] = {"value": round(N * a.cpu_steps / cpu_s, 1), "unit": "particle-steps/s",
                               "cores": cores, "kind": "port",
                               "sample": f"dam_break_1m, compact grid, steps 2..{1 + a.cpu_steps} after one "
                                         f"warm-up step, {cpu_s:.2f} s, OpenMP over particles",
                               "single_thread_value": round(N / cpu1_s, 1),
                               "single_thread_sample": f"step 2 on one thread, {cpu1_s:.2f} s",
                               "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
