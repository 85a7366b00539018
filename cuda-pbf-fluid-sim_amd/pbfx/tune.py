"""Autotuner for the reordered solver's launch knobs (include/pbf_sim.h PbfSimTuning).

The knobs change how the neighbour kernels walk candidates and how the radix sort
tiles its keys; every setting gives the same bits, so the tuner may try them on the
live state.  It snapshots the state, times a few steps per setting from that same
snapshot (coordinate descent over the knobs), restores the snapshot and leaves the
fastest setting applied.  Results are kept in a JSON cache keyed by the workload
(particle count, grid kind, solver iterations, GPU architecture), so a later run of
the same workload can apply the setting without timing anything.
"""
import ctypes
import json
import os
import time

from . import _lib

SPACE = {
    "candidate_mask": (True, False),
    "stream_ratio": (1, 2, 3, 6),
    "chunk_bits": (True, False),
    "sort_items": (0, 4, 8, 16),
    "sort_digit_bits": (8, 6, 4),
    "walk_block": (256, 128),
    "lane_order": (True, False),
}

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "pbfx", "tune.json")


def workload_key(sim, label=""):
    """Cache key: label, particles, grid kind, solver iterations, GPU architecture."""
    try:
        import torch
        arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:
        arch = "unknown"
    p = sim.params
    return f"{label}|n={sim.n}|grid={sim.grid}|iters={p.solverIters}|{arch}"


def load_cache(path=DEFAULT_CACHE):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(cache, path):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply_cached(sim, key, path=DEFAULT_CACHE):
    """Apply the cached setting for `key`; returns it, or None when there is none."""
    rec = load_cache(path).get(key)
    if not rec:
        return None
    sim.set_tuning(**rec["tuning"])
    return rec["tuning"]


class _Snapshot:
    """Device copy of a simulator's positions and velocities (torch tensors)."""

    def __init__(self, sim):
        import torch
        self.sim = sim
        self.X = torch.empty((sim.n, 4), dtype=torch.float32, device="cuda")
        self.V = torch.empty_like(self.X)
        if _lib.lib().pbfSimGetState(sim._h, ctypes.c_void_p(self.X.data_ptr()),
                                     ctypes.c_void_p(self.V.data_ptr()), 1) != 0:
            _lib.check("pbfSimGetState")
            raise _lib.PbfError("snapshot failed")

    def restore(self):
        if _lib.lib().pbfSimSetState(self.sim._h, ctypes.c_void_p(self.X.data_ptr()),
                                     ctypes.c_void_p(self.V.data_ptr()), self.sim.n, 1) != 0:
            _lib.check("pbfSimSetState")
            raise _lib.PbfError("restore failed")


class _StateSnapshot:
    """A caller's device state (X, V) as a snapshot of `sim` (restore() loads it)."""

    def __init__(self, sim, state):
        self.sim, self.X, self.V = sim, state[0], state[1]

    def restore(self):
        self.sim.set_state_device(self.X, self.V)


def autotune(sim, steps=3, sweeps=2, space=None, key=None, path=DEFAULT_CACHE, log=None, snapshot=None,
             clock=time.perf_counter, states=None):
    """Time `steps` steps per setting and keep the fastest.

    The steps start from the current state, or, with `states` (a list of device (X, V)
    tensor pairs, e.g. snapshots along a trajectory whose character changes), from each
    of them in turn, and a setting's time is the sum over them.  Returns (best setting,
    ms per step of it, [(setting, ms per step), ...]).  The simulator's state is the same
    before and after the call.  `snapshot(sim)` makes an object whose restore() puts the
    state back (default: a device copy)."""
    space = space or SPACE
    snap = (snapshot or _Snapshot)(sim)
    starts = [_StateSnapshot(sim, st) for st in states] if states else [snap]
    timed = {}

    def measure(cfg):
        k = tuple(sorted(cfg.items()))
        if k in timed:
            return timed[k]
        sim.set_tuning(**cfg)
        starts[0].restore()
        sim.step(1)              # first step after a change also captures the graph
        sim.synchronize()
        total = 0.0
        for st in starts:
            st.restore()
            sim.synchronize()
            t0 = clock()
            sim.step(steps)
            sim.synchronize()
            total += clock() - t0
        ms = total * 1e3 / (steps * len(starts))
        timed[k] = ms
        if log:
            log(f"[tune] {cfg} {ms:.3f} ms/step")
        return ms

    best = dict(sim.tuning())
    best_ms = measure(best)
    for _ in range(sweeps):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best[knob]:
                    continue
                cfg = dict(best, **{knob: v})
                ms = measure(cfg)
                if ms < best_ms:
                    best, best_ms, changed = cfg, ms, True
        if not changed:
            break
    sim.set_tuning(**best)
    snap.restore()
    table = [(dict(k), ms) for k, ms in timed.items()]
    if key is not None and path:
        cache = load_cache(path)
        cache[key] = {"tuning": best, "ms_per_step": best_ms, "steps": steps,
                      "tried": [{"tuning": t, "ms_per_step": ms} for t, ms in table]}
        _save_cache(cache, path)
    return best, best_ms, table
