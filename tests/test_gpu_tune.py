"""Launch knobs and the autotuner: every setting gives the same bits, and tuning
leaves the simulation where it found it."""
import itertools

import numpy as np
import pytest

from pbfx import params as P
from pbfx import tune
from pbfx.sim import Simulator

pytestmark = pytest.mark.gpu


def _collapsing_block():
    # 16^3 block, 40 steps in: dense neighbourhoods, so every walk mode is exercised
    sim = Simulator.from_scene(P.small_block((16, 16, 16)), grid="hashed")
    sim.step(40)
    return sim


def test_every_setting_gives_the_same_bits(dev):
    sim = _collapsing_block()
    try:
        X0, V0 = sim.state()
        ref = None
        for setting in itertools.product(*tune.SPACE.values()):
            knobs = dict(zip(tune.SPACE, setting))
            print(f"setting {knobs}", flush=True)
            sim.set_tuning(**knobs)
            sim.set_state(X0, V0)
            sim.step(3)
            X, V = sim.state()
            if ref is None:
                ref = (X, V)
            assert np.array_equal(X, ref[0]) and np.array_equal(V, ref[1])
    finally:
        sim.close()


def test_autotune_keeps_state_and_caches(dev, tmp_path):
    sim = _collapsing_block()
    twin = _collapsing_block()
    try:
        path = str(tmp_path / "tune.json")
        key = tune.workload_key(sim, "block16")
        best, ms, table = tune.autotune(sim, steps=2, key=key, path=path)
        assert ms > 0 and len(table) >= 2
        assert sim.tuning() == best
        # the tuned simulator continues exactly as the untouched twin
        sim.step(5)
        twin.step(5)
        assert all(np.array_equal(a, b) for a, b in zip(sim.state(), twin.state()))
        # the cache reproduces the choice
        twin.set_tuning(candidate_mask=not best["candidate_mask"])
        assert tune.apply_cached(twin, key, path=path) == best
        assert twin.tuning() == best
        assert tune.apply_cached(twin, "no such workload", path=path) is None
    finally:
        sim.close()
        twin.close()


def test_bad_knobs_are_rejected(dev):
    sim = Simulator.from_scene(P.small_block((4, 4, 4)))
    try:
        with pytest.raises(Exception):
            sim.set_tuning(stream_ratio=0)
        with pytest.raises(Exception):
            sim.set_tuning(sort_items=5)
        with pytest.raises(Exception):
            sim.set_tuning(sort_digit_bits=3)
        with pytest.raises(Exception):
            sim.set_tuning(walk_block=64)
    finally:
        sim.close()
