"""The autotuner's search and cache on a stand-in simulator (no GPU): it must find the
fastest setting of a separable cost, leave the state as it found it, and cache the
choice."""
import json

from pbfx import tune


class FakeSim:
    """Steps cost time by a separable cost model of the knobs; the state is a counter."""

    def __init__(self):
        self.knobs = {"candidate_mask": True, "stream_ratio": 3, "chunk_bits": True, "sort_items": 0,
                      "sort_digit_bits": 8, "walk_block": 256, "lane_order": False}
        self.t = 0.0
        self.state_value = 123
        self.n = 10
        self.grid = "dense"

        class _P:
            solverIters = 2
        self.params = _P()

    def cost(self):
        k = self.knobs
        return (1.0 + (0.0 if k["candidate_mask"] else 0.3) + {1: 0.0, 2: 0.1, 3: 0.2, 6: 0.5}[k["stream_ratio"]]
                + (0.05 if k["chunk_bits"] else 0.0) + {0: 0.02, 4: 0.0, 8: 0.01, 16: 0.03}[k["sort_items"]]
                + {8: 0.0, 6: 0.01, 4: 0.02}[k["sort_digit_bits"]] + {256: 0.01, 128: 0.0}[k["walk_block"]]
                + (0.0 if k["lane_order"] else 0.04))

    def tuning(self):
        return dict(self.knobs)

    def set_tuning(self, **kw):
        self.knobs.update(kw)

    def step(self, n=1):
        self.t += self.cost() * n * 1e-3
        self.state_value += n

    def synchronize(self):
        pass


class FakeSnapshot:
    def __init__(self, sim):
        self.sim, self.v = sim, sim.state_value

    def restore(self):
        self.sim.state_value = self.v


def test_autotune_finds_the_fastest_setting_and_restores(tmp_path):
    sim = FakeSim()
    path = str(tmp_path / "t.json")
    best, ms, table = tune.autotune(sim, steps=2, key="k", path=path, snapshot=FakeSnapshot, clock=lambda: sim.t)
    assert best == {"candidate_mask": True, "stream_ratio": 1, "chunk_bits": False, "sort_items": 4,
                    "sort_digit_bits": 8, "walk_block": 128, "lane_order": True}
    assert abs(ms - 1.0) < 1e-9
    assert sim.tuning() == best
    assert sim.state_value == 123
    assert len(table) == len({tuple(sorted(t.items())) for t, _ in table})  # each setting timed once
    rec = json.load(open(path))["k"]
    assert rec["tuning"] == best and len(rec["tried"]) == len(table)
    other = FakeSim()
    assert tune.apply_cached(other, "k", path=path) == best and other.tuning() == best
    assert tune.apply_cached(other, "missing", path=path) is None
