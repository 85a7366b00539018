"""Statistics kernels (LaunchComputeStats / LaunchComputeStatsBruteforce, pbfSimDensityStats)
against the oracle and a float64 brute force."""
import numpy as np
import pytest

from gpu_util import grid_of, predicted_state, to_dev
from pbfx import params as P

pytestmark = pytest.mark.gpu


def _brute(X, V, sp, stride=1):
    x = X[:, :3].astype(np.float64)
    sel = np.arange(len(X))[::stride] if stride > 1 else np.arange(len(X))
    nb, sw = [], []
    for i in sel:
        r2 = ((x - x[i]) ** 2).sum(1)
        m = (r2 < sp.kernel.h2)
        m[i] = False
        nb.append(m.sum())
        sw.append((sp.kernel.poly6 * (sp.kernel.h2 - r2[m]) ** 3).sum())
    speed = np.linalg.norm(V[sel, :3].astype(np.float64), axis=1)
    return np.mean(nb), np.mean(speed), np.mean(sw)


@pytest.mark.parametrize("stride", [1, 3])
def test_compute_stats_grid_and_bruteforce(oracle, dev, stride):
    from pbfx import launch as L
    scene = P.small_block((12, 12, 12))
    sp, X, V, Pp = predicted_state(oracle, scene, squeeze=(0.8, 0.1))
    rng = np.random.default_rng(2)
    V = rng.normal(0, 1, V.shape).astype(np.float32)
    g = grid_of(oracle, Pp, sp)
    tP, tV = to_dev(Pp, dev), to_dev(V, dev)
    got = L.compute_stats(tP, tV, to_dev(g["idxS"], dev), to_dev(g["start"], dev), to_dev(g["end"], dev), sp.grid,
                          sp.kernel, sp.particleMass, stride)
    nb, speed, sw = _brute(Pp, V, sp, stride)
    assert got[0] == pytest.approx(nb, rel=1e-6)
    assert got[1] == pytest.approx(speed, rel=1e-6)
    assert got[2] == pytest.approx(sw, rel=1e-5)
    assert got[3] == pytest.approx(sw * sp.particleMass, rel=1e-5)
    bf = L.compute_stats_bruteforce(tP, tV, sp.kernel, sp.particleMass, stride, len(Pp))
    assert bf[0] == pytest.approx(got[0], rel=1e-9)
    assert bf[2] == pytest.approx(got[2], rel=1e-6)


def test_density_stats_matches_oracle(oracle, dev):
    from pbfx.sim import Simulator
    scene = P.dam_break_32k()
    sp = P.build_sim_params(scene)
    sim = Simulator.from_scene(scene)
    sim.step(40)
    X, V = sim.state()
    got = sim.density_stats()
    ref = oracle.density_stats(X, V, sp)
    for k in ("mean_err", "max_err", "mean_neighbors", "mean_speed", "mean_pos_err"):
        assert got[k] == pytest.approx(ref[k], rel=1e-9, abs=1e-12), k
