"""Seeded sweep over solver parameters, grid kinds, step paths and launch knobs: every
configuration matches the oracle bit for bit after five steps.

The sweep crosses settings the targeted tests take one at a time:
- s_corr strength, Δq, exponent (4, 3, 2) and floor (including none): the kernels'
  fixed and general pair forms, and the two-instruction s_corr ratio or the division;
- 1–4 solver iterations, with or without the displacement clamp (the candidate mask
  and its skin follow from both);
- neighbour caps of 64, 32 and none, XSPH on or off, three time steps;
- dense, hashed and compact grids, the reordered and reference paths, graph replay or
  eager steps, and the walk knobs.

Reference step: sim/phases.cpp:108-117 with sim/post_ops.cpp:73-85.
"""
import numpy as np
import pytest

from pbfx import params as P

pytestmark = pytest.mark.gpu

CASES = 48


def _case(rng):
    dims = tuple(int(x) for x in rng.integers(6, 18, 3))
    scene = P.small_block(dims, table_log2=16)
    sp = P.build_sim_params(scene)
    sp.pbf.scorr_k = float(rng.choice([0.0005, 0.003, 0.01, 0.1]))
    sp.pbf.scorr_dq_h = float(rng.choice([0.1, 0.2, 0.3, 0.45, 0.6]))
    sp.pbf.scorr_n = float(rng.choice([4.0, 4.0, 4.0, 3.0, 2.0]))
    sp.pbf.scorr_min = float(rng.choice([-0.25, -0.05, -1.0, 0.0]))
    sp.pbf.relax_omega = float(rng.choice([0.75, 1.0, 0.5]))
    sp.pbf.disp_clamp_max_h = float(rng.choice([0.05, 0.02, 0.1]))
    sp.pbf.enable_disp_clamp = int(rng.choice([1, 1, 0]))
    sp.solverIters = int(rng.choice([1, 2, 3, 4]))
    sp.xsph_c = float(rng.choice([0.01, 0.0, 0.1]))
    sp.dt = float(rng.choice([1 / 240, 1 / 120, 1 / 60]))
    sp.maxNeighbors = int(rng.choice([64, 32, 0]))
    grid = str(rng.choice(["dense", "hashed", "compact"]))
    path = "reference" if (grid == "compact" and rng.random() < 0.5) else "reordered"
    use_graph = bool(rng.random() < 0.5)
    knobs = None
    if rng.random() < 0.5:
        knobs = dict(candidate_mask=bool(rng.random() < 0.7), walk_block=int(rng.choice([128, 256])),
                     lane_order=bool(rng.random() < 0.7))
    return scene, sp, grid, path, use_graph, knobs


def test_parameter_sweep_matches_oracle(oracle, dev):
    from pbfx.sim import Simulator
    rng = np.random.default_rng(20261204)
    for case in range(CASES):
        scene, sp, grid, path, use_graph, knobs = _case(rng)
        sim = Simulator(sp, grid=grid, table_log2=16, path=path, use_graph=use_graph)
        try:
            sim.seed_scene(scene)
            if knobs:
                sim.set_tuning(**knobs)
            sim.step(5)
            X, V = sim.state()
        finally:
            sim.close()
        X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
        eX, eV = oracle.run(X0, np.zeros_like(X0), sp, 5, compact=(grid != "dense"))
        what = (f"case {case}: {scene.num_particles} particles, {grid}/{path}, graph {use_graph}, knobs {knobs}, "
                f"iters {sp.solverIters}, s_corr k {sp.pbf.scorr_k} dq {sp.pbf.scorr_dq_h} n {sp.pbf.scorr_n} "
                f"min {sp.pbf.scorr_min}, clamp {sp.pbf.enable_disp_clamp}, xsph {sp.xsph_c}, cap {sp.maxNeighbors}")
        np.testing.assert_array_equal(X, eX, err_msg=what)
        np.testing.assert_array_equal(V, eV, err_msg=what)
