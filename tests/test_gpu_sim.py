"""Whole-step parity of the headless simulator (both pipelines, all grid kinds) with the oracle.

One step: positions within 1e-5*h and velocities within 1e-5*h/dt of the oracle.
Many steps: the dam break is chaotic (SI-unit PBF with a saturating displacement
clamp), so after 200 steps the density-constraint statistics must agree within 1%.
At full size the default run compares the 1M dam break (3 steps, hashed grid with the
candidate mask, the benchmark's setting) and the 8M splash (1 step, 2^24-slot table) with
the oracle bit for bit, checks the three grid kinds against each other and against
size-independent invariants; with PBF_LONG_PARITY=1 a 200-step 1M run is also compared
with the oracle bit for bit (minutes of oracle time).
"""
import os

import numpy as np
import pytest

from gpu_util import H_TOL, max_abs
from pbfx import params as P

pytestmark = pytest.mark.gpu


def _oracle_state(oracle, scene, steps, compact=False):
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = np.zeros_like(X)
    return oracle.run(X, V, sp, steps, compact=compact)


def _sim(scene, grid, path="reordered", use_graph=True):
    from pbfx.sim import Simulator
    return Simulator.from_scene(scene, path=path, use_graph=use_graph, grid=grid)


def test_seed_matches_oracle(oracle, dev):
    scene = P.dam_break_32k()
    sim = _sim(scene, "dense")
    X, V = sim.state()
    np.testing.assert_array_equal(X, P.seed_scene_reference(scene, oracle.seed_lattice))
    assert np.all(V == 0)


@pytest.mark.parametrize("path", ["reordered", "reference"])
@pytest.mark.parametrize("grid", ["dense", "compact", "hashed"])
def test_one_step_32k(oracle, dev, grid, path):
    if path == "reference" and grid == "hashed":
        pytest.skip("the hash table is a reordered-pipeline structure")
    scene = P.dam_break_32k(table_log2=20)
    sp = P.build_sim_params(scene)
    sim = _sim(scene, grid, path)
    sim.step(1)
    X, V = sim.state()
    eX, eV = _oracle_state(oracle, scene, 1, compact=(grid != "dense"))
    h = sp.kernel.h
    assert max_abs(X[:, :3], eX[:, :3]) <= H_TOL * h
    assert max_abs(V[:, :3], eV[:, :3]) <= H_TOL * h / sp.dt
    np.testing.assert_array_equal(X[:, 3], eX[:, 3])
    np.testing.assert_array_equal(V[:, 3], eV[:, 3])


@pytest.mark.parametrize("grid", ["dense", "hashed"])
def test_ten_steps_small_block(oracle, dev, grid):
    scene = P.small_block((16, 16, 16), table_log2=16)
    sp = P.build_sim_params(scene)
    sim = _sim(scene, grid)
    sim.step(10)
    X, V = sim.state()
    eX, eV = _oracle_state(oracle, scene, 10)
    # kernels and oracle share every rounding step (no FMA, IEEE div/sqrt, the same s_corr power),
    # so whole trajectories agree bit for bit
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 1, 1), (7, 3, 3), (13, 5, 1), (13, 11, 7)])
@pytest.mark.parametrize("path,grid", [("reordered", "hashed"), ("reordered", "dense"), ("reference", "compact")])
def test_particle_counts_off_the_wave_size(oracle, dev, dims, path, grid):
    # 1, 2, 63, 65 and 1001 particles: the last wave (and block) of every kernel is partly empty
    scene = P.small_block(dims, table_log2=14)
    sim = _sim(scene, grid, path)
    try:
        sim.step(4)
        X, V = sim.state()
    finally:
        sim.close()
    eX, eV = _oracle_state(oracle, scene, 4, compact=(grid != "dense"))
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def _fine_cells(sp, scene, frac=0.6):
    """Cells of frac * h: the neighbourhood then reaches two cells each way (25 rows)."""
    cell = np.float32(frac) * np.float32(sp.kernel.h)
    sp.grid.cellSize = float(cell)
    sp.grid.dim = P.Int3(*P.compute_dims(scene.box_min, scene.box_max, cell))
    return sp


@pytest.mark.parametrize("path,grid", [("reordered", "dense"), ("reordered", "hashed"), ("reference", "dense"),
                                       ("reference", "compact")])
def test_cells_smaller_than_h(oracle, dev, path, grid):
    from pbfx.sim import Simulator
    scene = P.small_block((12, 12, 12), table_log2=16)
    sp = _fine_cells(P.build_sim_params(scene), scene)
    sim = Simulator(sp, grid=grid, table_log2=scene.table_log2, path=path)
    try:
        sim.seed_scene(scene)
        sim.step(3)
        X, V = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    eX, eV = oracle.run(X0, np.zeros_like(X0), sp, 3, compact=(grid != "dense"))
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def test_graph_replay_equals_eager(dev):
    scene = P.small_block((16, 16, 16), table_log2=16)
    a = _sim(scene, "hashed", use_graph=True)
    b = _sim(scene, "hashed", use_graph=False)
    for _ in range(3):
        a.step(5)
        b.step(5)
        Xa, Va = a.state()
        Xb, Vb = b.state()
        np.testing.assert_array_equal(Xa, Xb)
        np.testing.assert_array_equal(Va, Vb)


@pytest.mark.timeout(300)
def test_restored_state_steps_alike_eager_and_graph(dev):
    """A state restored from device tensors (set_state_device) is in place before the
    next step starts: graph replays and eager steps from it give the same bits, every
    time.  The copies run on the simulator's stream: a blocking device-to-device copy
    on the null stream may return before it lands, and the simulator's stream does not
    wait for the null stream, so a graph launched right after it read half-copied
    positions.  The 1M dam break after 60 steps (collapsed, every step distinct)."""
    scene = P.dam_break_1m()
    graph = _sim(scene, "hashed", use_graph=True)
    graph.step(60)
    st = graph.state_device()
    st = (st[0].clone(), st[1].clone())
    eager = _sim(scene, "hashed", use_graph=False)
    eager.set_state_device(*st)
    eager.step(1)
    Xe, Ve = eager.state()
    eager.close()
    for _ in range(3):
        graph.set_state_device(*st)
        graph.step(1)
        Xg, Vg = graph.state()
        np.testing.assert_array_equal(Xg, Xe)
        np.testing.assert_array_equal(Vg, Ve)
        graph.set_state_device(*st)
        graph.profile_launches()
        Xp, Vp = graph.state()
        np.testing.assert_array_equal(Xp, Xe)
        np.testing.assert_array_equal(Vp, Ve)
    graph.close()


def test_paths_agree_bitwise(dev):
    scene = P.small_block((20, 20, 20))
    a = _sim(scene, "dense", "reordered")
    b = _sim(scene, "dense", "reference")
    a.step(3)
    b.step(3)
    Xa, Va = a.state()
    Xb, Vb = b.state()
    np.testing.assert_array_equal(Xa, Xb)
    np.testing.assert_array_equal(Va, Vb)


def test_density_statistics_200_steps(oracle, dev):
    scene = P.dam_break_32k()
    sp = P.build_sim_params(scene)
    sim = _sim(scene, "dense")
    sim.step(200)
    X, V = sim.state()
    eX, eV = _oracle_state(oracle, scene, 200)
    g = oracle.density_stats(X, V, sp)
    e = oracle.density_stats(eX, eV, sp)
    print("gpu", g, "\noracle", e, "\nmax |dx|", max_abs(X[:, :3], eX[:, :3]))
    for k in ("mean_err", "max_err", "mean_neighbors", "mean_pos_err"):
        assert g[k] == pytest.approx(e[k], rel=0.01), k
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(V))
    # stronger than the statistics criterion: the collapsed 200-step state is bit-identical
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def test_1m_grid_kinds_identical_and_invariants(dev):
    """dam_break_1m at full size: dense, compact and hashed lookups give identical bits."""
    scene = P.dam_break_1m()
    sp = P.build_sim_params(scene)
    states = {}
    for grid in ("hashed", "compact", "dense"):
        sim = _sim(scene, grid)
        sim.step(3)
        states[grid] = sim.state()
        if grid == "hashed":
            assert not sim.hash_overflow()
        sim.close()
    for grid in ("compact", "dense"):
        np.testing.assert_array_equal(states[grid][0], states["hashed"][0])
        np.testing.assert_array_equal(states[grid][1], states["hashed"][1])
    X, V = states["hashed"]
    lo = np.array(list(sp.grid.mins), np.float32)
    hi = np.array(list(sp.grid.maxs), np.float32)
    assert np.all(np.isfinite(X)) and np.all(X[:, :3] >= lo) and np.all(X[:, :3] <= hi)
    assert np.all(X[:, 3] == 1.0) and np.all(V[:, 3] == 0.0)


@pytest.mark.timeout(300)
def test_1m_three_steps_bit_identical_to_oracle(oracle, dev):
    """dam_break_1m at full size, 3 steps on the hashed grid with the candidate mask (the
    benchmark's setting) against the oracle's step (reference: phases.cpp:108-117), bit
    for bit after every step."""
    scene = P.dam_break_1m()
    sp = P.build_sim_params(scene)
    sim = _sim(scene, "hashed")
    try:
        sim.set_tuning(candidate_mask=True, stream_ratio=1, chunk_bits=False)
        X = P.seed_scene_reference(scene, oracle.seed_lattice)
        V = np.zeros_like(X)
        for t in range(1, 4):
            sim.step(1)
            X, V = oracle.run(X, V, sp, 1, compact=True)
            gX, gV = sim.state()
            print(f"1M step {t}: bit-identical {np.array_equal(gX, X) and np.array_equal(gV, V)}", flush=True)
            np.testing.assert_array_equal(gX, X)
            np.testing.assert_array_equal(gV, V)
        assert not sim.hash_overflow()
    finally:
        sim.close()


@pytest.mark.timeout(600)
def test_8m_one_step_bit_identical_to_oracle(oracle, dev):
    """splash_8m at full size (8,000,000 particles, 4 solver iterations, the 2^24-slot table
    smaller than its 32M-cell key space, so block-hashed homes): one step equals the
    oracle's, bit for bit."""
    scene = P.splash_8m()
    sp = P.build_sim_params(scene)
    sim = _sim(scene, "hashed")
    try:
        sim.step(1)
        gX, gV = sim.state()
        assert not sim.hash_overflow()
    finally:
        sim.close()
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    eX, eV = oracle.run(X, np.zeros_like(X), sp, 1, compact=True)
    print(f"8M step 1: bit-identical {np.array_equal(gX, eX) and np.array_equal(gV, eV)}", flush=True)
    np.testing.assert_array_equal(gX, eX)
    np.testing.assert_array_equal(gV, eV)


@pytest.mark.timeout(1200)
@pytest.mark.skipif(os.environ.get("PBF_LONG_PARITY") != "1",
                    reason="about 3 minutes of multi-threaded oracle time; set PBF_LONG_PARITY=1")
def test_1m_200_steps_bit_identical_to_oracle(oracle, dev):
    """dam_break_1m at full size, 200 steps on the hashed grid with the candidate mask (the
    benchmark's setting): bit-identical to the oracle after every 20 steps."""
    scene = P.dam_break_1m()
    sp = P.build_sim_params(scene)
    sim = _sim(scene, "hashed")
    sim.set_tuning(candidate_mask=True, stream_ratio=1, chunk_bits=False)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = np.zeros_like(X)
    for done in range(20, 201, 20):
        sim.step(20)
        X, V = oracle.run(X, V, sp, 20, compact=True)
        gX, gV = sim.state()
        print(f"step {done}: bit-identical {np.array_equal(gX, X) and np.array_equal(gV, V)}", flush=True)
        np.testing.assert_array_equal(gX, X)
        np.testing.assert_array_equal(gV, V)
    assert not sim.hash_overflow()
    sim.close()


def test_8m_splash_hashed_equals_compact(dev):
    """splash_8m at full size (8M particles, 4 iterations, 2^24-slot table): the hashed table and
    the compact binary search give identical bits; the state stays in the box and finite."""
    scene = P.splash_8m()
    sp = P.build_sim_params(scene)
    states = {}
    for grid in ("hashed", "compact"):
        sim = _sim(scene, grid)
        sim.step(2)
        states[grid] = sim.state()
        if grid == "hashed":
            assert not sim.hash_overflow()
        sim.close()
    np.testing.assert_array_equal(states["hashed"][0], states["compact"][0])
    np.testing.assert_array_equal(states["hashed"][1], states["compact"][1])
    X, V = states["hashed"]
    lo = np.array(list(sp.grid.mins), np.float32)
    hi = np.array(list(sp.grid.maxs), np.float32)
    assert np.all(np.isfinite(X)) and np.all(X[:, :3] >= lo) and np.all(X[:, :3] <= hi)


def _changed_params(scene):
    """Dynamic changes, which ParamChangeTracker (param_change_tracker.h:46-56) passes
    through to the captured step: step size, gravity, XSPH, s_corr, the displacement
    clamp (which also widens the candidate mask's skin), and the neighbour cap (which
    the reference counts as structural, param_change_tracker.h:39, a re-capture either
    way)."""
    sp = P.build_sim_params(scene)
    sp.dt = sp.dt * 0.75
    sp.gravity.x = 2.0
    sp.xsph_c = 0.03
    sp.pbf.scorr_k = 0.005
    sp.pbf.disp_clamp_max_h = 0.08
    sp.maxNeighbors = 48
    return sp


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("path,grid", [("reordered", "hashed"), ("reordered", "dense"), ("reference", "compact")])
def test_set_params_between_steps_matches_oracle(oracle, dev, use_graph, path, grid):
    scene = P.small_block((14, 14, 14), table_log2=16)
    spA = P.build_sim_params(scene)
    spB = _changed_params(scene)
    sim = _sim(scene, grid, path=path, use_graph=use_graph)
    try:
        sim.step(3)
        sim.set_params(spB)
        sim.step(3)
        X, V = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    compact = grid != "dense"
    eX, eV = oracle.run(X0, np.zeros_like(X0), spA, 3, compact=compact)
    eX, eV = oracle.run(eX, eV, spB, 3, compact=compact)
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def _structural(scene, kind):
    """A structural change of the reference's kind (param_change_tracker.h:37-44 and the
    re-allocation of simulator.cpp:583-597) and the scene it describes."""
    if kind == "iters":
        sp = P.build_sim_params(scene)
        sp.solverIters = 3
        return scene, sp
    if kind == "box":
        sc = P.small_block((8, 8, 8), box_max=(1.25, 1.0, 0.75), table_log2=16)
        return sc, P.build_sim_params(sc)
    if kind == "h":
        sc = P.small_block((8, 8, 8), h_over_r=2.2, table_log2=16)
        return sc, P.build_sim_params(sc)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["iters", "box", "h"])
@pytest.mark.parametrize("path,grid", [("reordered", "hashed"), ("reordered", "dense"), ("reference", "compact")])
def test_set_params_structural_changes_match_oracle(oracle, dev, kind, path, grid):
    """Solver iterations, the box (hence the grid dims) and the smoothing length (hence
    kernel constants, mass and cell size) changed between steps: the simulator
    re-allocates what depends on them, keeps the particles' state and re-captures, and
    equals the oracle run with the old then the new parameters, bit for bit."""
    scene = P.small_block((8, 8, 8), table_log2=16)
    spA = P.build_sim_params(scene)
    _, spB = _structural(scene, kind)
    sim = _sim(scene, grid, path=path)
    try:
        sim.step(2)
        sim.set_params(spB)
        sim.step(2)
        X, V = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    compact = grid != "dense"
    eX, eV = oracle.run(X0, np.zeros_like(X0), spA, 2, compact=compact)
    eX, eV = oracle.run(eX, eV, spB, 2, compact=compact)
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


@pytest.mark.parametrize("path,grid", [("reordered", "hashed"), ("reference", "compact")])
def test_set_params_more_particles(oracle, dev, path, grid):
    """A particle count beyond the buffers (reference: simulator.cpp:583-597 re-allocates):
    the first 512 particles keep their state, a second block is seeded at offset 512, and
    the run equals the oracle's on the concatenated state."""
    from pbfx import _lib
    scene = P.small_block((8, 8, 8), table_log2=16)
    spA = P.build_sim_params(scene)
    spB = P.build_sim_params(scene)
    spB.numParticles = spB.maxParticles = 1024
    sim = _sim(scene, grid, path=path)
    o2 = (0.4, 0.005, 0.005)
    amp = float(np.float32(np.float32(scene.jitter_frac_r) * np.float32(scene.radius)))
    try:
        sim.step(2)
        sim.set_params(spB)
        assert sim.n == 1024
        assert _lib.lib().pbfSimSeedLattice(sim._h, 512, 8, 8, 8, o2[0], o2[1], o2[2], scene.spacing, amp, 99) == 0
        sim.step(2)
        X, V = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    compact = grid != "dense"
    eX, eV = oracle.run(X0, np.zeros_like(X0), spA, 2, compact=compact)
    B = oracle.seed_lattice(8, 8, 8, o2, scene.spacing, spB.grid, amp, 99)
    eX = np.concatenate([eX, B])
    eV = np.concatenate([eV, np.zeros_like(B)])
    eX, eV = oracle.run(eX, eV, spB, 2, compact=compact)
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def test_set_params_structural_refusals(oracle, dev):
    """Refused with an error, the simulator stepping on with its own parameters: a
    structural change while adjoint tapes are held, and more particles than bound
    external position buffers hold."""
    import torch
    from pbfx import _lib
    scene = P.small_block((8, 8, 8), table_log2=16)
    sim = _sim(scene, "hashed")
    try:
        assert _lib.lib().pbfSimAdjointInit(sim._h, 1) == 0
        sp = P.build_sim_params(scene)
        sp.solverIters = 3
        with pytest.raises(_lib.PbfError):
            sim.set_params(sp)
    finally:
        sim.close()
    sim = _sim(scene, "hashed")
    try:
        n = sim.n
        ta = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        tb = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        sim.bind_external_positions(ta, tb)
        sp = P.build_sim_params(scene)
        sp.numParticles = sp.maxParticles = n + 1
        with pytest.raises(_lib.PbfError):
            sim.set_params(sp)
        sim.step(1)
        X1, _ = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    eX, _ = oracle.run(X0, np.zeros_like(X0), P.build_sim_params(scene), 1, compact=True)
    np.testing.assert_array_equal(X1, eX)
    del ta, tb


@pytest.mark.parametrize("grid,path,use_graph", [("dense", "reordered", True), ("hashed", "reordered", False),
                                                 ("compact", "reference", True)])
def test_non_finite_velocity_steps_like_oracle(oracle, dev, grid, path, use_graph):
    """A particle with a NaN velocity and one with a velocity of 3e38 (clamped to the box),
    in the 32,768-particle dam break at step 20: two steps give the oracle's bits, NaN where
    the oracle has NaN (the NaN spreads to a few hundred particles: the reference's range
    test, a skip when r2 > h2, admits a NaN distance).  On the reordered path the step's
    velocities then include NaN, so XSPH takes its general form, and positions that turn
    NaN in the first iteration make the second walk every candidate instead of the mask."""
    scene = P.dam_break_32k()
    sp = P.build_sim_params(scene)
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    X, V = oracle.run(X0, np.zeros_like(X0), sp, 20)
    V[100, :3] = [np.nan, 0.0, 0.0]
    V[5000, :3] = [3e38, -3e38, 3e38]
    sim = _sim(scene, grid, path=path, use_graph=use_graph)
    try:
        sim.set_state(X, V)
        sim.step(2)
        Xg, Vg = sim.state()
    finally:
        sim.close()
    eX, eV = oracle.run(X, V, sp, 2, compact=(grid != "dense"))
    assert np.isnan(eX[100, 0]) and np.isnan(eV).any()
    np.testing.assert_array_equal(Xg, eX)
    np.testing.assert_array_equal(Vg, eV)


@pytest.mark.parametrize("grid,path", [("dense", "reordered"), ("hashed", "reordered"), ("dense", "reference")])
def test_cell_size_far_below_h_is_refused(dev, grid, path):
    """A cell size 5000 times below h needs (2 * 5000 + 1)^2 neighbour rows a particle: the
    span table would hold more than 2^32 entries, so creation fails with a capacity error
    (no overflowed row count, no undersized table)."""
    from pbfx import _lib
    from pbfx.sim import Simulator
    scene = P.small_block((8, 8, 8), table_log2=16)
    sp = P.build_sim_params(scene)
    sp.kernel.h = sp.grid.cellSize * 5000.0
    with pytest.raises(_lib.PbfError, match="error 1003"):
        Simulator(sp, grid=grid, table_log2=16, path=path)
    # the library stays usable
    sim = _sim(scene, grid, path=path)
    try:
        sim.step(1)
    finally:
        sim.close()


@pytest.mark.parametrize("dims", [(16, 16, 16), (13, 11, 7)])
@pytest.mark.parametrize("path,grid", [("reordered", "hashed"), ("reordered", "dense"), ("reference", "compact")])
def test_simulator_grid_build_matches_oracle(oracle, dev, dims, path, grid):
    """The simulator's own grid build (pbfSimDeviceSortedKeys / Indices after a step):
    the cell keys of the predicted, clamped positions, stably sorted with their particle
    indices (phases.cpp:33-63), bit for bit."""
    scene = P.small_block(dims, table_log2=16)
    sp = P.build_sim_params(scene)
    sim = _sim(scene, grid, path=path)
    try:
        sim.step(1)
        keyS, idxS = sim.sorted_grid()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    V0 = np.zeros_like(X0)
    Pp = oracle.integrate_pred(X0, V0, (sp.gravity.x, sp.gravity.y, sp.gravity.z), sp.dt)
    Pp, _ = oracle.boundary(Pp, V0, sp.grid, 0.0)
    keys, idx = oracle.hash_keys(Pp, sp.grid)
    eK, eI = oracle.sort_pairs(keys, idx)
    assert np.all(np.diff(eK.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(keyS, eK)
    np.testing.assert_array_equal(idxS, eI)


def test_1m_grid_build_matches_oracle(oracle, dev):
    """The benchmark scene's grid build at full size (1,000,000 particles, hashed grid)."""
    scene = P.dam_break_1m()
    sp = P.build_sim_params(scene)
    sim = _sim(scene, "hashed")
    try:
        sim.step(1)
        keyS, idxS = sim.sorted_grid()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    V0 = np.zeros_like(X0)
    Pp = oracle.integrate_pred(X0, V0, (sp.gravity.x, sp.gravity.y, sp.gravity.z), sp.dt)
    Pp, _ = oracle.boundary(Pp, V0, sp.grid, 0.0)
    eK, eI = oracle.sort_pairs(*oracle.hash_keys(Pp, sp.grid))
    np.testing.assert_array_equal(keyS, eK)
    np.testing.assert_array_equal(idxS, eI)
