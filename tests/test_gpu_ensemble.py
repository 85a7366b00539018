"""Batched ensemble: every scene must evolve bit for bit as the oracle's run of that scene
alone, with the scene's own box, gravity, XSPH coefficient and restitution.

The batched step is the reference's single-scene step (sim/simulator.cpp:560-643, phase
order sim/phases.cpp:108-117) applied to many scenes at once, so each scene's expected
state is oracle.run on that scene's own SimParams.
"""
import numpy as np
import pytest

from pbfx import params as P

pytestmark = pytest.mark.gpu


def _scenes():
    from pbfx.ensemble import ensemble_scenes
    sc = ensemble_scenes(count=7, seed=11, n_min=300, n_max=9000)
    # one scene with restitution and one without XSPH, to exercise per-scene fields
    sc[1].restitution = 0.5
    sc[2].xsph_c = 0.0
    return sc


def _oracle_scene(oracle, sc, runs, compact):
    """The oracle's run of one scene alone: runs = [(SimParams, steps), ...] in order."""
    X = P.seed_scene_reference(sc, oracle.seed_lattice)
    V = np.zeros_like(X)
    for sp, steps in runs:
        X, V = oracle.run(X, V, sp, steps, compact=compact)
    return X, V


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("grid", ["hashed", "compact", "dense"])
def test_ensemble_scenes_match_oracle(oracle, dev, grid, use_graph):
    from pbfx.ensemble import Ensemble
    scenes = _scenes()
    ens = Ensemble(scenes, grid=grid, use_graph=use_graph)
    try:
        ens.step(12)
        for k, sc in enumerate(scenes):
            first, n, dims = ens.scene_info(k)
            assert n == sc.num_particles
            sp = P.build_sim_params(sc)
            assert dims == tuple(sp.grid.dim)
            eX, eV = _oracle_scene(oracle, sc, [(sp, 12)], compact=grid != "dense")
            X, V = ens.scene_state(k)
            np.testing.assert_array_equal(X, eX, err_msg=f"scene {k}")
            np.testing.assert_array_equal(V, eV, err_msg=f"scene {k}")
    finally:
        ens.close()


def test_ensemble_nan_in_one_scene(oracle, dev):
    """A NaN velocity in one scene of an ensemble: that scene evolves as the oracle's run
    of it alone (NaN where it has NaN), and every other scene as its own run, although
    the scenes share the step's flags (the NaN turns the candidate mask off for all of
    them: the full walk gives the same sums)."""
    from pbfx import _lib
    from pbfx.ensemble import Ensemble
    scenes = _scenes()
    ens = Ensemble(scenes, grid="hashed", use_graph=True)
    try:
        ens.step(6)
        total = sum(sc.num_particles for sc in scenes)
        X = np.zeros((total, 4), np.float32)
        V = np.zeros_like(X)
        assert _lib.lib().pbfSimGetState(ens._h, X.ctypes.data, V.ctypes.data, 0) == 0
        first, n, _ = ens.scene_info(3)
        V[first + n // 2, :3] = [np.nan, 0.0, 0.0]
        assert _lib.lib().pbfSimSetState(ens._h, X.ctypes.data, V.ctypes.data, total, 0) == 0
        ens.step(2)
        for k, sc in enumerate(scenes):
            f, n, _ = ens.scene_info(k)
            eX, eV = oracle.run(X[f:f + n], V[f:f + n], P.build_sim_params(sc), 2, compact=True)
            if k == 3:
                assert np.isnan(eX).any()
            gX, gV = ens.scene_state(k)
            np.testing.assert_array_equal(gX, eX, err_msg=f"scene {k}")
            np.testing.assert_array_equal(gV, eV, err_msg=f"scene {k}")
    finally:
        ens.close()


def test_ensemble_equals_solo_runs(dev):
    """The same scenes as single-scene simulators give the same bits (scene-by-scene path)."""
    from pbfx.ensemble import Ensemble, solo
    scenes = _scenes()
    ens = Ensemble(scenes, grid="hashed")
    try:
        ens.step(6)
        for k, sc in enumerate(scenes):
            s = solo(sc, grid="hashed")
            try:
                s.step(6)
                X, V = s.state()
            finally:
                s.close()
            eX, eV = ens.scene_state(k)
            np.testing.assert_array_equal(eX, X, err_msg=f"scene {k}")
            np.testing.assert_array_equal(eV, V, err_msg=f"scene {k}")
    finally:
        ens.close()


def _shared_change(sp):
    sp.dt = sp.dt * 0.75
    sp.pbf.scorr_k = 0.005
    sp.pbf.disp_clamp_max_h = 0.08
    sp.maxNeighbors = 40
    return sp


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("grid", ["hashed", "compact", "dense"])
def test_ensemble_set_params_matches_oracle(oracle, dev, grid, use_graph):
    """A change of the shared parameters between steps: every scene equals the oracle's
    run of that scene with its own parameters for 3 steps, then the changed ones for 3
    (reference: ParamChangeTracker, param_change_tracker.h:37-56).  Per-scene fields
    cannot be changed through the shared parameters."""
    from pbfx import _lib
    from pbfx.ensemble import Ensemble
    scenes = _scenes()[:4]
    ens = Ensemble(scenes, grid=grid, use_graph=use_graph)
    try:
        ens.step(3)
        base = ens.current_params()
        for field, value in (("xsph_c", 0.2), ("boundaryRestitution", 0.9)):
            sp = type(base).from_buffer_copy(base)
            setattr(sp, field, value)
            with pytest.raises(_lib.PbfError):
                ens.set_params(sp)
        sp = type(base).from_buffer_copy(base)
        sp.gravity.x = 1.0
        with pytest.raises(_lib.PbfError):
            ens.set_params(sp)
        ens.set_params(_shared_change(type(base).from_buffer_copy(base)))
        ens.step(3)
        for k, sc in enumerate(scenes):
            spA = P.build_sim_params(sc)
            spB = _shared_change(P.build_sim_params(sc))
            eX, eV = _oracle_scene(oracle, sc, [(spA, 3), (spB, 3)], compact=grid != "dense")
            X, V = ens.scene_state(k)
            np.testing.assert_array_equal(X, eX, err_msg=f"scene {k}")
            np.testing.assert_array_equal(V, eV, err_msg=f"scene {k}")
    finally:
        ens.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("grid", ["dense", "hashed"])
def test_ensemble_256_scenes_full_size(oracle, dev, grid):
    """The batched configuration at full size: 256 scenes, particle counts log-uniform in
    [4,096, 262,144] (18.1M particles), per-scene aspect, gravity and XSPH from seed 2024.
    After 3 steps every scene is finite, inside its own box, with w = 1 / 0; a seeded
    sample of scenes (always the smallest and the largest) equals the oracle's run of
    the scene alone, bit for bit."""
    from pbfx.ensemble import Ensemble, ensemble_scenes
    scenes = ensemble_scenes(256, seed=2024)
    sizes = np.array([sc.num_particles for sc in scenes])
    print(f"256-scene ensemble, {grid} grid: {int(sizes.sum())} particles", flush=True)
    steps = 3
    ens = Ensemble(scenes, grid=grid)
    try:
        ens.step(steps)
        X, V = ens.state()
        firsts = [ens.scene_info(k)[0] for k in range(len(scenes))]
        assert not (grid == "hashed" and ens.hash_overflow())
    finally:
        ens.close()
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(V))
    assert np.all(X[:, 3] == 1.0) and np.all(V[:, 3] == 0.0)
    for k, sc in enumerate(scenes):
        a, b = firsts[k], firsts[k] + sc.num_particles
        lo = np.array(sc.box_min, np.float32)
        hi = np.array(sc.box_max, np.float32)
        assert np.all(X[a:b, :3] >= lo) and np.all(X[a:b, :3] <= hi), f"scene {k} left its box"
    rng = np.random.default_rng(256)
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))} |
                    set(int(i) for i in rng.choice(len(scenes), 6, replace=False)))
    for k in sample:
        sc = scenes[k]
        a, b = firsts[k], firsts[k] + sc.num_particles
        eX, eV = _oracle_scene(oracle, sc, [(P.build_sim_params(sc), steps)], compact=grid != "dense")
        print(f"scene {k}: {sc.num_particles} particles", flush=True)
        np.testing.assert_array_equal(X[a:b], eX, err_msg=f"scene {k}")
        np.testing.assert_array_equal(V[a:b], eV, err_msg=f"scene {k}")
