"""Batched ensemble: every scene must evolve bit for bit as it does alone."""
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


@pytest.mark.parametrize("grid", ["hashed", "compact", "dense"])
def test_ensemble_equals_solo_runs(dev, grid):
    from pbfx.ensemble import Ensemble, solo
    scenes = _scenes()
    ens = Ensemble(scenes, grid=grid)
    ens.step(12)
    for k, sc in enumerate(scenes):
        first, n, dims = ens.scene_info(k)
        assert n == sc.num_particles
        assert dims == tuple(P.build_sim_params(sc).grid.dim)
        s = solo(sc, grid=grid)
        s.step(12)
        X, V = s.state()
        eX, eV = ens.scene_state(k)
        np.testing.assert_array_equal(eX, X, err_msg=f"scene {k}")
        np.testing.assert_array_equal(eV, V, err_msg=f"scene {k}")
        s.close()
    ens.close()
