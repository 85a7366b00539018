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


def _shared_change(sp):
    sp.dt = sp.dt * 0.75
    sp.pbf.scorr_k = 0.005
    sp.pbf.disp_clamp_max_h = 0.08
    return sp


def test_ensemble_set_params_equals_solo_runs(dev):
    """A change of the shared parameters between steps: every scene still evolves as a
    simulator of its own given the same change; per-scene fields cannot be changed."""
    from pbfx import _lib
    from pbfx.ensemble import Ensemble, solo
    scenes = _scenes()[:4]
    ens = Ensemble(scenes, grid="hashed")
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
            s = solo(sc, grid="hashed")
            try:
                s.step(3)
                s.set_params(_shared_change(P.build_sim_params(sc)))
                s.step(3)
                X, V = s.state()
            finally:
                s.close()
            eX, eV = ens.scene_state(k)
            np.testing.assert_array_equal(eX, X, err_msg=f"scene {k}")
            np.testing.assert_array_equal(eV, V, err_msg=f"scene {k}")
    finally:
        ens.close()
