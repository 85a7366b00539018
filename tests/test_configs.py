"""Scene configurations (no GPU): parameter derivation and the ensemble generator."""
import numpy as np

from pbfx import params as P


def test_dam_break_configs():
    for sc, n, dims, iters in ((P.dam_break_32k(), 32768, (81, 81, 81), 2),
                               (P.dam_break_1m(), 1_000_000, (161, 81, 81), 2),
                               (P.splash_8m(), 8_000_000, (400, 200, 400), 4)):
        sp = P.build_sim_params(sc)
        assert sp.numParticles == n and tuple(sp.grid.dim) == dims and sp.solverIters == iters
        assert sp.particleMass == np.float32(np.float32(np.float32(np.float32(1000.0) * sp.kernel.h) * sp.kernel.h)
                                             * sp.kernel.h)
        assert sp.maxNeighbors == 64 and sp.pbf.compliance == np.float32(2e-5)
        # every lattice fits in its box
        for nx, ny, nz, o in sc.lattices:
            top = np.array(o) + (np.array([nx, ny, nz]) - 1) * sc.spacing
            assert np.all(top <= np.array(sc.box_max))


def test_ensemble_scene_generator():
    from pbfx.ensemble import ensemble_scenes
    sc = ensemble_scenes(256, seed=2024)
    n = np.array([s.num_particles for s in sc])
    assert len(sc) == 256
    assert n.min() >= 0.8 * 4096 and n.max() <= 1.2 * 262144
    assert 40_000 < n.mean() < 90_000
    for s in sc:
        (nx, ny, nz, _), = s.lattices
        ext = np.array([nx, ny, nz]) * s.spacing
        assert np.all(np.array(s.box_max) >= ext)
        assert -12 <= s.gravity[1] <= -8 and 0.005 <= s.xsph_c <= 0.05
