"""CubeMix seeding and the external-buffer boundary on the GPU.

Seeding is checked bit for bit against the oracle's numpy restatement. External
position buffers must not change a trajectory: a simulator stepping in caller-owned
buffers (reference: bindExternalPosPingPong, simulator.cpp:676-757) gives the oracle's
states bit for bit, and the same bits as one using its own."""
import numpy as np
import pytest
import torch

from pbfx import params as P
from pbfx import scenes as S
from pbfx.sim import Simulator

pytestmark = pytest.mark.gpu


def cube_mix_params(groups=3, edge=6, box=((0, 0, 0), (1, 1, 1))):
    scene = P.small_block((1, 1, 1), box_min=box[0], box_max=box[1])
    sp = P.build_sim_params(scene)
    sp.numParticles = sp.maxParticles = groups * edge ** 3
    return sp


@pytest.mark.parametrize("jitter", [0.0, 0.0005])
def test_cube_mix_seed_matches_oracle(oracle, dev, jitter):
    lay = S.CubeMixLayout(group_count=5, layers=2, group_spacing=0.2, layer_spacing=0.15, base_height=0.1)
    centers = S.cube_mix_centers(lay)
    edge, spacing = 7, 0.0125 * 1.02
    sp = cube_mix_params(5, edge)
    sim = Simulator(sp, grid="dense", table_log2=16)
    try:
        sim.seed_cube_mix(centers, edge, spacing, jitter, seed=1234)
        X, V = sim.state()
    finally:
        sim.close()
    exp = oracle.seed_cube_mix((0, 0, 0), (1, 1, 1), centers, edge, spacing, jitter, 1234)
    assert np.array_equal(X, exp)
    assert np.all(V == 0)
    # groups are clamped into the box: the first centre sits on the floor corner
    assert X[:, :3].min() >= 0.0


def test_cube_mix_runs_and_rejects_overflow(dev):
    sp = cube_mix_params(2, 5)
    sim = Simulator(sp, grid="hashed", table_log2=16)
    try:
        c = np.array([(0.2, 0.2, 0.2), (0.6, 0.3, 0.5)], np.float32)
        sim.seed_cube_mix(c, 5, 0.0125, 0.0, 0)
        sim.step(5)
        X, _ = sim.state()
        assert np.isfinite(X).all()
        with pytest.raises(Exception):
            sim.seed_cube_mix(c, 6, 0.0125, 0.0, 0)  # 2 * 216 > 250 particles
    finally:
        sim.close()


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("path,grid", [("reordered", "dense"), ("reordered", "hashed"), ("reference", "compact")])
def test_external_positions_bit_identical(oracle, dev, use_graph, path, grid):
    scene = P.small_block((10, 10, 10), table_log2=16)
    sp = P.build_sim_params(scene)
    a = Simulator.from_scene(scene, use_graph=use_graph, path=path, grid=grid)
    b = Simulator.from_scene(scene, use_graph=use_graph, path=path, grid=grid)
    try:
        a.step(3)
        b.step(3)
        n = b.n
        ta = torch.full((n + 5, 4), float("nan"), dtype=torch.float32, device=dev)
        tb = torch.full((n + 5, 4), float("nan"), dtype=torch.float32, device=dev)
        b.bind_external_positions(ta, tb)
        a.step(4)
        b.step(4)
        xa, va = a.state()
        xb, vb = b.state()
        # the oracle's 7 steps of the same scene, with the binding between steps 3 and 4
        X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
        compact = grid != "dense"
        eX, eV = oracle.run(X0, np.zeros_like(X0), sp, 3, compact=compact)
        eX, eV = oracle.run(eX, eV, sp, 4, compact=compact)
        np.testing.assert_array_equal(xb, eX)
        np.testing.assert_array_equal(vb, eV)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        # the current state lives in one of the caller's buffers, viewed without a copy
        view = b.positions_view()
        assert view.data_ptr() in (ta.data_ptr(), tb.data_ptr())
        assert np.array_equal(view.cpu().numpy(), xb)
        assert np.array_equal(b.velocities_view().cpu().numpy(), vb)
    finally:
        a.close()
        b.close()
    del ta, tb


def test_publish_velocity_strided(dev):
    scene = P.small_block((6, 6, 6))
    sim = Simulator.from_scene(scene)
    try:
        sim.step(2)
        n = sim.n
        out = torch.zeros((n, 8), dtype=torch.float32, device=dev)  # 32-byte stride
        sim.bind_external_velocity(out, stride_bytes=32)
        sim.publish_velocity()
        sim.synchronize()
        _, V = sim.state()
        got = out.cpu().numpy()
        assert np.array_equal(got[:, :4], V) and np.all(got[:, 4:] == 0)
        # a buffer too small for every particle receives the rows that fit
        small = torch.zeros(40 * 4 + 4, dtype=torch.float32, device=dev)  # 10 rows at stride 64 B + 1
        sim.bind_external_velocity(small, stride_bytes=64)
        sim.publish_velocity()
        sim.synchronize()
        rows = small.cpu().numpy()[:160].reshape(10, 16)[:, :4]
        assert np.array_equal(rows, V[:10])
    finally:
        sim.close()
