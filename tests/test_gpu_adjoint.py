"""GPU adjoint of the reordered step against an independent float64 torch reference.

The reference (tests/torch_pbf.py) is differentiated by torch.autograd and is itself
pinned against float64 central differences (test_adjoint_reference.py), so agreement
here ties the hand-written reverse kernels to finite differences of the same step.
Tolerance: relative L2 error <= 1e-3 for every gradient (float32 kernels vs float64).
"""
import ctypes

import numpy as np
import pytest
import torch

from pbfx import _lib
from pbfx import params as P
from pbfx.diff import PBFRollout, Rollout, differentiable_scene, occupancy_grid, seeded_velocity
from torch_pbf import rollout as ref_rollout

pytestmark = pytest.mark.gpu

REL_TOL = 1e-3


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def block_state(oracle, n=8, vscale=0.05, seed=11):
    scene = P.small_block((n, n, n))
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = seeded_velocity(len(X), seed, vscale)
    return scene, sp, X, V


def ref_grads(scene, sp, X, V, c, steps, wX, wV):
    x0 = torch.tensor(X[:, :3], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(V[:, :3], dtype=torch.float64, requires_grad=True)
    cc = torch.tensor(float(c), dtype=torch.float64, requires_grad=True)
    xT, vT = ref_rollout(x0, v0, sp, cc, steps, scene.box_min, scene.box_max)
    L = (torch.tensor(wX, dtype=torch.float64) * xT).sum() + (torch.tensor(wV, dtype=torch.float64) * vT).sum()
    L.backward()
    return x0.grad.numpy(), v0.grad.numpy(), float(cc.grad), xT.detach().numpy()


@pytest.mark.parametrize("steps,ck", [(1, 1), (4, 2), (5, 3)])
def test_rollout_gradients_match_fp64_reference(oracle, dev, steps, ck):
    scene, sp, X, V = block_state(oracle)
    n = len(X)
    rng = np.random.default_rng(steps * 10 + ck)
    wX = rng.normal(0, 1, (n, 3))
    wV = 0.01 * rng.normal(0, 1, (n, 3))
    print(f"adjoint vs fp64: n={n} steps={steps} checkpoint={ck}", flush=True)
    gx_r, gv_r, gc_r, xT_r = ref_grads(scene, sp, X, V, sp.xsph_c, steps, wX, wV)

    roll = Rollout(sp, steps, checkpoint=ck)
    try:
        XT, VT = roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
        assert np.abs(XT[:, :3].cpu().numpy() - xT_r).max() < 1e-4 * sp.kernel.h
        gX = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        gV = torch.zeros_like(gX)
        gX[:, :3] = torch.from_numpy(wX).float().to(dev)
        gV[:, :3] = torch.from_numpy(wV).float().to(dev)
        gX0, gV0, gc = roll.backward(gX, gV)
    finally:
        roll.close()
    gX0 = gX0[:, :3].cpu().numpy()
    gV0 = gV0[:, :3].cpu().numpy()
    assert rel_err(gV0, gv_r) <= REL_TOL
    assert rel_err(gX0, gx_r) <= REL_TOL
    assert abs(gc - gc_r) <= REL_TOL * max(abs(gc_r), 1e-12)


def test_recorded_steps_equal_plain_steps(oracle, dev):
    """Recording a step (tape slot writes) must not change its result: bit-exact."""
    scene, sp, X, V = block_state(oracle, n=10)
    L = _lib.lib()
    from pbfx.sim import Simulator
    a = Simulator(sp, grid="dense", table_log2=16, use_graph=False)
    b = Simulator(sp, grid="dense", table_log2=16, use_graph=False)
    try:
        assert L.pbfSimAdjointInit(b._h, 3) == 0
        a.set_state(X, V)
        b.set_state(X, V)
        a.step(3)
        for t in range(3):
            assert L.pbfSimStepRecord(b._h, t) == 0
        xa, va = a.state()
        xb, vb = b.state()
    finally:
        a.close()
        b.close()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


def test_checkpoint_interval_does_not_change_gradients(oracle, dev):
    """The reverse pass recomputes each segment from its checkpoint; the gradient
    is bit-identical for every checkpoint interval (16,384 particles, 12 steps)."""
    scene = differentiable_scene(32, 32, 16)
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = seeded_velocity(len(X), 5, 0.1)
    g = torch.zeros((len(X), 4), dtype=torch.float32, device=dev)
    g[:, :3] = torch.from_numpy(np.random.default_rng(2).normal(0, 1, (len(X), 3))).float().to(dev)
    outs = []
    for ck in (1, 5, 12):
        print(f"checkpoint interval {ck}", flush=True)
        roll = Rollout(sp, 12, checkpoint=ck)
        try:
            roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
            gX0, gV0, gc = roll.backward(g)
            outs.append((gX0.cpu().numpy(), gV0.cpu().numpy(), gc))
        finally:
            roll.close()
    for o in outs[1:]:
        assert np.array_equal(o[0], outs[0][0]) and np.array_equal(o[1], outs[0][1]) and o[2] == outs[0][2]


def test_autograd_function_with_occupancy_loss(oracle, dev):
    """PBFRollout under torch.autograd with the Gaussian occupancy loss, against the
    float64 reference under the same loss."""
    scene, sp, X, V = block_state(oracle, n=8, seed=3)
    steps = 3
    target = torch.rand((16, 16, 16), dtype=torch.float64, generator=torch.Generator().manual_seed(9))

    def loss_of(xT, dtype, device):
        grid = occupancy_grid(xT.to(dtype), scene.box_min, scene.box_max, cells=16, sigma=0.02)
        return ((grid - target.to(dtype=dtype, device=device)) ** 2).sum()

    x0 = torch.tensor(X[:, :3], dtype=torch.float64)
    v0 = torch.tensor(V[:, :3], dtype=torch.float64, requires_grad=True)
    cc = torch.tensor(float(sp.xsph_c), dtype=torch.float64, requires_grad=True)
    xT, _ = ref_rollout(x0, v0, sp, cc, steps, scene.box_min, scene.box_max)
    loss_of(xT, torch.float64, "cpu").backward()

    roll = Rollout(sp, steps, checkpoint=2)
    try:
        v0g = torch.from_numpy(V[:, :3]).to(dev).requires_grad_(True)
        cg = torch.tensor(float(sp.xsph_c), dtype=torch.float32, device=dev, requires_grad=True)
        xTg = PBFRollout.apply(v0g, cg, torch.from_numpy(X[:, :3]).to(dev), roll)
        loss_of(xTg, torch.float64, dev).backward()
    finally:
        roll.close()
    assert rel_err(v0g.grad.cpu().numpy(), v0.grad.numpy()) <= REL_TOL
    assert abs(float(cg.grad) - float(cc.grad)) <= REL_TOL * abs(float(cc.grad))
