"""GPU adjoint of the reordered step against an independent float64 torch reference.

The reference (tests/torch_pbf.py) is differentiated by torch.autograd and is itself
pinned against float64 central differences (test_adjoint_reference.py), so agreement
here ties the hand-written reverse kernels to finite differences of the same step.
Tolerance: relative L2 error <= 1e-3 for every gradient (float32 kernels vs float64).

The dam-break dynamics are badly conditioned: the float64 gradient norm of the
512-particle block grows about 8x per step (0.96, 2.4, 92, 2.6e4, 3.3e7 at 1, 2, 4,
8, 16 steps), and a float32 trajectory parts from the float64 one at the same rate.
Whole-rollout comparisons are therefore kept to 1-3 steps; longer horizons are
checked step by step along the GPU's own trajectory (each step's float64 vector-
Jacobian product taken at the GPU state, chained backwards).

dL/dc is a sum of per-particle contributions of both signs (on this block their
sum is about 1/16 of the sum of their magnitudes), so its error is measured against
that L1 scale: the reference takes c as a per-particle vector c_k (all equal), whose
gradient holds the contributions; |gc - sum_k g_k| <= 1e-3 * sum_k |g_k|.
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


# the default block, a neighbour cap that cuts most lambda walks, the XSPH gate, and c = 0
VARIANTS = {"default": {}, "cap4": {"max_neighbors": 4},
            "xsph_gate": {"pbf": {"xsph_gate_enable": 1, "xsph_n_min": 2, "xsph_n_max": 7}, "xsph_c": 0.05},
            "xsph_off": {"xsph_c": 0.0}}


def block_state(oracle, n=8, vscale=0.05, seed=11, **kw):
    scene = P.small_block((n, n, n), **kw)
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = seeded_velocity(len(X), seed, vscale)
    return scene, sp, X, V


def ref_grads(scene, sp, X, V, c, steps, wX, wV):
    """float64 reference gradients; dL/dc as (sum, L1 scale) of its per-particle contributions."""
    x0 = torch.tensor(X[:, :3], dtype=torch.float64, requires_grad=True)
    v0 = torch.tensor(V[:, :3], dtype=torch.float64, requires_grad=True)
    cc = torch.full((len(X), 1), float(c), dtype=torch.float64, requires_grad=True)
    xT, vT = ref_rollout(x0, v0, sp, cc, steps, scene.box_min, scene.box_max)
    L = (torch.tensor(wX, dtype=torch.float64) * xT).sum() + (torch.tensor(wV, dtype=torch.float64) * vT).sum()
    L.backward()
    gcv = cc.grad.numpy()
    return x0.grad.numpy(), v0.grad.numpy(), (float(gcv.sum()), float(np.abs(gcv).sum())), xT.detach().numpy()


def gc_err(gc, ref):
    return abs(gc - ref[0]) / ref[1]


@pytest.mark.parametrize("variant,steps,ck", [("default", 1, 1), ("default", 2, 1), ("default", 3, 2),
                                              ("cap4", 1, 1), ("xsph_gate", 2, 1), ("xsph_off", 2, 1)])
def test_rollout_gradients_match_fp64_reference(oracle, dev, variant, steps, ck):
    scene, sp, X, V = block_state(oracle, **VARIANTS[variant])
    n = len(X)
    rng = np.random.default_rng(steps * 10 + ck)
    wX = rng.normal(0, 1, (n, 3))
    wV = 0.01 * rng.normal(0, 1, (n, 3))
    print(f"adjoint vs fp64: {variant} n={n} steps={steps} checkpoint={ck}", flush=True)
    gx_r, gv_r, gc_r, xT_r = ref_grads(scene, sp, X, V, sp.xsph_c, steps, wX, wV)

    roll = Rollout(sp, steps, checkpoint=ck)
    try:
        XT, VT = roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
        dx = float(np.abs(XT[:, :3].cpu().numpy() - xT_r).max())
        gX = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        gV = torch.zeros_like(gX)
        gX[:, :3] = torch.from_numpy(wX).float().to(dev)
        gV[:, :3] = torch.from_numpy(wV).float().to(dev)
        gX0, gV0, gc = roll.backward(gX, gV)
    finally:
        roll.close()
    gX0 = gX0[:, :3].cpu().numpy()
    gV0 = gV0[:, :3].cpu().numpy()
    ev, ex, ec = rel_err(gV0, gv_r), rel_err(gX0, gx_r), gc_err(gc, gc_r)
    print(f"max |x_T - x_T(fp64)| = {dx / sp.kernel.h:.3g} h; rel err gV0 {ev:.3g} gX0 {ex:.3g} gc {ec:.3g}")
    # the float32 trajectory stays close to the float64 one (the gradients are taken on it)
    assert dx < 1e-3 * sp.kernel.h
    assert ev <= REL_TOL and ex <= REL_TOL and ec <= REL_TOL


def _chain(states, sp, scene, wX, wV, dtype):
    """Vector-Jacobian products of the reference step chained backwards over the states
    (last first) in `dtype`, renormalised by a power of two after every step; returns
    (dL/dX_0, dL/dV_0, per-particle dL/dc) as float64 with the scale applied."""
    from torch_pbf import pbf_step
    n = len(wX)
    gx, gv = torch.tensor(wX, dtype=dtype), torch.tensor(wV, dtype=dtype)
    gcv = torch.zeros((n, 1), dtype=torch.float64)
    e_tot = 0
    for x, v in reversed(states):
        x = x.to(dtype).clone().requires_grad_(True)
        v = v.to(dtype).clone().requires_grad_(True)
        c = torch.full((n, 1), float(sp.xsph_c), dtype=dtype, requires_grad=True)
        xn, vn = pbf_step(x, v, sp, c, scene.box_min, scene.box_max)
        gx, gv, g_c = torch.autograd.grad([xn, vn], [x, v, c], [gx, gv])
        gcv += g_c.double() * 2.0 ** e_tot
        e = int(np.frexp(float(torch.maximum(gx.abs().max(), gv.abs().max())))[1])
        gx, gv = gx * 2.0 ** -e, gv * 2.0 ** -e
        e_tot += e
    return gx.double() * 2.0 ** e_tot, gv.double() * 2.0 ** e_tot, gcv


@pytest.mark.parametrize("variant,steps", [(v, 8) for v in sorted(VARIANTS)] + [("default", 60), ("xsph_gate", 60)])
def test_chained_step_gradients_along_gpu_trajectory(oracle, dev, variant, steps):
    """The GPU reverse pass (renormalised every step) against float64 vector-Jacobian
    products of the reference step evaluated at the GPU's own states X_t, V_t, chained.

    8 steps: within 1e-3 relative.  60 steps: the gradient grows by ~2^120 and the
    float32 forward quantities the Jacobian is built from (lambda, the clamp and cap
    decisions near their thresholds) part from their float64 re-evaluation, so no
    float32 evaluation of the step gets within 1e-3 of the float64 chain: the same
    reference evaluated in float32 by torch is 4e-3 (default) and 1.4e-2 (XSPH gate) off
    at 60 steps.  There the bar is that float32 torch chain, another float32 evaluation
    whose error is of the same order (a different draw of the same amplified rounding
    noise): the GPU's error may not exceed 3 times its error plus 1e-3, with 2e-2 as an
    absolute ceiling."""
    scene, sp, X, V = block_state(oracle, seed=21, **VARIANTS[variant])
    n = len(X)
    rng = np.random.default_rng(8)
    wX = rng.normal(0, 1, (n, 3))
    wV = 0.01 * rng.normal(0, 1, (n, 3))
    roll = Rollout(sp, steps, checkpoint=1)
    try:
        roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
        states = [(x[:, :3].double().cpu(), v[:, :3].double().cpu()) for x, v in roll._ckpts]
        gX = torch.zeros((n, 4), dtype=torch.float32, device=dev)
        gV = torch.zeros_like(gX)
        gX[:, :3] = torch.from_numpy(wX).float().to(dev)
        gV[:, :3] = torch.from_numpy(wV).float().to(dev)
        gX0, gV0, gc = roll.backward(gX, gV)
    finally:
        roll.close()
    gx, gv, gcv = _chain(states, sp, scene, wX, wV, torch.float64)
    ev = rel_err(gV0[:, :3].cpu().numpy(), gv.numpy())
    ex = rel_err(gX0[:, :3].cpu().numpy(), gx.numpy())
    ec = gc_err(gc, (float(gcv.sum()), float(gcv.abs().sum())))
    print(f"chained {variant} {steps} steps: |gV0| {float(gv.norm()):.4g} (2^{roll.log2_scale} scale); "
          f"rel err gV0 {ev:.3g} gX0 {ex:.3g} gc {ec:.3g}")
    if steps <= 8:
        assert ev <= REL_TOL and ex <= REL_TOL and ec <= REL_TOL
        return
    fx, fv, fc = _chain(states, sp, scene, wX, wV, torch.float32)
    bv, bx = rel_err(fv.numpy(), gv.numpy()), rel_err(fx.numpy(), gx.numpy())
    bc = gc_err(float(fc.sum()), (float(gcv.sum()), float(gcv.abs().sum())))
    print(f"float32 torch chain: rel err gV0 {bv:.3g} gX0 {bx:.3g} gc {bc:.3g}")
    for err, bar in ((ev, bv), (ex, bx), (ec, bc)):
        assert err <= min(2e-2, 3.0 * bar + REL_TOL)


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_recorded_steps_equal_plain_steps(oracle, dev, variant):
    """Recording a step (tape slot writes) must not change its result: bit-exact."""
    scene, sp, X, V = block_state(oracle, n=10, **VARIANTS[variant])
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


def test_forward_tapes_equal_re_recorded_tapes(oracle, dev):
    """forward() records its last segment as it runs and the first backward() starts
    from those tapes; a second backward() after the same forward re-records every
    segment from its checkpoint.  Both give the same bits, on a single scene and on a
    batched ensemble, and the seeds written on torch's stream just before the call are
    the ones the reverse kernels read (16,384 particles, 12 steps, checkpoint 5)."""
    from pbfx.diff import EnsembleRollout
    scene = differentiable_scene(32, 32, 16)
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = seeded_velocity(len(X), 5, 0.1)
    rng = np.random.default_rng(3)
    for kind in ("single", "ensemble"):
        print(kind, flush=True)
        if kind == "single":
            roll = Rollout(sp, 12, checkpoint=5)
            x0, v0 = torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev)
        else:
            scenes = [differentiable_scene(16, 16, 16, seed=7 + k) for k in range(3)]
            roll = EnsembleRollout(scenes, 12, checkpoint=5)
            X0, _ = roll.ensemble.state()
            x0 = torch.from_numpy(X0).to(dev)
            v0 = torch.from_numpy(seeded_velocity(len(X0), 9, 0.1)).to(dev)
        try:
            roll.forward(x0, v0, sp.xsph_c)
            g = torch.zeros((roll.n, 4), dtype=torch.float32, device=dev)
            g[:, :3] = torch.from_numpy(rng.normal(0, 1, (roll.n, 3))).float().to(dev)
            outs = []
            for _ in range(2):
                gX0, gV0, gc = roll.backward(g)
                outs.append((gX0.cpu().numpy(), gV0.cpu().numpy(), np.asarray(gc)))
            assert np.all(np.isfinite(outs[0][0])) and np.all(np.isfinite(outs[0][1]))
            assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
            assert np.array_equal(outs[0][2], outs[1][2])
        finally:
            roll.close()


@pytest.mark.parametrize("iters", [1, 2, 3])
def test_candidate_mask_does_not_change_gradients(oracle, dev, iters):
    """Recording with the candidate mask (the forward's later passes walk it) and without
    it gives bit-identical gradients (16,384 particles, 8 steps, 1-3 solver iterations)."""
    scene = differentiable_scene(32, 32, 16)
    sp = P.build_sim_params(scene)
    sp.solverIters = iters
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = seeded_velocity(len(X), 6, 0.1)
    g = torch.zeros((len(X), 4), dtype=torch.float32, device=dev)
    g[:, :3] = torch.from_numpy(np.random.default_rng(3).normal(0, 1, (len(X), 3))).float().to(dev)
    outs = []
    for mask in (True, False):
        print(f"solver iterations {iters}, candidate mask {mask}", flush=True)
        roll = Rollout(sp, 8, checkpoint=4)
        try:
            roll.sim.set_tuning(candidate_mask=mask)
            assert roll.sim.tuning()["candidate_mask"] == mask
            roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
            gX0, gV0, gc = roll.backward(g)
            outs.append((gX0.cpu().numpy(), gV0.cpu().numpy(), gc))
        finally:
            roll.close()
    assert np.all(np.isfinite(outs[0][0])) and np.abs(outs[0][0]).sum() > 0
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


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
    cc = torch.full((len(X), 1), float(sp.xsph_c), dtype=torch.float64, requires_grad=True)
    xT, _ = ref_rollout(x0, v0, sp, cc, steps, scene.box_min, scene.box_max)
    loss_of(xT, torch.float64, "cpu").backward()
    gc_ref = (float(cc.grad.sum()), float(cc.grad.abs().sum()))

    roll = Rollout(sp, steps, checkpoint=2)
    try:
        v0g = torch.from_numpy(V[:, :3]).to(dev).requires_grad_(True)
        cg = torch.tensor(float(sp.xsph_c), dtype=torch.float32, device=dev, requires_grad=True)
        xTg = PBFRollout.apply(v0g, cg, torch.from_numpy(X[:, :3]).to(dev), roll)
        loss_of(xTg, torch.float64, dev).backward()
    finally:
        roll.close()
    assert rel_err(v0g.grad.cpu().numpy(), v0.grad.numpy()) <= REL_TOL
    assert gc_err(float(cg.grad), gc_ref) <= REL_TOL


def test_ensemble_rollout_matches_single_scenes(oracle, dev):
    """Three scenes of different sizes, gravities and XSPH coefficients differentiated
    together as one ensemble give each scene's single-scene gradients (x0, v0 bit for
    bit; dL/dc up to summation order)."""
    from pbfx.diff import EnsembleRollout
    scenes = [P.small_block((6, 6, 6), xsph_c=0.02),
              P.small_block((8, 7, 6), xsph_c=0.05, gravity=(0.0, -8.0, 0.0)),
              P.small_block((5, 5, 9), xsph_c=0.0, gravity=(0.0, -11.0, 0.0))]
    steps, ck = 4, 2
    ens = EnsembleRollout(scenes, steps, checkpoint=ck)
    try:
        X0, _ = ens.ensemble.state()
        n = len(X0)
        V0 = seeded_velocity(n, 31, 0.05)
        rng = np.random.default_rng(5)
        gX = np.zeros((n, 4), np.float32)
        gV = np.zeros((n, 4), np.float32)
        gX[:, :3] = rng.normal(0, 1, (n, 3))
        gV[:, :3] = 0.01 * rng.normal(0, 1, (n, 3))
        ens.forward(torch.from_numpy(X0).to(dev), torch.from_numpy(V0).to(dev))
        eX, eV, egc = ens.backward(torch.from_numpy(gX).to(dev), torch.from_numpy(gV).to(dev))
        eX, eV = eX.cpu().numpy(), eV.cpu().numpy()
        firsts = ens.firsts
    finally:
        ens.close()
    assert len(egc) == 3
    for k, sc in enumerate(scenes):
        a, b = firsts[k], firsts[k] + sc.num_particles
        sp = P.build_sim_params(sc)
        roll = Rollout(sp, steps, checkpoint=ck)
        try:
            roll.forward(torch.from_numpy(X0[a:b]).to(dev), torch.from_numpy(V0[a:b]).to(dev), sc.xsph_c)
            sX, sV, sgc = roll.backward(torch.from_numpy(gX[a:b]).to(dev), torch.from_numpy(gV[a:b]).to(dev))
        finally:
            roll.close()
        print(f"scene {k}: gc ensemble {egc[k]:.9g} single {sgc:.9g}", flush=True)
        assert np.array_equal(eX[a:b], sX.cpu().numpy()) and np.array_equal(eV[a:b], sV.cpu().numpy())
        assert abs(egc[k] - sgc) <= 1e-9 * max(1.0, abs(sgc))


def _configured_loss_gradient(dev):
    """The differentiable-rollout configuration (65,536 particles, 64x32x32 lattice, 60
    steps, dt = 1/120, 2 iterations) and its occupancy loss: the target grid comes from
    a forward run with a seeded perturbed initial velocity; returns (scene params, x0,
    loss, dL/dX_T).  The gradient is computed once and handed to every reverse pass
    compared, so the comparison does not rest on the loss evaluating to the same bits
    twice."""
    from pbfx.diff import initial_state
    scene = differentiable_scene()
    sp = P.build_sim_params(scene)
    n, h = sp.numParticles, float(sp.kernel.h)
    X0, _ = initial_state(scene)
    x0 = torch.from_numpy(X0).to(dev)
    roll = Rollout(sp, 60, checkpoint=10)
    try:
        XT, _ = roll.forward(x0, torch.from_numpy(seeded_velocity(n, 77, 0.5)).to(dev), sp.xsph_c)
        target = occupancy_grid(XT[:, :3].double(), scene.box_min, scene.box_max, cells=32, sigma=h)
        XT, _ = roll.forward(x0, torch.zeros((n, 4), dtype=torch.float32, device=dev), sp.xsph_c)
        xT = XT[:, :3].double().requires_grad_(True)
        loss = ((occupancy_grid(xT, scene.box_min, scene.box_max, cells=32, sigma=h) - target) ** 2).sum()
        loss.backward()
    finally:
        roll.close()
    return sp, x0, float(loss.detach()), xT.grad.float()


@pytest.mark.timeout(300)
def test_configured_rollout_gradient_finite_and_checkpoint_invariant(dev):
    """The configured 65,536 x 60 rollout with the occupancy loss: the gradient w.r.t. v0
    and c is finite and non-zero (the reverse pass is renormalised, so it does not
    overflow float32 although it outgrows it), and identical for checkpoint intervals
    10 and 20 (each reverse segment is re-recorded from its checkpoint)."""
    sp, x0, loss, gXT = _configured_loss_gradient(dev)
    outs = {}
    for ck in (10, 20):
        print(f"65,536 x 60 rollout, checkpoint {ck}", flush=True)
        roll = Rollout(sp, 60, checkpoint=ck)
        try:
            roll.forward(x0, torch.zeros((sp.numParticles, 4), dtype=torch.float32, device=dev), sp.xsph_c)
            gX0, gV0, gc = roll.backward(gXT.clone())
            outs[ck] = (gX0.cpu().numpy(), gV0.cpu().numpy(), gc, roll.log2_scale)
        finally:
            roll.close()
    gX0, gV0, gc, e = outs[10]
    gnorm = float(np.linalg.norm(gV0[:, :3]))
    print(f"loss {loss:.6g}; |dL/dv0| {gnorm:.4g} (renormalised by 2^{e}); dL/dc {gc:.4g}", flush=True)
    assert np.isfinite(loss) and loss > 0
    assert np.all(np.isfinite(gX0)) and np.all(np.isfinite(gV0)) and np.isfinite(gc)
    assert gnorm > 0 and gc != 0
    for a, b in zip(outs[10], outs[20]):
        assert np.array_equal(a, b)


def test_parameter_change_after_forward_keeps_the_recorded_step(oracle, dev):
    """A parameter change between forward() and backward() does not change which
    trajectory is differentiated: each tape keeps the parameters it was recorded with
    and the rollout re-records with the forward's, so the gradients equal those of an
    unchanged rollout bit for bit, and the simulator keeps the new parameters."""
    scene, sp, X, V = block_state(oracle, n=8, seed=4)
    n = len(X)
    g = torch.zeros((n, 4), dtype=torch.float32, device=dev)
    g[:, :3] = torch.from_numpy(np.random.default_rng(4).normal(0, 1, (n, 3))).float().to(dev)
    outs = []
    for change in (False, True):
        roll = Rollout(sp, 4, checkpoint=2)
        try:
            roll.forward(torch.from_numpy(X).to(dev), torch.from_numpy(V).to(dev), sp.xsph_c)
            if change:
                sp2 = P.build_sim_params(scene)
                sp2.dt = sp2.dt * 0.5
                sp2.pbf.scorr_k = 0.003
                roll.sim.set_params(sp2)
            gX0, gV0, gc = roll.backward(g)
            outs.append((gX0.cpu().numpy(), gV0.cpu().numpy(), gc))
            if change:
                assert roll._params_now().dt == sp2.dt
        finally:
            roll.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


@pytest.mark.timeout(300)
def test_batched_rollouts_gradient_finite(dev):
    """The 64 batched rollouts of the configuration (4,194,304 particles, 60 steps,
    checkpoint 10) with the occupancy loss per rollout: every gradient finite.  Their
    collapsed corner piles hold near-coincident particles, where the reverse pass's
    spiky Hessian must not overflow."""
    from pbfx.diff import EnsembleRollout
    scenes = [differentiable_scene(seed=7 + k) for k in range(64)]
    roll = EnsembleRollout(scenes, 60, checkpoint=10)
    try:
        X0, _ = roll.ensemble.state()
        n = len(X0)
        h = float(roll.params.kernel.h)
        x0 = torch.from_numpy(X0).to(dev)
        vt = torch.from_numpy(np.concatenate([seeded_velocity(sc.num_particles, 1000 + k, 0.5)
                                              for k, sc in enumerate(scenes)])).to(dev)
        XT, _ = roll.forward(x0, vt)
        spans = [(f, f + sc.num_particles) for f, sc in zip(roll.firsts, scenes)]
        targets = [occupancy_grid(XT[a:b, :3].double(), sc.box_min, sc.box_max, cells=32, sigma=h)
                   for (a, b), sc in zip(spans, scenes)]
        XT, _ = roll.forward(x0, torch.zeros((n, 4), dtype=torch.float32, device=dev))
        xT = XT[:, :3].double().requires_grad_(True)
        loss = sum(((occupancy_grid(xT[a:b], sc.box_min, sc.box_max, cells=32, sigma=h) - tg) ** 2).sum()
                   for (a, b), sc, tg in zip(spans, scenes, targets))
        loss.backward()
        gX0, gV0, gc = roll.backward(xT.grad.float())
        print(f"64 rollouts: renormalised by 2^{roll.log2_scale}", flush=True)
        assert bool(torch.isfinite(gX0).all()) and bool(torch.isfinite(gV0).all())
        assert np.all(np.isfinite(gc))
        for a, b in spans:
            assert float(gV0[a:b, :3].abs().max()) > 0
    finally:
        roll.close()
