"""Pins of the float64 torch reference used for adjoint gradients (no GPU):
its forward agrees with the CPU oracle, and its autograd gradients agree with
float64 central finite differences."""
import numpy as np
import pytest
import torch

from pbfx import params as P
from torch_pbf import rollout


def small_case(n=8, steps=3, **kw):
    scene = P.small_block((n, n, n), **kw)
    sp = P.build_sim_params(scene)
    return scene, sp, steps


# the default block, a neighbour cap that cuts most walks, the XSPH gate, and c = 0
VARIANTS = {"default": {}, "cap4": {"max_neighbors": 4},
            "xsph_gate": {"pbf": {"xsph_gate_enable": 1, "xsph_n_min": 2, "xsph_n_max": 7}, "xsph_c": 0.05},
            "xsph_off": {"xsph_c": 0.0}}


def count_capped(oracle, scene, sp):
    """Particles whose first-step lambda walk reaches the cap (float64 count on predicted positions)."""
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    h2 = float(sp.kernel.h) ** 2
    p = X[:, :3].astype(np.float64) + np.array(list(sp.gravity)) * sp.dt * sp.dt
    d2 = ((p[:, None, :] - p[None, :, :]) ** 2).sum(-1)
    n = (d2 <= h2).sum(1) - 1
    return int((n >= sp.maxNeighbors).sum())


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_reference_forward_matches_oracle(oracle, variant):
    # 2 steps: float32 and float64 trajectories part ~10x per step later on
    scene, sp, steps = small_case(steps=2, **VARIANTS[variant])
    if variant == "cap4":
        assert count_capped(oracle, scene, sp) > 100
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    rng = np.random.default_rng(1)
    V = np.zeros_like(X)
    V[:, :3] = rng.normal(0, 0.05, (len(X), 3)).astype(np.float32)
    eX, eV = oracle.run(X, V, sp, steps)
    tX, tV = rollout(torch.tensor(X[:, :3], dtype=torch.float64), torch.tensor(V[:, :3], dtype=torch.float64), sp,
                     float(sp.xsph_c), steps, scene.box_min, scene.box_max)
    h = sp.kernel.h
    assert np.abs(tX.numpy() - eX[:, :3]).max() < 1e-4 * h
    assert np.abs(tV.numpy() - eV[:, :3]).max() < 1e-4 * h / sp.dt


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_reference_gradient_matches_finite_differences(oracle, variant):
    scene, sp, steps = small_case(n=6, steps=2, **VARIANTS[variant])
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    rng = np.random.default_rng(4)
    x0 = torch.tensor(X[:, :3], dtype=torch.float64)
    v0 = torch.tensor(rng.normal(0, 0.05, (len(X), 3)), dtype=torch.float64, requires_grad=True)
    c = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    w = torch.tensor(rng.normal(0, 1, (len(X), 3)), dtype=torch.float64)

    def loss(v, cc):
        xT, vT = rollout(x0, v, sp, cc, steps, scene.box_min, scene.box_max)
        return (w * xT).sum() + 0.01 * (w * vT).sum()

    L = loss(v0, c)
    L.backward()
    gv, gc = v0.grad.clone(), c.grad.clone()
    eps = 1e-7
    for k in range(3):
        e = torch.tensor(rng.normal(0, 1, v0.shape), dtype=torch.float64)
        with torch.no_grad():
            fd = (loss(v0 + eps * e, c) - loss(v0 - eps * e, c)) / (2 * eps)
        assert float((gv * e).sum()) == pytest.approx(float(fd), rel=1e-5, abs=1e-9)
    with torch.no_grad():
        fdc = (loss(v0, c + eps) - loss(v0, c - eps)) / (2 * eps)
    assert float(gc) == pytest.approx(float(fdc), rel=1e-5, abs=1e-9)
