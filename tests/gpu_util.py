"""Shared helpers of the GPU parity tests (numpy <-> torch, seeded inputs)."""
import numpy as np

from pbfx import params as P

H_TOL = 1e-5  # positions: 1e-5 * h; velocities: 1e-5 * h / dt


def to_dev(a, dev):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).to(dev)
    return torch.from_numpy(a).to(dev)


def to_host(t, dtype=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    if dtype == np.uint32:
        return a.view(np.uint32)
    return a


def empty_u32(n, dev):
    import torch
    return torch.zeros(max(n, 1), dtype=torch.int32, device=dev)[:n] if n else torch.zeros(0, dtype=torch.int32,
                                                                                             device=dev)


def predicted_state(oracle, scene, steps=0, squeeze=None):
    """Seeded scene, optionally advanced `steps` oracle steps, then predict + clamp:
    the positions the grid build and the solver see."""
    sp = P.build_sim_params(scene)
    X = P.seed_scene_reference(scene, oracle.seed_lattice)
    V = np.zeros_like(X)
    if steps:
        X, V = oracle.run(X, V, sp, steps)
    if squeeze is not None:
        X[:, :3] = (X[:, :3] * np.float32(squeeze[0]) + np.float32(squeeze[1])).astype(np.float32)
    Pp = oracle.integrate_pred(X, V, tuple(sp.gravity), sp.dt)
    Pp, V2 = oracle.boundary(Pp, V, sp.grid, 0.0)
    return sp, X, V, Pp


def grid_of(oracle, Pp, sp):
    keys, idx = oracle.hash_keys(Pp, sp.grid)
    keyS, idxS = oracle.sort_pairs(keys, idx)
    start, end = oracle.cell_ranges(keyS, sp.grid.num_cells)
    uniq, off, _ = oracle.cell_compact(keyS)
    return dict(keys=keys, idx=idx, keyS=keyS, idxS=idxS, start=start, end=end, uniq=uniq, off=off)


def max_abs(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)))) if a.size else 0.0
