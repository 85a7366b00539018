"""Parity of every drop-in launcher (include/pbf_hip.h) with the CPU oracle.

Integer work (keys, sort, ranges, compaction) must be bit-exact; floating-point
outputs must match to 1e-5*h (positions) / 1e-5*h/dt (velocities), and in
practice match bit for bit except where powf is involved.
"""
import numpy as np
import pytest

from gpu_util import H_TOL, grid_of, max_abs, predicted_state, to_dev, to_host
from pbfx import params as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    from pbfx import launch
    return launch


@pytest.fixture(params=[False, True], ids=["indexed", "staged"])
def staged(request, dev):
    """Run the neighbour launchers as the reference's phases would (reading through
    indicesSorted), or with staging buffers reserved for the stream (EnsureDropInStaging):
    both must give the oracle's bits."""
    import ctypes
    import torch
    from pbfx import _lib
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if request.param:
        assert _lib.lib().EnsureDropInStaging(1 << 20, s)
    yield request.param
    if request.param:
        torch.cuda.synchronize()
        _lib.lib().ReleaseDropInStaging(s)


@pytest.mark.parametrize("n", [0, 1, 1000, 65536 + 17])
def test_integrate_boundary_velocity(oracle, dev, L, n):
    import torch
    rng = np.random.default_rng(n)
    X = rng.uniform(-0.1, 1.1, (n, 4)).astype(np.float32)
    V = rng.normal(0, 3, (n, 4)).astype(np.float32)
    sp = P.build_sim_params(P.small_block())
    g, dt = tuple(sp.gravity), sp.dt
    tX, tV = to_dev(X, dev), to_dev(V, dev)
    tP = torch.zeros_like(tX)
    L.integrate_pred(tX, tV, tP, g, dt)
    Pp = oracle.integrate_pred(X, V, g, dt)
    np.testing.assert_array_equal(to_host(tP), Pp)
    L.boundary(tP, tV, sp.grid, 0.3)
    Pc, Vc = oracle.boundary(Pp, V, sp.grid, 0.3)
    np.testing.assert_array_equal(to_host(tP), Pc)
    np.testing.assert_array_equal(to_host(tV), Vc)
    tV2 = torch.zeros_like(tX)
    L.velocity(tV2, tX, tP, np.float32(1.0) / np.float32(dt))
    np.testing.assert_array_equal(to_host(tV2), oracle.velocity(np.zeros_like(X), X, Pc, np.float32(1.0) / np.float32(dt)))


def test_hash_keys_edge_cases(oracle, dev, L):
    import torch
    sp = P.build_sim_params(P.small_block())
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.3, 1.3, (4099, 4)).astype(np.float32)
    cs = np.float32(sp.grid.cellSize)
    pts[:40, 0] = (np.arange(40, dtype=np.float32) * cs).astype(np.float32)  # exactly on cell faces
    pts[40, :3] = [1.0, 1.0, 1.0]
    pts[41, :3] = [np.nan, 0.5, np.inf]
    pts[42, :3] = [-np.inf, 3e38, -3e38]
    keys = torch.zeros(len(pts), dtype=torch.int32, device=dev)
    idx = torch.zeros_like(keys)
    L.hash_keys(keys, idx, to_dev(pts, dev), sp.grid)
    ek, ei = oracle.hash_keys(pts, sp.grid)
    np.testing.assert_array_equal(to_host(keys, np.uint32), ek)
    np.testing.assert_array_equal(to_host(idx, np.uint32), ei)


@pytest.mark.parametrize("n,kmax", [(1, 10), (255, 3), (1024, 1 << 32), (4097, 100), (100003, 1 << 21),
                                    (1 << 20, 1 << 32), (3_000_017, 1 << 25)])
def test_sort_pairs_full_32_bits(oracle, dev, L, n, kmax):
    import torch
    rng = np.random.default_rng(n)
    keys = rng.integers(0, kmax, n, dtype=np.uint64).astype(np.uint32)
    vals = rng.permutation(n).astype(np.uint32)
    ko = torch.zeros(n, dtype=torch.int32, device=dev)
    vo = torch.zeros_like(ko)
    tk, tv = to_dev(keys, dev), to_dev(vals, dev)
    L.sort_pairs(tk, ko, tv, vo)
    order = np.lexsort((np.arange(n), keys))
    np.testing.assert_array_equal(to_host(ko, np.uint32), keys[order])
    np.testing.assert_array_equal(to_host(vo, np.uint32), vals[order])
    # inputs untouched
    np.testing.assert_array_equal(to_host(tk, np.uint32), keys)
    if n <= 100003:
        ek, ev = oracle.sort_pairs(keys, vals)
        np.testing.assert_array_equal(to_host(vo, np.uint32), ev)


@pytest.mark.parametrize("bits", [1, 7, 13, 21, 25])
def test_sort_pairs_bits_identity_values(dev, L, bits):
    import torch
    n = 300_001
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    ko = torch.zeros(n, dtype=torch.int32, device=dev)
    vo = torch.zeros_like(ko)
    L.sort_pairs_bits(to_dev(keys, dev), ko, None, vo, 0, bits)
    order = np.lexsort((np.arange(n), keys))
    np.testing.assert_array_equal(to_host(ko, np.uint32), keys[order])
    np.testing.assert_array_equal(to_host(vo, np.uint32), order.astype(np.uint32))


def test_sort_pairs_zero_length(dev, L):
    import torch
    z = torch.zeros(0, dtype=torch.int32, device=dev)
    L.sort_pairs(z, z.clone(), z.clone(), z.clone())


@pytest.mark.parametrize("case", ["lattice", "one_cell", "random"])
def test_cell_ranges_and_compact(oracle, dev, L, case):
    import torch
    sp = P.build_sim_params(P.small_block())
    C = sp.grid.num_cells
    rng = np.random.default_rng(11)
    if case == "lattice":
        _, _, _, Pp = predicted_state(oracle, P.small_block((16, 16, 16)))
        keys, _ = oracle.hash_keys(Pp, sp.grid)
        keyS = np.sort(keys)
    elif case == "one_cell":
        keyS = np.full(5000, 1234, np.uint32)
    else:
        keyS = np.sort(rng.integers(0, C, 200_000)).astype(np.uint32)
    n = len(keyS)
    tks = to_dev(keyS, dev)
    start = torch.zeros(C, dtype=torch.int32, device=dev)
    end = torch.zeros_like(start)
    L.cell_ranges(start, end, tks, C)
    es, ee = oracle.cell_ranges(keyS, C)
    np.testing.assert_array_equal(to_host(start, np.uint32), es)
    np.testing.assert_array_equal(to_host(end, np.uint32), ee)
    uniq = torch.zeros(n, dtype=torch.int32, device=dev)
    off = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    L.cell_ranges_compact(uniq, off, cnt, tks)
    eu, eo, m = oracle.cell_compact(keyS)
    assert int(to_host(cnt)[0]) == m
    np.testing.assert_array_equal(to_host(uniq, np.uint32)[:m], eu)
    np.testing.assert_array_equal(to_host(off, np.uint32)[: m + 1], eo)


def test_compact_zero_length(dev, L):
    import torch
    off = torch.full((1,), 77, dtype=torch.int32, device=dev)
    cnt = torch.full((1,), 77, dtype=torch.int32, device=dev)
    z = torch.zeros(0, dtype=torch.int32, device=dev)
    L.cell_ranges_compact(z, off, cnt, z)
    assert int(to_host(off)[0]) == 0 and int(to_host(cnt)[0]) == 0


def _solver_inputs(oracle, dev, scene, steps=0, squeeze=None):
    sp, X, V, Pp = predicted_state(oracle, scene, steps, squeeze)
    g = grid_of(oracle, Pp, sp)
    t = {k: to_dev(v, dev) for k, v in g.items()}
    t["cnt"] = to_dev(np.array([len(g["uniq"])], np.uint32), dev)
    return sp, X, V, Pp, g, t


@pytest.mark.parametrize("grid", ["dense", "compact"])
@pytest.mark.parametrize("scene_name,steps,squeeze", [("block16", 0, None), ("block16", 0, (0.8, 0.1)),
                                                      ("block20", 15, None)])
def test_lambda_delta_xsph(oracle, dev, L, staged, grid, scene_name, steps, squeeze):
    import torch
    n = 16 if scene_name == "block16" else 20
    scene = P.small_block((n, n, n))
    sp, X, V, Pp, g, t = _solver_inputs(oracle, dev, scene, steps, squeeze)
    dp = P.make_device_params(sp)
    N = len(Pp)
    tP = to_dev(Pp, dev)
    lam = torch.zeros(N, dtype=torch.float32, device=dev)
    if grid == "dense":
        L.lambda_dense(lam, tP, t["idxS"], t["keyS"], t["start"], t["end"], dp)
        elam = oracle.lambda_dense(Pp, g["idxS"], g["keyS"], g["start"], g["end"], dp)
    else:
        L.lambda_compact(lam, tP, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        elam = oracle.lambda_compact(Pp, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
    glam = to_host(lam)
    # lambda involves no powf: same operations, same order -> same bits
    np.testing.assert_array_equal(glam, elam)

    delta = torch.zeros_like(tP)
    if grid == "dense":
        L.delta_apply_dense(tP, delta, lam, t["idxS"], t["keyS"], t["start"], t["end"], dp)
        eP, edelta = oracle.delta_apply_dense(Pp, elam, g["idxS"], g["keyS"], g["start"], g["end"], dp)
    else:
        L.delta_apply_compact(tP, delta, lam, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eP, edelta = oracle.delta_apply_compact(Pp, elam, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
    gP = to_host(tP)
    h = sp.kernel.h
    assert max_abs(gP[:, :3], eP[:, :3]) <= H_TOL * h
    np.testing.assert_array_equal(gP[:, 3], eP[:, 3])
    assert max_abs(to_host(delta)[:, :3], edelta[:, :3]) <= H_TOL * h

    rng = np.random.default_rng(3)
    Vin = rng.normal(0, 0.5, (N, 4)).astype(np.float32)
    tVin = to_dev(Vin, dev)
    tVout = torch.zeros_like(tVin)
    if grid == "dense":
        L.xsph_dense(tVout, tVin, to_dev(eP, dev), t["idxS"], t["keyS"], t["start"], t["end"], dp)
        eV = oracle.xsph_dense(Vin, eP, g["idxS"], g["keyS"], g["start"], g["end"], dp)
    else:
        L.xsph_compact(tVout, tVin, to_dev(eP, dev), t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eV = oracle.xsph_compact(Vin, eP, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
    np.testing.assert_array_equal(to_host(tVout), eV)


@pytest.mark.parametrize("grid", ["dense", "compact"])
def test_lambda_delta_xsph_reach_two(oracle, dev, L, staged, grid):
    # cells of 0.6 h: every neighbour row spans five cells and there are 25 rows
    import torch
    scene = P.small_block((12, 12, 12))
    sp, X, V, Pp = predicted_state(oracle, scene, 0, (0.7, 0.15))
    cell = np.float32(0.6) * np.float32(sp.kernel.h)
    sp.grid.cellSize = float(cell)
    sp.grid.dim = P.Int3(*P.compute_dims(scene.box_min, scene.box_max, cell))
    g = grid_of(oracle, Pp, sp)
    t = {k: to_dev(v, dev) for k, v in g.items()}
    t["cnt"] = to_dev(np.array([len(g["uniq"])], np.uint32), dev)
    dp = P.make_device_params(sp)
    N = len(Pp)
    tP = to_dev(Pp, dev)
    lam = torch.zeros(N, dtype=torch.float32, device=dev)
    delta = torch.zeros_like(tP)
    rng = np.random.default_rng(5)
    Vin = rng.normal(0, 0.5, (N, 4)).astype(np.float32)
    tVin = to_dev(Vin, dev)
    tVout = torch.zeros_like(tVin)
    if grid == "dense":
        look, elook = (t["start"], t["end"]), (g["start"], g["end"])
        L.lambda_dense(lam, tP, t["idxS"], t["keyS"], *look, dp)
        elam = oracle.lambda_dense(Pp, g["idxS"], g["keyS"], *elook, dp)
        L.delta_apply_dense(tP, delta, lam, t["idxS"], t["keyS"], *look, dp)
        eP, edelta = oracle.delta_apply_dense(Pp, elam, g["idxS"], g["keyS"], *elook, dp)
        L.xsph_dense(tVout, tVin, tP, t["idxS"], t["keyS"], *look, dp)
        eV = oracle.xsph_dense(Vin, eP, g["idxS"], g["keyS"], *elook, dp)
    else:
        L.lambda_compact(lam, tP, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        elam = oracle.lambda_compact(Pp, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
        L.delta_apply_compact(tP, delta, lam, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eP, edelta = oracle.delta_apply_compact(Pp, elam, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
        L.xsph_compact(tVout, tVin, tP, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eV = oracle.xsph_compact(Vin, eP, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
    np.testing.assert_array_equal(to_host(lam), elam)
    np.testing.assert_array_equal(to_host(tP), eP)
    np.testing.assert_array_equal(to_host(delta), edelta)
    np.testing.assert_array_equal(to_host(tVout), eV)


@pytest.mark.parametrize("grid", ["dense", "compact"])
def test_lambda_delta_xsph_partial_wave(oracle, dev, L, staged, grid):
    # 1001 particles (13 x 11 x 7), squeezed: the last wave has 41 lanes
    import torch
    scene = P.small_block((13, 11, 7))
    sp, X, V, Pp = predicted_state(oracle, scene, 0, (0.7, 0.15))
    g = grid_of(oracle, Pp, sp)
    t = {k: to_dev(v, dev) for k, v in g.items()}
    t["cnt"] = to_dev(np.array([len(g["uniq"])], np.uint32), dev)
    dp = P.make_device_params(sp)
    N = len(Pp)
    assert N % 64 != 0
    tP = to_dev(Pp, dev)
    lam = torch.zeros(N, dtype=torch.float32, device=dev)
    delta = torch.zeros_like(tP)
    Vin = np.random.default_rng(7).normal(0, 0.5, (N, 4)).astype(np.float32)
    tVout = torch.zeros_like(tP)
    if grid == "dense":
        look, elook = (t["start"], t["end"]), (g["start"], g["end"])
        L.lambda_dense(lam, tP, t["idxS"], t["keyS"], *look, dp)
        elam = oracle.lambda_dense(Pp, g["idxS"], g["keyS"], *elook, dp)
        L.delta_apply_dense(tP, delta, lam, t["idxS"], t["keyS"], *look, dp)
        eP, edelta = oracle.delta_apply_dense(Pp, elam, g["idxS"], g["keyS"], *elook, dp)
        L.xsph_dense(tVout, to_dev(Vin, dev), tP, t["idxS"], t["keyS"], *look, dp)
        eV = oracle.xsph_dense(Vin, eP, g["idxS"], g["keyS"], *elook, dp)
    else:
        L.lambda_compact(lam, tP, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        elam = oracle.lambda_compact(Pp, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
        L.delta_apply_compact(tP, delta, lam, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eP, edelta = oracle.delta_apply_compact(Pp, elam, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
        L.xsph_compact(tVout, to_dev(Vin, dev), tP, t["idxS"], t["keyS"], t["uniq"], t["off"], t["cnt"], dp)
        eV = oracle.xsph_compact(Vin, eP, g["idxS"], g["keyS"], g["uniq"], g["off"], dp)
    np.testing.assert_array_equal(to_host(lam), elam)
    np.testing.assert_array_equal(to_host(tP), eP)
    np.testing.assert_array_equal(to_host(delta), edelta)
    np.testing.assert_array_equal(to_host(tVout), eV)


def test_lambda_neighbor_cap(oracle, dev, L, staged):
    import torch
    scene = P.small_block((12, 12, 12))
    sp, X, V, Pp, g, t = _solver_inputs(oracle, dev, scene, 0, (0.6, 0.2))
    sp.maxNeighbors = 7
    dp = P.make_device_params(sp)
    lam = torch.zeros(len(Pp), dtype=torch.float32, device=dev)
    L.lambda_dense(lam, to_dev(Pp, dev), t["idxS"], t["keyS"], t["start"], t["end"], dp)
    elam = oracle.lambda_dense(Pp, g["idxS"], g["keyS"], g["start"], g["end"], dp)
    np.testing.assert_array_equal(to_host(lam), elam)


def test_staging_smaller_than_the_launch_falls_back(oracle, dev, L):
    """Staging reserved for fewer particles than a launch holds, grown, then released:
    every launch gives the oracle's bits (the walk reads through indicesSorted whenever
    the stream's staging is too small or absent)."""
    import ctypes
    import torch
    from pbfx import _lib
    scene = P.small_block((12, 12, 12))
    sp, X, V, Pp, g, t = _solver_inputs(oracle, dev, scene, 0, (0.8, 0.1))
    dp = P.make_device_params(sp)
    N = len(Pp)
    elam = oracle.lambda_dense(Pp, g["idxS"], g["keyS"], g["start"], g["end"], dp)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    try:
        for reserve in (N // 2, N, 0):
            if reserve:
                assert _lib.lib().EnsureDropInStaging(reserve, s)
            else:
                torch.cuda.synchronize()
                _lib.lib().ReleaseDropInStaging(s)
            lam = torch.zeros(N, dtype=torch.float32, device=dev)
            L.lambda_dense(lam, to_dev(Pp, dev), t["idxS"], t["keyS"], t["start"], t["end"], dp)
            np.testing.assert_array_equal(to_host(lam), elam)
    finally:
        torch.cuda.synchronize()
        _lib.lib().ReleaseDropInStaging(s)


def test_xsph_rejects_in_place(dev, L):
    import torch
    from pbfx import PbfError
    sp = P.build_sim_params(P.small_block())
    dp = P.make_device_params(sp)
    v = torch.zeros((8, 4), dtype=torch.float32, device=dev)
    z = torch.zeros(8, dtype=torch.int32, device=dev)
    c = torch.zeros(sp.grid.num_cells, dtype=torch.int32, device=dev)
    with pytest.raises(PbfError):
        L.xsph_dense(v, v, v, z, z, c, c, dp)
