"""Scene generators (host side, no GPU): the library's auto lattice dims and CubeMix
centres against closed forms and the oracle's numpy restatement."""
import numpy as np
import pytest

from pbfx import scenes as S


def test_auto_dims_closed_forms():
    # cube-root split: 1000 -> 10^3; the origin is pulled 0.25 spacing inside the box
    dims, o = S.lattice_auto_dims(1000, (0, 0, 0), (1, 1, 1), 0.01, (-1.0, 0.5, 2.0))
    assert dims == (10, 10, 10)
    assert o == pytest.approx((0.0025, 0.5, 1.0 - 0.0025), abs=1e-7)
    # 1001: nx = ceil(cbrt) = 11, rem = 91, ny = 10, nz = 10 (1100 >= 1001)
    assert S.lattice_auto_dims(1001, (0, 0, 0), (1, 1, 1), 0.01)[0] == (11, 10, 10)
    # a box 5 spacings wide in x caps nx; the loop then grows y, then z
    assert S.lattice_auto_dims(1000, (0, 0, 0), (0.0501, 1, 1), 0.01)[0] == (5, 15, 14)
    # box too small for the request: every axis at its cap, fewer particles than asked
    assert S.lattice_auto_dims(10 ** 6, (0, 0, 0), (0.0301, 0.0301, 0.0301), 0.01)[0] == (3, 3, 3)
    # total 0 and 1 give one particle
    assert S.lattice_auto_dims(0, (0, 0, 0), (1, 1, 1), 0.01)[0] == (1, 1, 1)


@pytest.mark.parametrize("total", [1, 7, 64, 999, 4096, 12345, 262144])
@pytest.mark.parametrize("box", [((0, 0, 0), (1, 1, 1)), ((0, 0, 0), (2, 1, 1)), ((-1, 0, -1), (0.2, 3, 0.05))])
def test_auto_dims_match_oracle(oracle, total, box):
    for spacing in (0.01, 0.0125, 0.2):
        got = S.lattice_auto_dims(total, box[0], box[1], spacing, (0.1, 0.2, 0.3))
        exp = oracle.lattice_auto_dims(total, box[0], box[1], spacing, (0.1, 0.2, 0.3))
        assert got[0] == exp[0]
        assert np.array_equal(np.float32(got[1]), np.float32(exp[1]))


def test_cube_mix_centres_closed_forms():
    # the reference's defaults: 8 groups over 2 layers, 2x2 per layer
    c = S.cube_mix_centers(S.CubeMixLayout())
    exp = [(0, 200, 0), (200, 200, 0), (0, 200, 200), (200, 200, 200),
           (0, 400, 0), (200, 400, 0), (0, 400, 200), (200, 400, 200)]
    assert np.array_equal(c, np.array(exp, np.float32))
    # 5 groups over 2 layers: 2 then 3; the second layer is 2 wide
    c = S.cube_mix_centers(S.CubeMixLayout(group_count=5, layers=2, group_spacing=1, layer_spacing=0.5,
                                           base_height=0.25))
    assert np.array_equal(c, np.array([(0, .25, 0), (1, .25, 0), (0, .75, 0), (1, .75, 0), (0, .75, 1)], np.float32))
    # more layers than groups: one group per layer until the groups run out
    c = S.cube_mix_centers(S.CubeMixLayout(group_count=2, layers=3, group_spacing=1, layer_spacing=0.5,
                                           base_height=0.25))
    assert np.array_equal(c, np.array([(0, .25, 0), (0, .75, 0)], np.float32))


@pytest.mark.parametrize("groups,layers", [(1, 1), (8, 2), (13, 3), (2, 5), (512, 4)])
@pytest.mark.parametrize("jitter", [False, True])
def test_cube_mix_centres_match_oracle(oracle, groups, layers, jitter):
    lay = S.CubeMixLayout(group_count=groups, layers=layers, group_spacing=0.3, layer_spacing=0.45,
                          base_height=0.2, layer_jitter=jitter, layer_jitter_amp=0.01, layer_jitter_seed=99)
    got = S.cube_mix_centers(lay)
    exp = oracle.cube_mix_centers(groups, layers, 0.3, 0.45, 0.2, jitter, 0.01, 99)
    assert got.shape == (groups, 3)
    assert np.array_equal(got, exp)
    if jitter and layers > 1 and groups > 1:
        ys = np.unique(got[:, 1])
        assert np.all(np.abs(ys - (0.2 + 0.45 * np.arange(len(ys)))) <= 0.01 + 1e-6)
