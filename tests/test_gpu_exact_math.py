"""The pair arithmetic's unscaled IEEE sequences (csrc/pbf_math.hpp): the same bits as
sqrtf and division wherever the kernels use them, checked on the device; the scenes of
the parity tests take that form, and parameters outside its range take the general
one and still match the oracle bit for bit."""
import ctypes

import numpy as np
import pytest

from pbfx import _lib
from pbfx import params as P

pytestmark = pytest.mark.gpu

SCENES = {
    "dam_break_1m": P.dam_break_1m,
    "dam_break_32k": P.dam_break_32k,
    "small_block": P.small_block,
    "splash_8m": P.splash_8m,
}


def _unscaled(sp):
    return _lib.lib().pbfDebugPairUnscaled(P.make_device_params(sp))


def test_unscaled_sequences_match_ieee(dev):
    out = (ctypes.c_ulonglong * 5)()
    samples = 1 << 30
    rc = _lib.lib().pbfDebugCheckUnscaledMath(samples, 0x9E3779B97F4A7C15, ctypes.addressof(out))
    assert rc == 0, _lib.lib().pbfGetLastError()
    names = ("sqrt mismatches (all floats >= 2^-96)", "tiny-input roots out of [0, 2^-47]",
             f"division mismatches ({samples} pairs)", "zero numerators not giving +0", "integer max mismatches")
    print(dict(zip(names, list(out))))
    assert list(out) == [0, 0, 0, 0, 0], dict(zip(names, list(out)))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_scenes_take_the_unscaled_form(dev, name):
    assert _unscaled(P.build_sim_params(SCENES[name]())) == 1


@pytest.mark.parametrize("field,value", [("grad_r_eps", 0.0), ("grad_r_eps", 1e-30)])
def test_out_of_range_parameters_take_the_general_form(dev, field, value):
    sp = P.build_sim_params(P.small_block())
    setattr(sp.pbf, field, value)
    assert _unscaled(sp) == 0


def test_general_form_matches_oracle(oracle, dev):
    """grad_r_eps = 1e-30 leaves the lattice's physics unchanged but takes the kernels off
    the unscaled form (max(r, eps) no longer hides tiny roots): both forms agree with the
    oracle, and with each other, bit for bit."""
    from pbfx.sim import Simulator
    states = []
    for eps in (1e-6, 1e-30):
        scene = P.small_block((14, 14, 14), table_log2=16)
        sp = P.build_sim_params(scene)
        sp.pbf.grad_r_eps = eps
        assert _unscaled(sp) == (1 if eps == 1e-6 else 0)
        sim = Simulator(sp, grid="dense", table_log2=scene.table_log2)
        try:
            sim.seed_scene(scene)
            sim.step(4)
            X, V = sim.state()
        finally:
            sim.close()
        X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
        eX, eV = oracle.run(X0, np.zeros_like(X0), sp, 4)
        np.testing.assert_array_equal(X, eX)
        np.testing.assert_array_equal(V, eV)
        states.append((X, V))
    np.testing.assert_array_equal(states[0][0], states[1][0])


# s_corr settings outside the common one (exponent 4 with a floor): the kernels take the
# general pair form, whose branches are decided per pair
SCORR_VARIANTS = {"scorr_off": {"scorr_enable": 0}, "n3": {"scorr_n": 3.0}, "n1": {"scorr_n": 1.0},
                  "no_floor": {"scorr_min": 0.0}}


@pytest.mark.parametrize("path", ["reordered", "reference"])
@pytest.mark.parametrize("variant", sorted(SCORR_VARIANTS))
def test_general_pair_form_matches_oracle(oracle, dev, variant, path):
    from pbfx.sim import Simulator
    scene = P.small_block((12, 12, 12), table_log2=16)
    sp = P.build_sim_params(scene)
    for k, v in SCORR_VARIANTS[variant].items():
        setattr(sp.pbf, k, v)
    sim = Simulator(sp, grid="dense", table_log2=scene.table_log2, path=path)
    try:
        sim.seed_scene(scene)
        sim.step(4)
        X, V = sim.state()
    finally:
        sim.close()
    X0 = P.seed_scene_reference(scene, oracle.seed_lattice)
    eX, eV = oracle.run(X0, np.zeros_like(X0), sp, 4)
    np.testing.assert_array_equal(X, eX)
    np.testing.assert_array_equal(V, eV)


def _ratio_record(sp):
    out = (ctypes.c_float * 12)()
    assert _lib.lib().pbfDebugScorrRatio(P.make_device_params(sp), ctypes.addressof(out)) == 0
    return np.array(list(out), np.float32)


@pytest.mark.parametrize("name", sorted(SCENES))
def test_scenes_take_the_two_instruction_ratio(dev, name):
    """Every scene's s_corr ratio passes the exhaustive check (fma(W, zh, W * zl) gives the
    division's s_corr term for every float W from +0 to poly6 h^6), and the record holds
    zh = RN(1 / w_q) and zl = RN(1 / w_q - zh) for the scene's w_q = poly6 (h^2 - dq^2)^3."""
    sp = P.build_sim_params(SCENES[name]())
    zh, zl, ok, wq, poly6, h2, negk, smin = _ratio_record(sp)[:8]
    f = np.float32
    hr2q = f(f(sp.kernel.h2) - f(f(f(sp.pbf.scorr_dq_h) * f(sp.kernel.h)) ** 2))
    assert wq == f(f(f(f(sp.kernel.poly6) * hr2q) * hr2q) * hr2q)
    assert ok == 1.0
    # zh is the float nearest 1 / w_q: the exact error |1 - w_q zh| / w_q is at most half
    # an ulp of zh (products of two floats are exact in float64)
    e = abs(1.0 - float(wq) * float(zh)) / float(wq)
    assert e <= 0.5 * float(np.spacing(zh))
    assert zl == f((1.0 - float(wq) * float(zh)) / float(wq))
    assert (poly6, h2, negk, smin) == (f(sp.kernel.poly6), f(sp.kernel.h2), f(-sp.pbf.scorr_k), f(sp.pbf.scorr_min))


@pytest.mark.parametrize("name", sorted(SCENES))
def test_scenes_take_the_rsq_started_reciprocal(dev, name):
    """Every scene's spiky coefficient passes the exhaustive check: with the reciprocal of
    max(r, eps) started from the square root's rsq (min(rsq(r2), 1/eps)) it equals the one
    started from v_rcp for every float r2 from +0 to h2, and the record holds the
    constants it was made for."""
    sp = P.build_sim_params(SCENES[name]())
    rec = _ratio_record(sp)
    rsq_ok, h, eps, neg3 = rec[8:]
    f = np.float32
    assert rsq_ok == 1.0
    assert (h, eps, neg3) == (f(sp.kernel.h), f(sp.pbf.grad_r_eps), f(f(-3.0) * f(sp.kernel.spiky)))


@pytest.mark.parametrize("field,value", [("scorr_enable", 0), ("scorr_n", 3.0), ("scorr_min", 0.0)])
def test_ratio_check_declines_outside_the_common_form(dev, field, value):
    """Outside the delta kernels' common form (s_corr off, another exponent, no floor below
    zero) the check records no two-instruction ratio."""
    sp = P.build_sim_params(P.small_block())
    setattr(sp.pbf, field, value)
    rec = _ratio_record(sp)
    assert rec[2] == 0.0 and rec[8] == 0.0
