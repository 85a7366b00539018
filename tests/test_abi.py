"""The C-ABI library loads, exports every entry point its headers declare, and
receives by-value parameter structs with the reference's layout (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from pbfx import _lib
from pbfx import params as P

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", h) for h in ("pbf_hip.h", "pbf_sim.h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"//.*", "", text)
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*:]*?\b(\w+)\s*\(", text, flags=re.M):
            if m.group(1) not in ("if", "defined"):
                names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert "LaunchLambda" in names and "pbfSimCreate" in names and len(names) >= 35
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = sorted(names - exported)
    assert not missing, missing
    # and the Python bindings know every one of them
    assert not sorted(names - set(_lib.SIGNATURES))


def test_library_loads_and_binds():
    L = _lib.lib()
    for name in _lib.SIGNATURES:
        assert getattr(L, name) is not None
    assert L.pbfGetLastError() == 0


def test_struct_sizes():
    assert ctypes.sizeof(P.SimParams) == 212
    assert ctypes.sizeof(P.DeviceParams) == 192
    assert P.DeviceParams.pbf.offset == 100 and P.DeviceParams.xsph_c.offset == 188
    assert P.SimParams.grid.offset == 168


def test_by_value_structs_round_trip():
    L = _lib.lib()
    sp = P.build_sim_params(P.dam_break_1m())
    sp.pbf.xsph_gate_enable = 1
    sp.pbf.relax_omega = 0.625
    dp = P.make_device_params(sp)
    out = P.DeviceParams()
    L.pbfDebugEchoDeviceParams(dp, ctypes.byref(out))
    assert bytes(out) == bytes(dp)
    g = P.GridBounds()
    L.pbfDebugEchoGridBounds(sp.grid, ctypes.byref(g))
    assert bytes(g) == bytes(sp.grid)
    f = (ctypes.c_float * 3)()
    L.pbfDebugEchoFloat3(P.Float3(1.5, -9.81, 3.25), f)
    assert list(f) == [1.5, np.float32(-9.81), 3.25]


def test_invalid_arguments_recorded_without_gpu():
    L = _lib.lib()
    sp = P.build_sim_params(P.small_block())
    sp.grid.dim = P.Int3(0, 0, 0)
    h = L.pbfSimCreate(ctypes.byref(sp), 0, 22, 0, 1)
    assert not h
    assert L.pbfGetLastError() != 0
    assert L.pbfGetLastError() == 0  # cleared by the read
    with pytest.raises(KeyError):
        from pbfx.sim import GRID_KINDS
        GRID_KINDS["octree"]
