"""Scene generators over the C ABI: auto-sized box lattices and CubeMix groups
(reference: Simulator::seedBoxLatticeAuto / seedCubeMix, console GenerateCubeMixCenters)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


def lattice_auto_dims(total, box_min, box_max, spacing, origin=(0.0, 0.0, 0.0)):
    """(nx, ny, nz) for about `total` particles at `spacing` in the box, and the origin
    pulled a quarter spacing inside it (reference: simulator.cpp:882-941)."""
    lo = (ctypes.c_float * 3)(*box_min)
    hi = (ctypes.c_float * 3)(*box_max)
    o = (ctypes.c_float * 3)(*origin)
    dims = (ctypes.c_uint32 * 3)()
    if _lib.lib().pbfLatticeAutoDims(int(total), lo, hi, float(spacing), o, dims) != 0:
        raise _lib.PbfError("pbfLatticeAutoDims: spacing must be > 0")
    return (dims[0], dims[1], dims[2]), (o[0], o[1], o[2])


class CubeMixLayoutC(ctypes.Structure):
    _fields_ = [("groupCount", ctypes.c_uint32), ("layers", ctypes.c_uint32), ("groupSpacing", ctypes.c_float),
                ("layerSpacing", ctypes.c_float), ("baseHeight", ctypes.c_float), ("layerJitter", ctypes.c_int),
                ("layerJitterAmp", ctypes.c_float), ("layerJitterSeed", ctypes.c_uint32)]


@dataclass
class CubeMixLayout:
    """The reference's CubeMix settings (console.h:184-208), in world units."""
    group_count: int = 8
    layers: int = 2
    group_spacing: float = 200.0
    layer_spacing: float = 200.0
    base_height: float = 200.0
    layer_jitter: bool = False
    layer_jitter_amp: float = 0.0
    layer_jitter_seed: int = 0x13579BDF

    def as_c(self):
        return CubeMixLayoutC(self.group_count, self.layers, self.group_spacing, self.layer_spacing,
                              self.base_height, int(self.layer_jitter), self.layer_jitter_amp,
                              self.layer_jitter_seed)


def cube_mix_centers(layout: CubeMixLayout):
    """(groups, 3) float32 group centres (reference: console.cpp:277-316)."""
    out = np.zeros((layout.group_count, 3), np.float32)
    c = layout.as_c()
    if _lib.lib().pbfCubeMixCenters(ctypes.byref(c), out.ctypes.data) != 0:
        raise _lib.PbfError("pbfCubeMixCenters failed")
    return out
