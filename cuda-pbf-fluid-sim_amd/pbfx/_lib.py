"""Loader and ctypes signatures of libpbfx.so (the gfx950 HIP library).

There is no CPU fallback: if the library is missing or fails to load, every
entry point raises.
"""
import ctypes
import os

from .params import DeviceParams, Float3, GridBounds, KernelCoeffs, SimParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbfx.so")

_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_I = ctypes.c_int
_F = ctypes.c_float

# Every C-ABI entry point the library exports, with (restype, argtypes).
SIGNATURES = {
    # drop-in launchers (include/pbf_hip.h)
    "LaunchIntegratePred": (None, [_P, _P, _P, Float3, _F, _U32, _P]),
    "LaunchBoundary": (None, [_P, _P, GridBounds, _F, _U32, _P]),
    "LaunchVelocity": (None, [_P, _P, _P, _F, _U32, _P]),
    "LaunchHashKeys": (None, [_P, _P, _P, GridBounds, _U32, _P]),
    "LaunchSortPairsQuery": (None, [_P, _P, _P, _P, _P, _U32, _P]),
    "LaunchSortPairs": (None, [_P, ctypes.c_size_t, _P, _P, _P, _P, _U32, _P]),
    "LaunchCellRanges": (None, [_P, _P, _P, _U32, _U32, _P]),
    "EnsureCellCompactScratch": (ctypes.c_bool, [_U32, _U32]),
    "EnsureDropInStaging": (ctypes.c_bool, [_U32, _P]),
    "ReleaseDropInStaging": (None, [_P]),
    "LaunchCellRangesCompact": (None, [_P, _P, _P, _P, _U32, _P]),
    "LaunchCellRangesCompactNoAlloc": (None, [_P, _P, _P, _P, _U32, _P]),
    "LaunchLambda": (None, [_P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchLambdaCompact": (None, [_P, _P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchDeltaApply": (None, [_P, _P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchDeltaApplyCompact": (None, [_P, _P, _P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchXSPH": (None, [_P, _P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchXSPHCompact": (None, [_P, _P, _P, _P, _P, _P, _P, _P, DeviceParams, _U32, _P]),
    "LaunchComputeStats": (ctypes.c_bool, [_P, _P, _P, _P, _P, GridBounds, KernelCoeffs, _F, _U32, _U32, _U32,
                                            _P, _P, _P, _P, _P]),
    "LaunchComputeStatsBruteforce": (ctypes.c_bool, [_P, _P, KernelCoeffs, _F, _U32, _U32, _U32, _P, _P, _P, _P,
                                                      _P]),
    "pbfGetLastError": (_I, []),
    "pbfGetLastErrorString": (ctypes.c_char_p, []),
    "pbfSortPairsBits": (None, [_P, ctypes.POINTER(ctypes.c_size_t), _P, _P, _P, _P, _U32, _I, _I, _P]),
    # headless simulator (include/pbf_sim.h)
    "pbfSimCreate": (_P, [ctypes.POINTER(SimParams), _I, _I, _I, _I]),
    "pbfSimDestroy": (None, [_P]),
    "pbfSimSeedLattice": (_I, [_P, _U32, _U32, _U32, _U32, _F, _F, _F, _F, _F, ctypes.c_uint64]),
    "pbfSimSetState": (_I, [_P, _P, _P, _U32, _I]),
    "pbfSimGetState": (_I, [_P, _P, _P, _I]),
    "pbfSimStep": (_I, [_P, _I]),
    "pbfSimSetParams": (_I, [_P, ctypes.POINTER(SimParams)]),
    "pbfSimGetParams": (_I, [_P, ctypes.POINTER(SimParams)]),
    "pbfSimSynchronize": (_I, [_P]),
    "pbfSimStream": (_P, [_P]),
    "pbfSimDevicePositions": (_P, [_P]),
    "pbfSimDeviceVelocities": (_P, [_P]),
    "pbfSimDeviceSortedKeys": (_P, [_P]),
    "pbfSimDeviceSortedIndices": (_P, [_P]),
    "pbfSimProfilePhases": (_I, [_P, _P, _I, _P, _I]),
    "pbfSimDensityStats": (_I, [_P, _P]),
    "pbfSimHashOverflow": (_I, [_P]),
    "pbfSimAdjointInit": (_I, [_P, _I]),
    "pbfSimStepRecord": (_I, [_P, _I]),
    "pbfSimBackwardStep": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "pbfSimBackwardStepScaled": (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    "pbfEnsembleCreate": (_P, [ctypes.POINTER(SimParams), _P, _U32, _I, _I, _I]),
    "pbfEnsembleSeedLattice": (_I, [_P, _U32, _U32, _U32, _U32, _F, _F, _F, _F, _F, ctypes.c_uint64]),
    "pbfEnsembleGetScene": (_I, [_P, _U32, _P, _P]),
    "pbfEnsembleSceneInfo": (_I, [_P, _U32, _P, _P, _P]),
    # scene generators and external buffers
    "pbfLatticeAutoDims": (_I, [_U32, _P, _P, _F, _P, _P]),
    "pbfCubeMixCenters": (_I, [_P, _P]),
    "pbfSimSeedCubeMix": (_I, [_P, _U32, _U32, _P, _U32, _F, _F, ctypes.c_uint64]),
    "pbfSimBindExternalPositions": (_I, [_P, _P, _P, _U32]),
    "pbfSimBindExternalVelocity": (_I, [_P, _P, ctypes.c_size_t, _U32]),
    "pbfSimPublishVelocity": (_I, [_P, _U32]),
    "pbfSimSetTuning": (_I, [_P, _P]),
    "pbfSimGetTuning": (_I, [_P, _P]),
    "pbfDebugEchoDeviceParams": (None, [DeviceParams, ctypes.POINTER(DeviceParams)]),
    "pbfDebugEchoGridBounds": (None, [GridBounds, ctypes.POINTER(GridBounds)]),
    "pbfDebugEchoFloat3": (None, [Float3, ctypes.POINTER(ctypes.c_float)]),
    "pbfDebugPairUnscaled": (_I, [DeviceParams]),
    "pbfDebugCheckUnscaledMath": (_I, [ctypes.c_uint64, ctypes.c_uint64, _P]),
    "pbfDebugScorrRatio": (_I, [DeviceParams, _P]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with cuda-pbf-fluid-sim_amd/build.py "
                               "(or __graft_entry__.build())")
        # torch ships its own HIP runtime under the same soname; loading torch first makes the
        # library bind to that one, so the process holds a single HIP runtime (device pointers,
        # streams and the null-stream ordering are then shared with torch).
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class PbfError(RuntimeError):
    pass


def check(where=""):
    """Raise PbfError if the library recorded an error since the last check."""
    L = lib()
    code = L.pbfGetLastError()
    if code != 0:
        msg = L.pbfGetLastErrorString()
        raise PbfError(f"{where}: pbfx error {code} ({msg.decode() if msg else ''})")
