"""Python front end of the drop-in launchers (include/pbf_hip.h) over torch tensors.

torch provides device memory and the stream; every call goes through the C ABI
to the HIP kernels.  Tensors: positions/velocities float32 (N, 4) on the GPU,
keys/indices int32 (N,) reinterpreted as uint32.  Each wrapper enqueues on the
current torch stream and raises if the library recorded an error.
"""
import ctypes

import torch

from . import _lib
from .params import Float3


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _n(t):
    return int(t.shape[0])


def integrate_pred(pos, vel, pos_pred, gravity, dt):
    _lib.lib().LaunchIntegratePred(_p(pos), _p(vel), _p(pos_pred), Float3(*gravity), dt, _n(pos), _stream())
    _lib.check("LaunchIntegratePred")


def boundary(pos_pred, vel, grid, restitution):
    _lib.lib().LaunchBoundary(_p(pos_pred), _p(vel), grid, restitution, _n(pos_pred), _stream())
    _lib.check("LaunchBoundary")


def velocity(vel, pos, pos_pred, inv_dt):
    _lib.lib().LaunchVelocity(_p(vel), _p(pos), _p(pos_pred), inv_dt, _n(vel), _stream())
    _lib.check("LaunchVelocity")


def hash_keys(keys, idx, pos, grid):
    _lib.lib().LaunchHashKeys(_p(keys), _p(idx), _p(pos), grid, _n(pos), _stream())
    _lib.check("LaunchHashKeys")


def sort_pairs_temp_bytes(n):
    b = ctypes.c_size_t(0)
    _lib.lib().LaunchSortPairsQuery(ctypes.byref(b), None, None, None, None, n, _stream())
    return b.value


def sort_pairs(keys_in, keys_out, vals_in, vals_out, temp=None):
    n = _n(keys_in)
    if temp is None:
        temp = torch.empty(max(sort_pairs_temp_bytes(n), 1), dtype=torch.uint8, device=keys_in.device)
    _lib.lib().LaunchSortPairs(_p(temp), temp.numel(), _p(keys_in), _p(keys_out), _p(vals_in), _p(vals_out), n,
                               _stream())
    _lib.check("LaunchSortPairs")


def sort_pairs_bits(keys_in, keys_out, vals_in, vals_out, begin_bit, end_bit):
    """Stable sort over key bits [begin_bit, end_bit); vals_in None = identity values."""
    n = _n(keys_in)
    b = ctypes.c_size_t(0)
    L = _lib.lib()
    L.pbfSortPairsBits(None, ctypes.byref(b), None, None, None, None, n, begin_bit, end_bit, _stream())
    temp = torch.empty(max(b.value, 1), dtype=torch.uint8, device=keys_in.device)
    L.pbfSortPairsBits(_p(temp), ctypes.byref(b), _p(keys_in), _p(keys_out), _p(vals_in), _p(vals_out), n,
                       begin_bit, end_bit, _stream())
    _lib.check("pbfSortPairsBits")


def cell_ranges(start, end, keys_sorted, num_cells):
    _lib.lib().LaunchCellRanges(_p(start), _p(end), _p(keys_sorted), _n(keys_sorted), num_cells, _stream())
    _lib.check("LaunchCellRanges")


def cell_ranges_compact(uniq, offsets, count, keys_sorted):
    n = _n(keys_sorted)
    if not _lib.lib().EnsureCellCompactScratch(n, 256):
        _lib.check("EnsureCellCompactScratch")
        raise _lib.PbfError("EnsureCellCompactScratch failed")
    _lib.lib().LaunchCellRangesCompact(_p(uniq), _p(offsets), _p(count), _p(keys_sorted), n, _stream())
    _lib.check("LaunchCellRangesCompact")


def lambda_dense(lam, pos_pred, idx_s, key_s, start, end, dp):
    _lib.lib().LaunchLambda(_p(lam), _p(pos_pred), _p(idx_s), _p(key_s), _p(start), _p(end), dp, _n(pos_pred),
                            _stream())
    _lib.check("LaunchLambda")


def lambda_compact(lam, pos_pred, idx_s, key_s, uniq, offsets, count, dp):
    _lib.lib().LaunchLambdaCompact(_p(lam), _p(pos_pred), _p(idx_s), _p(key_s), _p(uniq), _p(offsets), _p(count), dp,
                                   _n(pos_pred), _stream())
    _lib.check("LaunchLambdaCompact")


def delta_apply_dense(pos_pred, delta, lam, idx_s, key_s, start, end, dp):
    _lib.lib().LaunchDeltaApply(_p(pos_pred), _p(delta), _p(lam), _p(idx_s), _p(key_s), _p(start), _p(end), dp,
                                _n(pos_pred), _stream())
    _lib.check("LaunchDeltaApply")


def delta_apply_compact(pos_pred, delta, lam, idx_s, key_s, uniq, offsets, count, dp):
    _lib.lib().LaunchDeltaApplyCompact(_p(pos_pred), _p(delta), _p(lam), _p(idx_s), _p(key_s), _p(uniq),
                                       _p(offsets), _p(count), dp, _n(pos_pred), _stream())
    _lib.check("LaunchDeltaApplyCompact")


def xsph_dense(vel_out, vel_in, pos_pred, idx_s, key_s, start, end, dp):
    _lib.lib().LaunchXSPH(_p(vel_out), _p(vel_in), _p(pos_pred), _p(idx_s), _p(key_s), _p(start), _p(end), dp,
                          _n(pos_pred), _stream())
    _lib.check("LaunchXSPH")


def xsph_compact(vel_out, vel_in, pos_pred, idx_s, key_s, uniq, offsets, count, dp):
    _lib.lib().LaunchXSPHCompact(_p(vel_out), _p(vel_in), _p(pos_pred), _p(idx_s), _p(key_s), _p(uniq), _p(offsets),
                                 _p(count), dp, _n(pos_pred), _stream())
    _lib.check("LaunchXSPHCompact")


def compute_stats(pos, vel, idx_s, start, end, grid, kc, mass, sample_stride=1):
    """LaunchComputeStats: (avg neighbours, avg speed, avg sum W, avg rho); synchronises."""
    out = [ctypes.c_double() for _ in range(4)]
    ok = _lib.lib().LaunchComputeStats(_p(pos), _p(vel), _p(idx_s), _p(start), _p(end), grid, kc, mass, _n(pos),
                                       int(start.shape[0]), sample_stride, *[ctypes.byref(o) for o in out],
                                       _stream())
    _lib.check("LaunchComputeStats")
    assert ok
    return tuple(o.value for o in out)


def compute_stats_bruteforce(pos, vel, kc, mass, sample_stride=1, max_samples=None):
    out = [ctypes.c_double() for _ in range(4)]
    n = _n(pos)
    ok = _lib.lib().LaunchComputeStatsBruteforce(_p(pos), _p(vel), kc, mass, n, sample_stride,
                                                 n if max_samples is None else max_samples,
                                                 *[ctypes.byref(o) for o in out], _stream())
    _lib.check("LaunchComputeStatsBruteforce")
    assert ok
    return tuple(o.value for o in out)
