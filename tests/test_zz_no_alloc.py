"""LaunchCellRangesCompactNoAlloc: the compaction launcher for callers that cannot
allocate.  Its outputs must equal LaunchCellRangesCompact's (and the oracle's) bit for
bit, on two calls in a row (the second smaller, over the scratch the first left
behind), and it must refuse sizes beyond its static scratch without launching."""
import ctypes

import numpy as np
import pytest

from gpu_util import to_dev, to_host
from pbfx import _lib

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 26


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _sorted_keys(n, seed, cells):
    rng = np.random.default_rng(seed)
    return np.sort(rng.integers(0, cells, n, dtype=np.int64)).astype(np.uint32)


def _compact(fn, keys_dev, n, dev):
    import torch
    uniq = torch.full((max(n, 1),), -1, dtype=torch.int32, device=dev)
    off = torch.full((n + 1,), -1, dtype=torch.int32, device=dev)
    count = torch.full((1,), -1, dtype=torch.int32, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    fn(_p(uniq), _p(off), _p(count), _p(keys_dev), n, stream)
    _lib.check(fn.__name__)
    m = int(to_host(count, np.uint32)[0])
    return to_host(uniq, np.uint32)[:m], to_host(off, np.uint32)[:m + 1], m


def test_no_alloc_equals_allocating_launcher(oracle, dev):
    L = _lib.lib()
    for n, seed, cells in ((1_000_003, 1, 400_000), (70_001, 2, 9_000), (1, 3, 5)):
        print(f"compaction of {n} sorted keys over {cells} cells", flush=True)
        keys = _sorted_keys(n, seed, cells)
        kd = to_dev(keys, dev)
        assert L.EnsureCellCompactScratch(n, 256)
        ref = _compact(L.LaunchCellRangesCompact, kd, n, dev)
        got = _compact(L.LaunchCellRangesCompactNoAlloc, kd, n, dev)
        eu, eo, em = oracle.cell_compact(keys)
        assert got[2] == ref[2] == em
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[0], eu[:em])
        np.testing.assert_array_equal(got[1], eo[:em + 1])


def test_no_alloc_empty_and_oversized(dev):
    import torch
    L = _lib.lib()
    got = _compact(L.LaunchCellRangesCompactNoAlloc, torch.zeros(1, dtype=torch.int32, device=dev), 0, dev)
    assert got[2] == 0 and list(got[1]) == [0]
    # more keys than the static scratch holds: refused on the host, nothing launched
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.LaunchCellRangesCompactNoAlloc(None, None, None, None, MAX_KEYS + 1, stream)
    assert L.pbfGetLastError() == 1003
    assert L.pbfGetLastError() == 0
