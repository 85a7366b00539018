"""ctypes front end of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Loaded by tests/, the smoke check and bench.py's cpu_baseline leg, as the
checker.  The product package never imports this module.

Arrays are numpy: positions/velocities float32 (N, 4), keys/indices uint32 (N,).
Parameter blocks are any ctypes object laid out as the reference's SimParams /
DeviceParams / GridBounds (reference: sim/parameters.h:9-92); they are passed by
pointer.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_lib = None


def build(verbose=False):
    """Compile liboracle.so from pbf_oracle.cpp (g++, no FMA contraction, OpenMP)."""
    cmd = ["make", "-C", _HERE, "-B", "liboracle.so"]
    out = subprocess.run(cmd, capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if verbose:
        print(out.stdout)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = ctypes.c_void_p
_U32 = ctypes.c_uint32
_F = ctypes.c_float


def _declare(L):
    sig = {
        "oracle_make_kernel_coeffs": (None, [_F, _P]),
        "oracle_make_device_params": (None, [_P, _P]),
        "oracle_compute_dims": (None, [_P, _P]),
        "oracle_integrate_pred": (None, [_P, _P, _P, _F, _F, _F, _F, _U32]),
        "oracle_boundary": (None, [_P, _P, _P, _F, _U32]),
        "oracle_velocity": (None, [_P, _P, _P, _F, _U32]),
        "oracle_hash_keys": (None, [_P, _P, _P, _P, _U32]),
        "oracle_sort_pairs": (None, [_P, _P, _P, _P, _U32]),
        "oracle_cell_ranges": (None, [_P, _P, _P, _U32, _U32]),
        "oracle_cell_compact": (None, [_P, _P, _P, _P, _U32]),
        "oracle_compact_cell_range": (ctypes.c_int, [_P, _P, _U32, _U32, _P, _P]),
        "oracle_lambda_dense": (None, [_P] * 7 + [_U32]),
        "oracle_lambda_compact": (None, [_P] * 8 + [_U32]),
        "oracle_delta_apply_dense": (None, [_P] * 8 + [_U32]),
        "oracle_delta_apply_compact": (None, [_P] * 9 + [_U32]),
        "oracle_xsph_dense": (None, [_P] * 8 + [_U32]),
        "oracle_xsph_compact": (None, [_P] * 9 + [_U32]),
        "oracle_run": (ctypes.c_int, [_P, _P, _P, ctypes.c_int, ctypes.c_int]),
        "oracle_density_stats": (None, [_P, _P, _P, _P]),
        "oracle_seed_lattice": (None, [_P, _U32, _U32, _U32, _F, _F, _F, _F, _P, _F, ctypes.c_uint64]),
        "oracle_jitter_uniform": (_F, [ctypes.c_uint64, ctypes.c_uint64, _U32]),
        "oracle_pow_scorr": (_F, [_F, _F]),
        "oracle_num_threads": (ctypes.c_int, []),
        "oracle_set_num_threads": (None, [ctypes.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return ctypes.addressof(a)


def _f4(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    return a


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- parameters
def make_kernel_coeffs(h, out):
    lib().oracle_make_kernel_coeffs(float(h), _ptr(out))
    return out


def make_device_params(sp, out):
    lib().oracle_make_device_params(_ptr(sp), _ptr(out))
    return out


def compute_dims(grid):
    d = (ctypes.c_int * 3)()
    lib().oracle_compute_dims(_ptr(grid), ctypes.addressof(d))
    return tuple(d)


# ------------------------------------------------------- per-particle phases
def integrate_pred(X, V, g, dt):
    X, V = _f4(X), _f4(V)
    P = np.zeros_like(X)
    lib().oracle_integrate_pred(_ptr(X), _ptr(V), _ptr(P), g[0], g[1], g[2], dt, len(X))
    return P


def boundary(P, V, grid, e):
    P, V = _f4(P).copy(), _f4(V).copy()
    lib().oracle_boundary(_ptr(P), _ptr(V), _ptr(grid), e, len(P))
    return P, V


def velocity(V, X, P, inv_dt):
    V = _f4(V).copy()
    X, P = _f4(X), _f4(P)
    lib().oracle_velocity(_ptr(V), _ptr(X), _ptr(P), inv_dt, len(V))
    return V


# ----------------------------------------------------------------- grid build
def hash_keys(P, grid):
    P = _f4(P)
    n = len(P)
    keys = np.zeros(n, np.uint32)
    idx = np.zeros(n, np.uint32)
    lib().oracle_hash_keys(_ptr(keys), _ptr(idx), _ptr(P), _ptr(grid), n)
    return keys, idx


def sort_pairs(keys, vals):
    keys, vals = _u32(keys), _u32(vals)
    ko = np.zeros_like(keys)
    vo = np.zeros_like(vals)
    lib().oracle_sort_pairs(_ptr(keys), _ptr(ko), _ptr(vals), _ptr(vo), len(keys))
    return ko, vo


def cell_ranges(keyS, num_cells):
    keyS = _u32(keyS)
    start = np.zeros(num_cells, np.uint32)
    end = np.zeros(num_cells, np.uint32)
    lib().oracle_cell_ranges(_ptr(start), _ptr(end), _ptr(keyS), len(keyS), num_cells)
    return start, end


def cell_compact(keyS):
    keyS = _u32(keyS)
    n = len(keyS)
    uniq = np.zeros(max(n, 1), np.uint32)
    off = np.zeros(max(n, 1) + 1, np.uint32)
    count = np.zeros(1, np.uint32)
    lib().oracle_cell_compact(_ptr(uniq), _ptr(off), _ptr(count), _ptr(keyS), n)
    m = int(count[0])
    return uniq[:m].copy(), off[: m + 1].copy(), m


def compact_cell_range(uniq, off, key):
    uniq, off = _u32(uniq), _u32(off)
    b = ctypes.c_uint32()
    e = ctypes.c_uint32()
    ok = lib().oracle_compact_cell_range(_ptr(uniq), _ptr(off), len(uniq), key, ctypes.addressof(b),
                                         ctypes.addressof(e))
    return bool(ok), b.value, e.value


# ---------------------------------------------------------- neighbour solves
def lambda_dense(P, idxS, keyS, start, end, dp):
    P, idxS, keyS, start, end = _f4(P), _u32(idxS), _u32(keyS), _u32(start), _u32(end)
    lam = np.zeros(len(P), np.float32)
    lib().oracle_lambda_dense(_ptr(lam), _ptr(P), _ptr(idxS), _ptr(keyS), _ptr(start), _ptr(end), _ptr(dp), len(P))
    return lam


def lambda_compact(P, idxS, keyS, uniq, off, dp):
    P, idxS, keyS, uniq, off = _f4(P), _u32(idxS), _u32(keyS), _u32(uniq), _u32(off)
    lam = np.zeros(len(P), np.float32)
    cnt = np.array([len(uniq)], np.uint32)
    lib().oracle_lambda_compact(_ptr(lam), _ptr(P), _ptr(idxS), _ptr(keyS), _ptr(uniq), _ptr(off), _ptr(cnt),
                                _ptr(dp), len(P))
    return lam


def delta_apply_dense(P, lam, idxS, keyS, start, end, dp):
    P = _f4(P).copy()
    lam, idxS, keyS, start, end = (np.ascontiguousarray(lam, np.float32), _u32(idxS), _u32(keyS), _u32(start),
                                   _u32(end))
    delta = np.zeros_like(P)
    lib().oracle_delta_apply_dense(_ptr(P), _ptr(delta), _ptr(lam), _ptr(idxS), _ptr(keyS), _ptr(start), _ptr(end),
                                   _ptr(dp), len(P))
    return P, delta


def delta_apply_compact(P, lam, idxS, keyS, uniq, off, dp):
    P = _f4(P).copy()
    lam, idxS, keyS, uniq, off = (np.ascontiguousarray(lam, np.float32), _u32(idxS), _u32(keyS), _u32(uniq),
                                  _u32(off))
    delta = np.zeros_like(P)
    cnt = np.array([len(uniq)], np.uint32)
    lib().oracle_delta_apply_compact(_ptr(P), _ptr(delta), _ptr(lam), _ptr(idxS), _ptr(keyS), _ptr(uniq), _ptr(off),
                                     _ptr(cnt), _ptr(dp), len(P))
    return P, delta


def xsph_dense(V, P, idxS, keyS, start, end, dp):
    V, P = _f4(V), _f4(P)
    out = V.copy()
    lib().oracle_xsph_dense(_ptr(out), _ptr(V), _ptr(P), _ptr(_u32(idxS)), _ptr(_u32(keyS)), _ptr(_u32(start)),
                            _ptr(_u32(end)), _ptr(dp), len(P))
    return out


def xsph_compact(V, P, idxS, keyS, uniq, off, dp):
    V, P = _f4(V), _f4(P)
    out = V.copy()
    uniq = _u32(uniq)
    cnt = np.array([len(uniq)], np.uint32)
    lib().oracle_xsph_compact(_ptr(out), _ptr(V), _ptr(P), _ptr(_u32(idxS)), _ptr(_u32(keyS)), _ptr(uniq),
                              _ptr(_u32(off)), _ptr(cnt), _ptr(dp), len(P))
    return out


# ----------------------------------------------------------------- full step
def run(X, V, sp, steps, compact=False):
    """Advance (X, V) by `steps` steps; returns new copies."""
    X, V = _f4(X).copy(), _f4(V).copy()
    assert sp.numParticles == len(X)
    rc = lib().oracle_run(_ptr(X), _ptr(V), _ptr(sp), int(bool(compact)), int(steps))
    if rc != 0:
        raise RuntimeError("oracle_run failed")
    return X, V


def density_stats(X, V, sp):
    X, V = _f4(X), _f4(V)
    out = np.zeros(5, np.float64)
    lib().oracle_density_stats(_ptr(X), _ptr(V), _ptr(sp), _ptr(out))
    return {"mean_err": out[0], "max_err": out[1], "mean_neighbors": out[2], "mean_speed": out[3],
            "mean_pos_err": out[4]}


def seed_lattice(nx, ny, nz, origin, spacing, grid, jitter_amp, seed):
    n = nx * ny * nz
    X = np.zeros((n, 4), np.float32)
    lib().oracle_seed_lattice(_ptr(X), nx, ny, nz, origin[0], origin[1], origin[2], spacing, _ptr(grid),
                              jitter_amp, seed)
    return X


def jitter_uniform(seed, particle, axis):
    return lib().oracle_jitter_uniform(seed, particle, axis)


def pow_scorr(x, n):
    return lib().oracle_pow_scorr(float(x), float(n))


def num_threads():
    return lib().oracle_num_threads()


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


# ------------------------------------------------------------ scene generators
# Restated in numpy / Python integers (float32 where the reference computes in float).

def _splitmix_uniform(seed, index):
    """The seeders' variate, vectorised: SplitMix64 finaliser over seed + golden * (index + 1)
    (index = 3 * particle + axis), top 24 bits as a float in [0, 1)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (np.asarray(index, np.uint64) + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(40)).astype(np.uint32).astype(np.float32) * np.float32(1.0 / 16777216.0)


def lattice_auto_dims(total, box_min, box_max, spacing, origin):
    """reference: simulator.cpp:882-941 (seedBoxLatticeAuto): returns ((nx, ny, nz), origin)."""
    f = np.float32
    sp = f(spacing)
    size = [max(f(0), f(box_max[a]) - f(box_min[a])) for a in range(3)]
    mx = [max(1, int(np.floor(f(s) / sp))) for s in size]
    T = max(1, int(total))
    nx = min(mx[0], int(np.ceil(np.power(float(T), 1.0 / 3.0))))
    rem = (T + nx - 1) // nx
    ny = min(mx[1], int(np.ceil(np.sqrt(float(rem)))))
    nz = min(mx[2], int(np.ceil(float(rem) / max(1, ny))))
    while nx * ny * nz < T:
        if nx < mx[0]:
            nx += 1
        elif ny < mx[1]:
            ny += 1
        elif nz < mx[2]:
            nz += 1
        else:
            break
    m = sp * f(0.25)
    o = tuple(float(min(max(f(origin[a]), f(box_min[a]) + m), f(box_max[a]) - m)) for a in range(3))
    return (nx, ny, nz), o


def cube_mix_centers(group_count, layers, group_spacing, layer_spacing, base_height, layer_jitter=False,
                     layer_jitter_amp=0.0, layer_jitter_seed=0x13579BDF):
    """reference: console.cpp:277-316 (GenerateCubeMixCenters).  Per-layer counts use
    uint32 arithmetic (remaining wraps when layers outnumber groups); the layer jitter
    is one seeder variate per layer, from seed ^ layer."""
    f = np.float32
    counts, remaining = [], group_count & 0xFFFFFFFF
    for L in range(layers):
        take = remaining // (layers - L) or 1
        counts.append(take)
        remaining = (remaining - take) & 0xFFFFFFFF
    out = []
    for layer in range(layers):
        if len(out) >= group_count:
            break
        n = counts[layer]
        nx = int(np.ceil(np.sqrt(float(n))))
        y = f(base_height) + f(layer) * f(layer_spacing)
        if layer_jitter:
            u = _splitmix_uniform((layer_jitter_seed ^ layer) & 0xFFFFFFFFFFFFFFFF, 1)
            y = f(y + (f(2) * u - f(1)) * f(layer_jitter_amp))
        for k in range(n):
            if len(out) >= group_count:
                break
            out.append((f(k % nx) * f(group_spacing), y, f(k // nx) * f(group_spacing)))
    return np.array(out, np.float32).reshape(-1, 3)


def seed_cube_mix(box_min, box_max, centers, edge, spacing, jitter_amp, seed):
    """reference: simulator.cpp:920-975 (seedCubeMix): (groups * edge^3, 4) float32, x fastest,
    w = group index, clamped; per-axis jitter of +-jitter_amp, clamped again."""
    f = np.float32
    lo = np.array(box_min, f)
    hi = np.array(box_max, f)
    sp = f(spacing)
    start = f(-0.5) * sp * f(edge - 1)
    iz, iy, ix = np.meshgrid(np.arange(edge), np.arange(edge), np.arange(edge), indexing="ij")
    offs = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1).astype(f) * sp
    blocks = []
    for g, c in enumerate(np.asarray(centers, f).reshape(-1, 3)):
        p = (c + start) + offs
        p = np.minimum(np.maximum(p, lo), hi)
        blocks.append(np.concatenate([p, np.full((len(p), 1), f(g))], 1))
    X = np.concatenate(blocks).astype(f) if blocks else np.zeros((0, 4), f)
    if jitter_amp > 0:
        n = len(X)
        idx = np.arange(n, dtype=np.uint64)[:, None] * np.uint64(3) + np.arange(3, dtype=np.uint64)[None, :]
        j = (f(2) * _splitmix_uniform(seed, idx) - f(1)) * f(jitter_amp)
        X[:, :3] = np.minimum(np.maximum(X[:, :3] + j, lo), hi)
    return X
