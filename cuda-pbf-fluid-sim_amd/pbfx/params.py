"""ctypes mirrors of the solver's parameter block and the scene configurations.

The structures lay out exactly as ``sim::KernelCoeffs``, ``GridBounds``,
``PbfTuning``, ``SimParams`` and ``DeviceParams`` in include/pbf_params.h
(reference: sim/parameters.h:9-92): 24, 40, 88, 212 and 192 bytes.

``build_sim_params`` restates how the reference application turns its runtime
console into SimParams (reference: engine/core/console.cpp:198-267): h from the
particle radius, mass from the uniform-lattice rule ``rho0 * h^3``, cell size h,
grid dims ``max(1, ceil(extent / cell))``.  All arithmetic is float32, one
rounding per operation, in the reference's order.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

f32 = np.float32


class Float3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class Int3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class KernelCoeffs(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("h", "inv_h", "h2", "poly6", "spiky", "visc")]


class GridBounds(ctypes.Structure):
    _fields_ = [("mins", Float3), ("maxs", Float3), ("cellSize", ctypes.c_float), ("dim", Int3)]

    @property
    def num_cells(self):
        return int(self.dim.x) * int(self.dim.y) * int(self.dim.z)


_PBF_FIELDS = [
    ("scorr_enable", ctypes.c_int, 1), ("scorr_k", ctypes.c_float, 0.003), ("scorr_n", ctypes.c_float, 4.0),
    ("scorr_dq_h", ctypes.c_float, 0.3), ("wq_min", ctypes.c_float, 1e-12), ("scorr_min", ctypes.c_float, -0.25),
    ("grad_r_eps", ctypes.c_float, 1e-6), ("lambda_denom_eps", ctypes.c_float, 1e-4),
    ("compliance", ctypes.c_float, 0.0), ("xpbd_enable", ctypes.c_int, 0),
    ("enable_lambda_clamp", ctypes.c_int, 1), ("lambda_max_abs", ctypes.c_float, 50.0),
    ("enable_disp_clamp", ctypes.c_int, 1), ("disp_clamp_max_h", ctypes.c_float, 0.05),
    ("enable_relax", ctypes.c_int, 1), ("relax_omega", ctypes.c_float, 0.75),
    ("xsph_gate_enable", ctypes.c_int, 0), ("xsph_n_min", ctypes.c_int, 0), ("xsph_n_max", ctypes.c_int, 8),
    ("lambda_warm_start_enable", ctypes.c_int, 0), ("lambda_warm_start_decay", ctypes.c_float, 0.5),
    ("semi_implicit_integration_enable", ctypes.c_int, 0),
]


class PbfTuning(ctypes.Structure):
    _fields_ = [(n, t) for n, t, _ in _PBF_FIELDS]

    def __init__(self, **kw):
        super().__init__()
        for n, _, d in _PBF_FIELDS:
            setattr(self, n, kw.pop(n, d))
        if kw:
            raise TypeError(f"unknown PbfTuning fields {sorted(kw)}")


class SimParams(ctypes.Structure):
    _fields_ = [
        ("numParticles", ctypes.c_uint32), ("maxParticles", ctypes.c_uint32), ("gravity", Float3),
        ("dt", ctypes.c_float), ("cfl", ctypes.c_float), ("restDensity", ctypes.c_float),
        ("solverIters", ctypes.c_int), ("maxNeighbors", ctypes.c_int), ("sortEveryN", ctypes.c_int),
        ("boundaryRestitution", ctypes.c_float), ("particleMass", ctypes.c_float), ("pbf", PbfTuning),
        ("xsph_c", ctypes.c_float), ("kernel", KernelCoeffs), ("grid", GridBounds),
        ("ghostParticleCount", ctypes.c_uint32),
    ]


class DeviceParams(ctypes.Structure):
    _fields_ = [
        ("kernel", KernelCoeffs), ("grid", GridBounds), ("gravity", Float3), ("restDensity", ctypes.c_float),
        ("dt", ctypes.c_float), ("inv_dt", ctypes.c_float), ("boundaryRestitution", ctypes.c_float),
        ("maxNeighbors", ctypes.c_int), ("particleMass", ctypes.c_float), ("pbf", PbfTuning),
        ("xsph_c", ctypes.c_float),
    ]


assert ctypes.sizeof(KernelCoeffs) == 24
assert ctypes.sizeof(GridBounds) == 40
assert ctypes.sizeof(PbfTuning) == 88
assert ctypes.sizeof(DeviceParams) == 192
assert DeviceParams.pbf.offset == 100 and DeviceParams.xsph_c.offset == 188
assert ctypes.sizeof(SimParams) == 212


def make_kernel_coeffs(h):
    """float32 restatement of MakeKernelCoeffs (reference: sim/parameters.h:94-107)."""
    h = f32(h)
    pi = f32(3.14159265358979323846)
    kc = KernelCoeffs()
    h2 = f32(h * h)
    h3 = f32(h * h2)
    h6 = f32(h3 * h3)
    h9 = f32(h6 * h3)
    kc.h = h
    kc.inv_h = f32(f32(1.0) / h)
    kc.h2 = h2
    kc.poly6 = f32(f32(315.0) / f32(f32(f32(64.0) * pi) * h9))
    kc.spiky = f32(f32(15.0) / f32(pi * h6))
    kc.visc = f32(f32(15.0) / f32(f32(f32(2.0) * pi) * h3))
    return kc


def make_device_params(sp):
    """SimParams -> DeviceParams (reference: sim/parameters.h:109-123)."""
    dp = DeviceParams()
    dp.kernel = sp.kernel
    dp.grid = sp.grid
    dp.gravity = sp.gravity
    dp.restDensity = sp.restDensity
    dp.dt = sp.dt
    dp.inv_dt = float(f32(1.0) / f32(sp.dt)) if sp.dt > 0 else 0.0
    dp.boundaryRestitution = sp.boundaryRestitution
    dp.maxNeighbors = sp.maxNeighbors
    dp.particleMass = sp.particleMass
    dp.pbf = sp.pbf
    dp.xsph_c = sp.xsph_c
    return dp


def compute_dims(mins, maxs, cell, clamp_ge1=True):
    """Grid dims ceil((max - min) / cell) per axis in float32 (reference: console.cpp:258-265)."""
    out = []
    for lo, hi in zip(mins, maxs):
        size = f32(f32(hi) - f32(lo))
        d = int(math.ceil(float(f32(size / f32(cell)))))
        out.append(max(1, d) if clamp_ge1 else max(0, d))
    return tuple(out)


@dataclass
class Scene:
    """A particle scene: lattice block(s) inside a box, plus the physics constants."""
    name: str
    lattices: list            # [(nx, ny, nz, (ox, oy, oz))], x fastest
    box_min: tuple
    box_max: tuple
    radius: float
    h_over_r: float = 2.5
    rest_density: float = 1000.0
    gravity: tuple = (0.0, -9.81, 0.0)
    dt: float = 1.0 / 120.0
    solver_iters: int = 2
    xsph_c: float = 0.01
    max_neighbors: int = 64
    restitution: float = 0.0
    compliance: float = 2e-5
    jitter_frac_r: float = 0.1
    seed: int = 0x5EED
    grid: str = "dense"          # "dense" | "compact" | "hashed"
    table_log2: int = 22
    steps: int = 200
    pbf: dict = field(default_factory=dict)

    @property
    def num_particles(self):
        return sum(nx * ny * nz for nx, ny, nz, _ in self.lattices)

    @property
    def spacing(self):
        return float(f32(f32(2.0) * f32(self.radius)))


def build_sim_params(scene):
    """SimParams as the reference app derives them (reference: console.cpp:198-267), overriding only
    the scene's stated constants.  XPBD compliance and the unused warm-start / semi-implicit flags take
    the app's defaults (reference: console.h:215-220)."""
    sp = SimParams()
    n = scene.num_particles
    sp.numParticles = n
    sp.maxParticles = n
    sp.dt = float(f32(scene.dt))
    sp.cfl = 0.45
    sp.gravity = Float3(*scene.gravity)
    sp.restDensity = scene.rest_density
    sp.solverIters = scene.solver_iters
    sp.maxNeighbors = scene.max_neighbors
    sp.sortEveryN = 1
    sp.boundaryRestitution = scene.restitution
    pbf = dict(xpbd_enable=1, compliance=scene.compliance, lambda_warm_start_enable=1,
               lambda_warm_start_decay=0.5, semi_implicit_integration_enable=1)
    pbf.update(scene.pbf)
    sp.pbf = PbfTuning(**pbf)
    sp.xsph_c = scene.xsph_c
    r = f32(max(f32(1e-8), f32(scene.radius)))
    h = f32(f32(scene.h_over_r) * r)
    sp.kernel = make_kernel_coeffs(h)
    spacing = f32(f32(1.0) * h)
    sp.particleMass = float(f32(f32(f32(f32(scene.rest_density) * spacing) * spacing) * spacing))
    cell = f32(max(f32(1e-6), f32(f32(1.0) * h)))
    g = GridBounds()
    g.mins = Float3(*scene.box_min)
    g.maxs = Float3(*scene.box_max)
    g.cellSize = float(cell)
    g.dim = Int3(*compute_dims(scene.box_min, scene.box_max, cell))
    sp.grid = g
    sp.ghostParticleCount = 0
    return sp


def seed_scene_reference(scene, seeder):
    """Seed every lattice of a scene with `seeder(nx, ny, nz, origin, spacing, grid, amp, seed)`."""
    sp = build_sim_params(scene)
    blocks = []
    amp = float(f32(f32(scene.jitter_frac_r) * f32(scene.radius)))
    for bi, (nx, ny, nz, origin) in enumerate(scene.lattices):
        blocks.append(seeder(nx, ny, nz, origin, scene.spacing, sp.grid, amp, scene.seed + bi))
    return np.concatenate(blocks, axis=0) if blocks else np.zeros((0, 4), np.float32)


# ------------------------------------------------------------------ configs
def dam_break_32k(**kw):
    """32,768 particles, 32^3 lattice at spacing 2r (r = 0.005 m), corner of a 1 m cube, dense grid."""
    r = 0.005
    o = (r, r, r)
    args = dict(name="dam_break_32k", lattices=[(32, 32, 32, o)], box_min=(0.0, 0.0, 0.0),
                box_max=(1.0, 1.0, 1.0), radius=r, grid="dense", steps=200)
    args.update(kw)
    return Scene(**args)


def dam_break_1m(**kw):
    """1,000,000 particles, 100^3 lattice in a 2x1x1 m box, hashed grid (2^22 table), 500 steps."""
    r = 0.005
    o = (r, r, r)
    args = dict(name="dam_break_1m", lattices=[(100, 100, 100, o)], box_min=(0.0, 0.0, 0.0),
                box_max=(2.0, 1.0, 1.0), radius=r, grid="hashed", table_log2=22, steps=500)
    args.update(kw)
    return Scene(**args)


def splash_8m(**kw):
    """8,000,000 particles in two 100x100x400 columns (laid along z so they fit the 2 m tall box),
    r = 0.004 m, 4x2x4 m box, 4 iterations, dt = 1/240, hashed grid (2^24 table), 300 steps."""
    r = 0.004
    sp = 2 * r
    cols = [(100, 100, 400, (0.5 + r, 1.0 + r, 0.4 + r)), (100, 100, 400, (2.7 + r, 1.0 + r, 0.4 + r))]
    del sp
    args = dict(name="splash_8m", lattices=cols, box_min=(0.0, 0.0, 0.0), box_max=(4.0, 2.0, 4.0), radius=r,
                solver_iters=4, dt=1.0 / 240.0, grid="hashed", table_log2=24, steps=300)
    args.update(kw)
    return Scene(**args)


def small_block(n=(8, 8, 8), **kw):
    """A small dam-break block for parity tests (same constants as the 32k case)."""
    r = 0.005
    o = (r, r, r)
    args = dict(name="small_block", lattices=[(n[0], n[1], n[2], o)], box_min=(0.0, 0.0, 0.0),
                box_max=(1.0, 1.0, 1.0), radius=r, grid="dense", steps=10)
    args.update(kw)
    return Scene(**args)
