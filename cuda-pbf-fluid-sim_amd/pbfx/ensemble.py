"""Batched ensembles: many independent dam-break scenes advanced by one launch per phase.

``ensemble_scenes`` draws the 256-scene configuration from a seed:
- particle counts log-uniform in [4096, 262144];
- square-based columns with height/width aspect uniform in [0.5, 2];
- gravity_y uniform in [-12, -8];
- XSPH c uniform in [0.005, 0.05].
Every scene keeps r = 0.005 m and the 32k dam break's constants. Its box is
max(1 m, 1.25 x column extent) per axis, so the largest columns fit. Columns stand
in the box corner at (r, r, r), like the single-scene dam breaks. The draws use
numpy's PCG64 generator.
"""
import ctypes
import math

import numpy as np

from . import _lib
from .params import Scene, build_sim_params, f32
from .sim import GRID_KINDS, Simulator


class SceneDesc(ctypes.Structure):
    _fields_ = [("boxMin", ctypes.c_float * 3), ("boxMax", ctypes.c_float * 3), ("gravity", ctypes.c_float * 3),
                ("xsph_c", ctypes.c_float), ("restitution", ctypes.c_float), ("numParticles", ctypes.c_uint32)]


assert ctypes.sizeof(SceneDesc) == 48


def ensemble_scenes(count=256, seed=2024, n_min=4096, n_max=262144, radius=0.005):
    """The batched configuration: a list of single-scene `Scene`s drawn from `seed`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    scenes = []
    sp2 = 2 * radius
    for k in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        aspect = float(rng.uniform(0.5, 2.0))
        gy = float(rng.uniform(-12.0, -8.0))
        c = float(rng.uniform(0.005, 0.05))
        nx = max(1, int(round((n / aspect) ** (1.0 / 3.0))))
        ny = max(1, int(round(n / (nx * nx))))
        ext = (nx * sp2, ny * sp2, nx * sp2)
        box = tuple(max(1.0, 1.25 * e) for e in ext)
        scenes.append(Scene(name=f"ensemble_{k}", lattices=[(nx, ny, nx, (radius, radius, radius))],
                            box_min=(0.0, 0.0, 0.0), box_max=box, radius=radius, gravity=(0.0, gy, 0.0),
                            xsph_c=c, grid="hashed", seed=seed * 1000 + k, steps=200))
    return scenes


class Ensemble:
    """Handle over pbfEnsemble* (a pbfSim with a scene table)."""

    def __init__(self, scenes, grid="hashed", table_log2=None, use_graph=True):
        assert scenes
        self.scenes = scenes
        base = build_sim_params(scenes[0])
        self.base = base
        descs = (SceneDesc * len(scenes))()
        total = 0
        for d, sc in zip(descs, scenes):
            d.boxMin[:] = sc.box_min
            d.boxMax[:] = sc.box_max
            d.gravity[:] = sc.gravity
            d.xsph_c = sc.xsph_c
            d.restitution = sc.restitution
            d.numParticles = sc.num_particles
            total += sc.num_particles
        self.n = total
        self.grid = grid
        self.params = base
        if table_log2 is None:
            table_log2 = max(16, int(math.ceil(math.log2(max(total, 1)))) + 1)
        L = _lib.lib()
        h = L.pbfEnsembleCreate(ctypes.byref(base), ctypes.addressof(descs), len(scenes), GRID_KINDS[grid],
                                int(table_log2), int(use_graph))
        if not h:
            _lib.check("pbfEnsembleCreate")
            raise _lib.PbfError("pbfEnsembleCreate failed")
        self._h = ctypes.c_void_p(h)
        for k, sc in enumerate(scenes):
            amp = float(f32(f32(sc.jitter_frac_r) * f32(sc.radius)))
            (nx, ny, nz, o), = sc.lattices
            if L.pbfEnsembleSeedLattice(self._h, k, nx, ny, nz, o[0], o[1], o[2], sc.spacing, amp, sc.seed) != 0:
                _lib.check("pbfEnsembleSeedLattice")
                raise _lib.PbfError("seed failed")

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().pbfSimDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps=1):
        if _lib.lib().pbfSimStep(self._h, int(steps)) != 0:
            _lib.check("pbfSimStep")
            raise _lib.PbfError("step failed")

    def current_params(self):
        """A copy of the parameters the ensemble steps with: its shared parameters, with
        the ensemble's particle count and key space (the base for set_params)."""
        out = type(self.params)()
        if _lib.lib().pbfSimGetParams(self._h, ctypes.byref(out)) != 0:
            raise _lib.PbfError("get_params failed")
        return out

    def set_params(self, params):
        """New shared parameters (dt, solver coefficients, clamps, neighbour cap) for the
        following steps, given as a changed copy of current_params().  Box, gravity, XSPH
        and restitution are per scene; changing them raises PbfError."""
        if _lib.lib().pbfSimSetParams(self._h, ctypes.byref(params)) != 0:
            _lib.check("pbfSimSetParams")
            raise _lib.PbfError("set_params refused")

    def synchronize(self):
        if _lib.lib().pbfSimSynchronize(self._h) != 0:
            raise _lib.PbfError("synchronize failed")

    def scene_state(self, k):
        n = self.scenes[k].num_particles
        pos = np.zeros((n, 4), np.float32)
        vel = np.zeros((n, 4), np.float32)
        if _lib.lib().pbfEnsembleGetScene(self._h, k, pos.ctypes.data, vel.ctypes.data) != 0:
            _lib.check("pbfEnsembleGetScene")
            raise _lib.PbfError("get scene failed")
        return pos, vel

    def scene_info(self, k):
        first = ctypes.c_uint32()
        cnt = ctypes.c_uint32()
        dims = (ctypes.c_int * 3)()
        if _lib.lib().pbfEnsembleSceneInfo(self._h, k, ctypes.byref(first), ctypes.byref(cnt), dims) != 0:
            raise _lib.PbfError("scene info failed")
        return first.value, cnt.value, tuple(dims)

    def profile_launches(self):
        return Simulator.profile_launches(self)

    # whole-ensemble state (all scenes, scene order) and launch knobs, as for one scene
    state = Simulator.state
    set_state = Simulator.set_state
    tuning = Simulator.tuning
    set_tuning = Simulator.set_tuning
    hash_overflow = Simulator.hash_overflow
    state_device = Simulator.state_device
    set_state_device = Simulator.set_state_device


def solo(scene, grid="hashed", use_graph=True):
    """The same scene as a single-scene simulator (the scene-by-scene path)."""
    sp = build_sim_params(scene)
    tl = max(16, int(math.ceil(math.log2(max(scene.num_particles, 1)))) + 1)
    sim = Simulator(sp, grid=grid, table_log2=tl, use_graph=use_graph)
    sim.seed_scene(scene)
    return sim
