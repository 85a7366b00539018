"""Headless simulator handle (include/pbf_sim.h) with numpy host I/O."""
import ctypes

import numpy as np

from . import _lib
from .params import Scene, build_sim_params, f32

GRID_KINDS = {"dense": 0, "compact": 1, "hashed": 2}
PATHS = {"reordered": 0, "reference": 1}


class _Tuning(ctypes.Structure):
    _fields_ = [("candidateMask", ctypes.c_int), ("streamRatio", ctypes.c_uint32), ("chunkBits", ctypes.c_int),
                ("sortItems", ctypes.c_int), ("sortDigitBits", ctypes.c_int), ("walkBlock", ctypes.c_int),
                ("laneOrder", ctypes.c_int)]


class _CudaArray:
    """__cuda_array_interface__ holder for a device pointer owned by a simulator."""

    def __init__(self, ptr, shape, typestr, owner):
        self.owner = owner
        self.__cuda_array_interface__ = {"shape": shape, "typestr": typestr, "data": (int(ptr or 0), False),
                                         "version": 2, "strides": None}


def _device_view(ptr, n, owner, shape=None, typestr="<f4"):
    import torch
    if not ptr:
        raise _lib.PbfError("no device buffer")
    return torch.as_tensor(_CudaArray(ptr, shape or (n, 4), typestr, owner), device="cuda")


class Simulator:
    """sim::Simulator over the C ABI: initialize, seed, step, read back."""

    def __init__(self, params, grid="dense", table_log2=22, path="reordered", use_graph=True):
        self.params = params
        self.n = int(params.numParticles)
        L = _lib.lib()
        h = L.pbfSimCreate(ctypes.byref(params), GRID_KINDS[grid], int(table_log2), PATHS[path], int(use_graph))
        if not h:
            _lib.check("pbfSimCreate")
            raise _lib.PbfError("pbfSimCreate failed")
        self._h = ctypes.c_void_p(h)
        self.grid = grid
        self.path = path

    @classmethod
    def from_scene(cls, scene: Scene, path="reordered", use_graph=True, grid=None):
        sp = build_sim_params(scene)
        sim = cls(sp, grid=grid or scene.grid, table_log2=scene.table_log2, path=path, use_graph=use_graph)
        sim.seed_scene(scene)
        return sim

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().pbfSimDestroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def seed_scene(self, scene):
        amp = float(f32(f32(scene.jitter_frac_r) * f32(scene.radius)))
        first = 0
        for bi, (nx, ny, nz, o) in enumerate(scene.lattices):
            rc = _lib.lib().pbfSimSeedLattice(self._h, first, nx, ny, nz, o[0], o[1], o[2], scene.spacing, amp,
                                              scene.seed + bi)
            if rc != 0:
                _lib.check("pbfSimSeedLattice")
                raise _lib.PbfError("seed failed")
            first += nx * ny * nz

    def seed_cube_mix(self, centers, edge, spacing, jitter_amp=0.0, seed=0, first=0):
        """groups * edge^3 particles in cubes around `centers` ((G, 3)); see pbfSimSeedCubeMix."""
        c = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        if _lib.lib().pbfSimSeedCubeMix(self._h, int(first), len(c), c.ctypes.data, int(edge), float(spacing),
                                        float(jitter_amp), int(seed)) != 0:
            _lib.check("pbfSimSeedCubeMix")
            raise _lib.PbfError("seed_cube_mix failed")

    def bind_external_positions(self, a, b):
        """Adopt two caller-owned (>= N, 4) float32 CUDA tensors as the position ping-pong
        (reference: bindExternalPosPingPong).  The tensors must outlive the simulator."""
        import torch
        for t in (a, b):
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.dim() == 2
                    and t.shape[1] == 4 and t.shape[0] >= self.n):
                raise ValueError("external positions: contiguous (>= N, 4) float32 CUDA tensors")
        torch.cuda.synchronize()
        if _lib.lib().pbfSimBindExternalPositions(self._h, a.data_ptr(), b.data_ptr(), min(a.shape[0],
                                                                                            b.shape[0])) != 0:
            _lib.check("pbfSimBindExternalPositions")
            raise _lib.PbfError("bind_external_positions failed")
        self._external = (a, b)

    def bind_external_velocity(self, buf, stride_bytes=16):
        """Caller-owned CUDA tensor that publish_velocity() fills, one float4 every stride_bytes."""
        if _lib.lib().pbfSimBindExternalVelocity(self._h, buf.data_ptr(), buf.numel() * buf.element_size(),
                                                 int(stride_bytes)) != 0:
            _lib.check("pbfSimBindExternalVelocity")
            raise _lib.PbfError("bind_external_velocity failed")
        self._external_vel = buf

    def publish_velocity(self, count=None):
        if _lib.lib().pbfSimPublishVelocity(self._h, self.n if count is None else int(count)) != 0:
            _lib.check("pbfSimPublishVelocity")
            raise _lib.PbfError("publish_velocity failed")

    def set_tuning(self, candidate_mask=None, stream_ratio=None, chunk_bits=None, sort_items=None,
                   sort_digit_bits=None, walk_block=None, lane_order=None):
        """Change launch knobs (results are bit-identical for every setting)."""
        t = self.tuning()
        if candidate_mask is not None:
            t["candidate_mask"] = bool(candidate_mask)
        if stream_ratio is not None:
            t["stream_ratio"] = int(stream_ratio)
        if chunk_bits is not None:
            t["chunk_bits"] = bool(chunk_bits)
        if sort_items is not None:
            t["sort_items"] = int(sort_items)
        if sort_digit_bits is not None:
            t["sort_digit_bits"] = int(sort_digit_bits)
        if walk_block is not None:
            t["walk_block"] = int(walk_block)
        if lane_order is not None:
            t["lane_order"] = bool(lane_order)
        c = _Tuning(int(t["candidate_mask"]), t["stream_ratio"], int(t["chunk_bits"]), t["sort_items"],
                    t["sort_digit_bits"], t["walk_block"], int(t["lane_order"]))
        if _lib.lib().pbfSimSetTuning(self._h, ctypes.byref(c)) != 0:
            _lib.check("pbfSimSetTuning")
            raise _lib.PbfError("set_tuning failed")

    def tuning(self):
        c = _Tuning()
        if _lib.lib().pbfSimGetTuning(self._h, ctypes.byref(c)) != 0:
            raise _lib.PbfError("get_tuning failed")
        return {"candidate_mask": bool(c.candidateMask), "stream_ratio": int(c.streamRatio),
                "chunk_bits": bool(c.chunkBits), "sort_items": int(c.sortItems),
                "sort_digit_bits": int(c.sortDigitBits), "walk_block": int(c.walkBlock),
                "lane_order": bool(c.laneOrder)}

    def positions_view(self):
        """(N, 4) float32 CUDA tensor viewing the current position buffer (no copy).  The
        buffer alternates between steps; take a new view after stepping.  Steps run on the
        simulator's own stream: synchronize() before reading the view on another stream,
        and finish writes to it there before the next step."""
        return _device_view(_lib.lib().pbfSimDevicePositions(self._h), self.n, self)

    def velocities_view(self):
        """(N, 4) float32 CUDA tensor viewing the velocity buffer (no copy; the same stream
        rule as positions_view)."""
        return _device_view(_lib.lib().pbfSimDeviceVelocities(self._h), self.n, self)

    def sorted_grid(self):
        """Host copies (uint32) of the last grid build's sorted cell keys and particle
        indices (reference: d_cellKeys_sorted / d_indices_sorted, phases.cpp:44-63)."""
        self.synchronize()
        out = []
        for fn in (_lib.lib().pbfSimDeviceSortedKeys, _lib.lib().pbfSimDeviceSortedIndices):
            t = _device_view(fn(self._h), self.n, self, shape=(self.n,), typestr="<i4")
            out.append(t.cpu().numpy().view(np.uint32).copy())
        return tuple(out)

    def set_state(self, pos, vel):
        pos = np.ascontiguousarray(pos, np.float32)
        vel = np.ascontiguousarray(vel, np.float32)
        assert pos.shape == (self.n, 4) and vel.shape == (self.n, 4)
        if _lib.lib().pbfSimSetState(self._h, pos.ctypes.data, vel.ctypes.data, self.n, 0) != 0:
            _lib.check("pbfSimSetState")
            raise _lib.PbfError("set_state failed")

    def state(self):
        pos = np.zeros((self.n, 4), np.float32)
        vel = np.zeros((self.n, 4), np.float32)
        if _lib.lib().pbfSimGetState(self._h, pos.ctypes.data, vel.ctypes.data, 0) != 0:
            _lib.check("pbfSimGetState")
            raise _lib.PbfError("get_state failed")
        return pos, vel

    def state_device(self):
        """Device copies of the current state: two (N, 4) float32 CUDA tensors."""
        import torch
        pos = torch.empty((self.n, 4), dtype=torch.float32, device="cuda")
        vel = torch.empty_like(pos)
        torch.cuda.synchronize()
        if _lib.lib().pbfSimGetState(self._h, ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(vel.data_ptr()),
                                     1) != 0:
            _lib.check("pbfSimGetState")
            raise _lib.PbfError("get_state failed")
        return pos, vel

    def set_state_device(self, pos, vel):
        """Replace the state from device tensors ((N, 4) float32, contiguous, CUDA)."""
        import torch
        for t in (pos, vel):
            assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and tuple(t.shape) == (self.n, 4)
        torch.cuda.synchronize()
        if _lib.lib().pbfSimSetState(self._h, ctypes.c_void_p(pos.data_ptr()), ctypes.c_void_p(vel.data_ptr()),
                                     self.n, 1) != 0:
            _lib.check("pbfSimSetState")
            raise _lib.PbfError("set_state failed")

    def step(self, steps=1):
        if _lib.lib().pbfSimStep(self._h, int(steps)) != 0:
            _lib.check("pbfSimStep")
            raise _lib.PbfError("step failed")

    def set_params(self, params):
        """New parameters for the following steps (reference: Simulator::step's parameter
        handling, simulator.cpp:575-607, ParamChangeTracker param_change_tracker.h:23-61).
        Coefficients, gravity, dt and the like take effect at the next step (the graph is
        re-captured); a new particle count, grid, h or iteration count re-allocates what
        depends on it and keeps the particles' state.  Raises PbfError when refused (see
        pbfSimSetParams)."""
        if _lib.lib().pbfSimSetParams(self._h, ctypes.byref(params)) != 0:
            _lib.check("pbfSimSetParams")
            raise _lib.PbfError("set_params refused")
        self.params = params
        self.n = int(params.numParticles)

    def synchronize(self):
        if _lib.lib().pbfSimSynchronize(self._h) != 0:
            raise _lib.PbfError("synchronize failed")
        _lib.check("pbfSimSynchronize")

    @property
    def stream(self):
        return _lib.lib().pbfSimStream(self._h)

    def profile_launches(self):
        """One eager step with an event after every launch: [(label, ms), ...] in launch order."""
        ms = (ctypes.c_float * 64)()
        names = ctypes.create_string_buffer(4096)
        n = _lib.lib().pbfSimProfilePhases(self._h, ctypes.addressof(ms), 64, ctypes.addressof(names), 4096)
        if n < 0:
            _lib.check("pbfSimProfilePhases")
            raise _lib.PbfError("profile failed")
        labels = names.raw.split(b"\0")[:n]
        return [(l.decode(), float(ms[i])) for i, l in enumerate(labels)]

    def density_stats(self):
        """Density-constraint statistics of the current state, computed on the GPU."""
        out = np.zeros(5, np.float64)
        if _lib.lib().pbfSimDensityStats(self._h, out.ctypes.data) != 0:
            _lib.check("pbfSimDensityStats")
            raise _lib.PbfError("density_stats failed")
        return {"mean_err": out[0], "max_err": out[1], "mean_neighbors": out[2], "mean_speed": out[3],
                "mean_pos_err": out[4]}

    def hash_overflow(self):
        return bool(_lib.lib().pbfSimHashOverflow(self._h))
