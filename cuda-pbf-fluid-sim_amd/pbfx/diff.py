"""Differentiable PBF rollouts: the hand-written adjoint kernels behind torch.autograd.

``Rollout`` runs T steps on the GPU and keeps (X, V) checkpoints every k steps.
The reverse pass walks the segments from the last one back. For each segment it
restores the checkpoint, re-runs the segment's forward steps while recording them
(sort permutation, spans, clamp bits, every iteration's positions and lambdas), and
applies the reverse kernel of each recorded step, last first.  The forward pass records
its last segment as it runs, so the reverse pass starts from those tapes.
``PBFRollout`` exposes X_T as a differentiable function of the initial velocity
and of the XSPH coefficient c. Gradients reach x0 as well.

``occupancy_grid`` is the loss's Gaussian splat: a G^3 grid over the box, each
particle contributing exp(-|x_c - p|^2 / (2 sigma^2)) to the cells within 3 sigma.
It is written in torch and differentiated by torch.
"""
import ctypes
import math
import time

import numpy as np
import torch

from . import _lib
from .params import build_sim_params, f32
from .sim import Simulator


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _as4(t):
    """(N, 3) or (N, 4) float tensor -> contiguous (N, 4) float32 on the same device."""
    if t.shape[1] == 4:
        return t.float().contiguous()
    out = torch.zeros((t.shape[0], 4), dtype=torch.float32, device=t.device)
    out[:, :3] = t
    return out


class Rollout:
    """T-step forward on the GPU, checkpointed every `checkpoint` steps, with a reverse pass."""

    def __init__(self, params, steps, checkpoint=10, grid="dense", table_log2=None, sim=None, scenes=1):
        self.params = params
        self.steps = int(steps)
        self.ck = max(1, int(checkpoint))
        if sim is None:
            tl = table_log2 or max(16, int(math.ceil(math.log2(max(params.numParticles, 1)))) + 1)
            sim = Simulator(params, grid=grid, table_log2=tl, use_graph=False)
        self.sim = sim
        self.scenes = int(scenes)  # dL/dc has one entry per scene
        if _lib.lib().pbfSimAdjointInit(self.sim._h, self.ck) != 0:
            _lib.check("pbfSimAdjointInit")
            raise _lib.PbfError("adjoint init failed")
        self.n = int(sim.n)
        self._ckpts = []
        self._tail_taped = False  # the tape slots hold the last segment of the latest forward()
        self.timed = False    # when set, backward() synchronises between its phases and fills `timing`
        self.timing = {}

    def close(self):
        self.sim.close()

    def _clock(self):
        if not self.timed:
            return 0.0
        self.sim.synchronize()
        return time.perf_counter()

    def _set_c(self, c):
        if c is None:
            return
        if float(f32(c)) != self.params.xsph_c:
            self.params.xsph_c = float(c)
            if _lib.lib().pbfSimSetParams(self.sim._h, ctypes.byref(self.params)) != 0:
                _lib.check("pbfSimSetParams")
                raise _lib.PbfError("set params failed")

    def _set_state(self, X, V):
        L = _lib.lib()
        torch.cuda.synchronize()
        if L.pbfSimSetState(self.sim._h, _p(X), _p(V), self.n, 1) != 0:
            _lib.check("pbfSimSetState")
            raise _lib.PbfError("set state failed")

    def _get_state(self):
        X = torch.empty((self.n, 4), dtype=torch.float32, device="cuda")
        V = torch.empty_like(X)
        if _lib.lib().pbfSimGetState(self.sim._h, _p(X), _p(V), 1) != 0:
            _lib.check("pbfSimGetState")
            raise _lib.PbfError("get state failed")
        return X, V

    def _params_now(self):
        out = type(self.params)()
        if _lib.lib().pbfSimGetParams(self.sim._h, ctypes.byref(out)) != 0:
            raise _lib.PbfError("get params failed")
        return out

    def _apply_params(self, sp):
        if _lib.lib().pbfSimSetParams(self.sim._h, ctypes.byref(sp)) != 0:
            _lib.check("pbfSimSetParams")
            raise _lib.PbfError("set params failed")

    def forward(self, x0, v0, c):
        """Run the rollout; returns (X_T, V_T) as (N, 4) float32 CUDA tensors."""
        self._set_c(c)
        X, V = _as4(x0), _as4(v0)
        self._set_state(X, V)
        self._ckpts = []
        self._tail_taped = False
        # backward() re-records the segments with these parameters, whatever changes later
        self._fwd_params = self._params_now()
        L = _lib.lib()
        done = 0
        while done < self.steps:
            self._ckpts.append(self._get_state())
            k = min(self.ck, self.steps - done)
            if done + k < self.steps:
                if L.pbfSimStep(self.sim._h, k) != 0:
                    _lib.check("pbfSimStep")
                    raise _lib.PbfError("step failed")
            else:
                # the last segment is recorded as it runs (a recorded step is the same
                # step, with its tape written beside it), so backward() starts from these
                # tapes instead of re-running the segment
                for t in range(k):
                    if L.pbfSimStepRecord(self.sim._h, t) != 0:
                        _lib.check("pbfSimStepRecord")
                        raise _lib.PbfError("record failed")
                self._tail_taped = True
            done += k
        return self._get_state()

    def backward(self, gX, gV=None):
        """dL/dX_0, dL/dV_0 ((N, 4) float64 CUDA tensors) and dL/dc from dL/dX_T and dL/dV_T
        (a float, or one float per scene for an ensemble rollout).

        The reverse pass is renormalised every step (pbfSimBackwardStepScaled): its float32
        buffers hold the gradients times a power of two whose exponent stays on the device,
        so a long rollout whose gradients outgrow float32 (the dam break amplifies them
        about 8x a step) still ends finite.  The result is that exponent applied in float64;
        ``log2_scale`` (one exponent, or one per scene of an ensemble) and ``scaled_grads``
        keep the exponents and the float32 mantissas."""
        if not self._ckpts:
            raise _lib.PbfError("backward() before forward()")
        L = _lib.lib()
        gX = _as4(gX).clone()
        gV = torch.zeros_like(gX) if gV is None else _as4(gV).clone()
        gX[:, 3] = 0
        gV[:, 3] = 0
        oX, oV = torch.empty_like(gX), torch.empty_like(gV)
        gc = torch.zeros(self.scenes, dtype=torch.float64, device="cuda")
        E = torch.zeros(self.scenes, dtype=torch.int32, device="cuda")
        now = self._params_now()
        changed = bytes(now) != bytes(self._fwd_params)
        if changed:
            self._apply_params(self._fwd_params)
        starts = list(range(0, self.steps, self.ck))
        t_rec = t_rev = 0.0
        # the reverse kernels run on the simulator's stream and read what torch's stream
        # wrote above (the seeds, the zeroed dL/dc and exponents)
        torch.cuda.synchronize()
        try:
            for seg in reversed(range(len(starts))):
                k = min(self.ck, self.steps - starts[seg])
                t0 = self._clock()
                if seg == len(starts) - 1 and self._tail_taped:
                    # forward() left this segment's tapes in the slots
                    self._tail_taped = False
                else:
                    self._tail_taped = False  # the slots are overwritten from here on
                    self._set_state(*self._ckpts[seg])
                    t0 = self._clock()
                    for t in range(k):
                        if L.pbfSimStepRecord(self.sim._h, t) != 0:
                            _lib.check("pbfSimStepRecord")
                            raise _lib.PbfError("record failed")
                t1 = self._clock()
                for t in reversed(range(k)):
                    if L.pbfSimBackwardStepScaled(self.sim._h, t, _p(gX), _p(gV), _p(oX), _p(oV), _p(gc),
                                                  _p(E)) != 0:
                        _lib.check("pbfSimBackwardStepScaled")
                        raise _lib.PbfError("backward failed")
                    gX, oX = oX, gX
                    gV, oV = oV, gV
                t2 = self._clock()
                t_rec += t1 - t0
                t_rev += t2 - t1
        finally:
            if changed:
                self._apply_params(now)
        if self.timed:
            self.timing = {"record_s": t_rec, "reverse_s": t_rev}
        self.sim.synchronize()
        _lib.check("Rollout.backward")
        exps = [int(e) for e in E.cpu()]
        self.scaled_grads = (gX, gV)
        if self.scenes == 1:
            self.log2_scale = exps[0]
            return _ldexp64(gX, exps[0]), _ldexp64(gV, exps[0]), float(gc.item())
        self.log2_scale = exps
        gX64 = torch.empty(gX.shape, dtype=torch.float64, device=gX.device)
        gV64 = torch.empty_like(gX64)
        for (a, b), e in zip(self.spans, exps):
            gX64[a:b] = _ldexp64(gX[a:b], e)
            gV64[a:b] = _ldexp64(gV[a:b], e)
        return gX64, gV64, gc.cpu().numpy()


def _ldexp64(t, e):
    """t * 2^e in float64 (exact; overflows to +-inf only past the float64 range)."""
    out = t.double()
    while e > 1000:
        out = out * 2.0 ** 1000
        e -= 1000
    while e < -1000:
        out = out * 2.0 ** -1000
        e += 1000
    return out * 2.0 ** e


class EnsembleRollout(Rollout):
    """Many rollouts advanced together as one batched ensemble (one launch per phase for
    all scenes, forward and reverse); dL/dc comes back per scene.  The XSPH coefficient
    of each scene is the one in its scene description."""

    def __init__(self, scenes, steps, checkpoint=10, grid="dense"):
        from .ensemble import Ensemble
        ens = Ensemble(scenes, grid=grid, use_graph=False)
        super().__init__(ens.base, steps, checkpoint, sim=ens, scenes=len(scenes))
        self.ensemble = ens
        self.firsts = [ens.scene_info(k)[0] for k in range(len(scenes))]
        self.spans = [(f, f + sc.num_particles) for f, sc in zip(self.firsts, scenes)]

    def forward(self, x0, v0, c=None):
        return super().forward(x0, v0, None)


class PBFRollout(torch.autograd.Function):
    """X_T = rollout(x0, v0, c) on the GPU; backward through the adjoint kernels."""

    @staticmethod
    def forward(ctx, v0, c, x0, roll):
        XT, VT = roll.forward(x0, v0, float(c))
        ctx.roll = roll
        return XT[:, :3].clone()

    @staticmethod
    def backward(ctx, gXT):
        gX0, gV0, gc = ctx.roll.backward(gXT.contiguous())
        return (gV0[:, :3].to(gXT.dtype), torch.tensor(gc, dtype=gXT.dtype, device=gXT.device),
                gX0[:, :3].to(gXT.dtype), None)


def occupancy_grid(pos, box_min, box_max, cells=32, sigma=0.0125, reach=3.0):
    """Gaussian splat of (N, 3) positions on a cells^3 grid spanning the box (cell centres)."""
    dev, dt = pos.device, pos.dtype
    lo = torch.tensor(box_min, dtype=dt, device=dev)
    hi = torch.tensor(box_max, dtype=dt, device=dev)
    size = (hi - lo) / cells
    R = int(math.ceil(reach * sigma / float(size.min())))
    base = torch.floor((pos - lo) / size).long()
    offs = torch.stack(torch.meshgrid(*[torch.arange(-R, R + 1, device=dev)] * 3, indexing="ij"), -1).reshape(-1, 3)
    idx = base[:, None, :] + offs[None, :, :]
    valid = ((idx >= 0) & (idx < cells)).all(-1)
    centres = lo + (idx.to(dt) + 0.5) * size
    d2 = ((centres - pos[:, None, :]) ** 2).sum(-1)
    w = torch.exp(-d2 / (2.0 * sigma * sigma)) * valid.to(dt)
    flat = (idx[..., 2].clamp(0, cells - 1) * cells + idx[..., 1].clamp(0, cells - 1)) * cells + \
        idx[..., 0].clamp(0, cells - 1)
    # accumulate through index_put, which sorts the contributions by cell and sums each
    # run: deterministic, and no queue of atomics on the few cells a collapsed pile covers
    # (index_add_ took 0.75 s for 65,536 piled particles on MI355X, this 4 ms)
    grid = torch.zeros(cells ** 3, dtype=dt, device=dev)
    grid = grid.index_put((flat.reshape(-1),), w.reshape(-1), accumulate=True)
    return grid.reshape(cells, cells, cells)


def differentiable_scene(nx=64, ny=32, nz=32, seed=7, box=(1.0, 1.0, 1.0)):
    """The differentiable rollout configuration: an nx*ny*nz lattice (spacing 2r, r = 0.005 m,
    seeded jitter) in the corner of the box, dt = 1/120, 2 iterations."""
    from .params import Scene
    r = 0.005
    return Scene(name="diff_rollout", lattices=[(nx, ny, nz, (r, r, r))], box_min=(0.0, 0.0, 0.0),
                 box_max=box, radius=r, grid="dense", seed=seed, steps=60)


def initial_state(scene):
    """Seeded lattice positions (the simulator's own seeding) as a host array."""
    sim = Simulator.from_scene(scene, use_graph=False)
    X, V = sim.state()
    sim.close()
    return X, V


def seeded_velocity(n, seed, scale):
    rng = np.random.Generator(np.random.PCG64(seed))
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0.0, scale, (n, 3)).astype(np.float32)
    return v


__all__ = ["Rollout", "PBFRollout", "occupancy_grid", "differentiable_scene", "initial_state", "seeded_velocity",
           "build_sim_params"]
