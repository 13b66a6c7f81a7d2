"""Differentiable SPH rollout: torch.autograd.Function over the hand-written
adjoint of include/rtps_c.h (rtps_adj_*).  torch provides the device memory
and the stream; forward and backward are HIP kernels.

    sim = AdjointSPH(settings, num_particles, checkpoint_every=20, max_steps=200)
    pos_T = sph_rollout(sim, pos0, vel0, viscosity, gas_constant, nsteps)
    loss = torch.nn.functional.mse_loss(pos_T[:, :3], target[:, :3])
    loss.backward()   # -> vel0.grad, viscosity.grad, gas_constant.grad

Many scenes at once (rtps_badj_*): BatchedAdjointSPH and batched_rollout,
with one viscosity and gas constant per scene.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N


class AdjointSPH:
    """Owns an rtps_adj handle (checkpoints and per-segment records)."""

    def __init__(self, settings: dict, num_particles: int, checkpoint_every: int = 20, max_steps: int = 200):
        ka, va, n = N.settings_arrays(settings)
        h = C.c_void_p()
        N.check(N.lib().rtps_adj_create(C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p), n, num_particles,
                                        checkpoint_every, max_steps, C.byref(h)))
        self._h = h
        self.num = num_particles
        # forward passes run so far: the handle keeps the checkpoints of the
        # last one only, which is what a backward replays
        self.generation = 0

    def close(self):
        if getattr(self, "_h", None):
            N.lib().rtps_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t: torch.Tensor, name: str):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and tuple(t.shape) == (self.num, 4)):
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of shape ({self.num}, 4)")

    def forward(self, pos0, vel0, nsteps: int, viscosity: float, gas_constant: float, collision: bool = True):
        self._check(pos0, "pos0")
        self._check(vel0, "vel0")
        pos_t = torch.empty_like(pos0)
        vel_t = torch.empty_like(pos0)
        st = torch.cuda.current_stream().cuda_stream
        N.check(N.lib().rtps_adj_forward(self._h, pos0.data_ptr(), vel0.data_ptr(), int(nsteps), float(viscosity),
                                         float(gas_constant), int(collision), pos_t.data_ptr(), vel_t.data_ptr(),
                                         st))
        self.generation += 1
        return pos_t, vel_t

    def backward(self, grad_pos_t, grad_vel_t=None):
        self._check(grad_pos_t, "grad_pos_T")
        dv0 = torch.empty_like(grad_pos_t)
        dx0 = torch.empty_like(grad_pos_t)
        dmu, dK = C.c_double(), C.c_double()
        st = torch.cuda.current_stream().cuda_stream
        gv = None
        if grad_vel_t is not None:
            self._check(grad_vel_t, "grad_vel_T")
            gv = grad_vel_t.data_ptr()
        N.check(N.lib().rtps_adj_backward(self._h, grad_pos_t.data_ptr(), gv, dv0.data_ptr(), dx0.data_ptr(),
                                          C.byref(dmu), C.byref(dK), st))
        return dv0, dx0, dmu.value, dK.value


def _replay_if_stale(ctx, visc, gas, pos0, vel0, scalars):
    """The handle holds the checkpoints of its latest forward only.  When
    another forward ran on it after this rollout's (a second rollout in the
    same graph, an evaluation pass), run this rollout's forward again from its
    saved inputs so the backward replays the right trajectory."""
    if ctx.sim.generation != ctx.generation:
        ctx.sim.forward(pos0.detach().contiguous(), vel0.detach().contiguous(), ctx.nsteps, scalars(visc),
                        scalars(gas), ctx.collision)
        ctx.generation = ctx.sim.generation


class _Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, viscosity, gas_constant, sim, pos0, nsteps, collision):
        pos_t, _ = sim.forward(pos0.detach().contiguous(), vel0.detach().contiguous(), nsteps,
                               float(viscosity), float(gas_constant), collision)
        ctx.sim = sim
        ctx.generation = sim.generation
        ctx.nsteps, ctx.collision = nsteps, collision
        ctx.save_for_backward(viscosity, gas_constant, pos0, vel0)
        return pos_t

    @staticmethod
    def backward(ctx, grad_pos_t):
        visc, gas, pos0, vel0 = ctx.saved_tensors
        _replay_if_stale(ctx, visc, gas, pos0, vel0, float)
        g = grad_pos_t.detach().contiguous().clone()
        g[:, 3] = 0.0
        dv0, _dx0, dmu, dK = ctx.sim.backward(g)
        dv0[:, 3] = 0.0
        return (dv0, torch.tensor(dmu, dtype=visc.dtype, device=visc.device),
                torch.tensor(dK, dtype=gas.dtype, device=gas.device), None, None, None, None)


def sph_rollout(sim: AdjointSPH, pos0: torch.Tensor, vel0: torch.Tensor, viscosity: torch.Tensor,
                gas_constant: torch.Tensor, nsteps: int, collision: bool = True) -> torch.Tensor:
    """Final positions after nsteps, differentiable in vel0, viscosity and gas_constant."""
    return _Rollout.apply(vel0, viscosity, gas_constant, sim, pos0, int(nsteps), bool(collision))


class BatchedAdjointSPH:
    """Owns an rtps_badj handle: forward and adjoint of many scenes at once.

    Scene s holds exactly the ``max_num_particles`` of its settings; the
    particle arrays are the scenes' arrays concatenated (``offsets[s]`` to
    ``offsets[s + 1]``).  Viscosity and gas constant are per scene.
    """

    def __init__(self, scene_settings: list, checkpoint_every: int = 20, max_steps: int = 100):
        flat_k, flat_v, counts = [], [], []
        for st in scene_settings:
            st = dict(st)
            counts.append(len(st))
            flat_k += [str(k).encode() for k in st]
            flat_v += [str(v).encode() for v in st.values()]
        ka = (C.c_char_p * max(1, len(flat_k)))(*flat_k)
        va = (C.c_char_p * max(1, len(flat_v)))(*flat_v)
        ca = (C.c_int * len(counts))(*counts)
        h = C.c_void_p()
        N.check(N.lib().rtps_badj_create(len(counts), ca, C.cast(ka, C.c_void_p), C.cast(va, C.c_void_p),
                                         checkpoint_every, max_steps, C.byref(h)))
        self._h = h
        ns, tot = C.c_int(), C.c_int()
        offs = (C.c_int * (len(counts) + 1))()
        N.check(N.lib().rtps_badj_layout(h, C.byref(ns), C.byref(tot), offs))
        self.num_scenes = ns.value
        self.num = tot.value
        self.offsets = list(offs)
        self.generation = 0   # forward passes so far (the handle keeps the latest one's checkpoints)

    def close(self):
        if getattr(self, "_h", None):
            N.lib().rtps_badj_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, t: torch.Tensor, name: str):
        if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and tuple(t.shape) == (self.num, 4)):
            raise ValueError(f"{name} must be a contiguous float32 CUDA tensor of shape ({self.num}, 4)")

    def _scalars(self, x, name: str):
        a = [float(v) for v in (x.tolist() if isinstance(x, torch.Tensor) else x)]
        if len(a) != self.num_scenes:
            raise ValueError(f"{name} needs one value per scene ({self.num_scenes})")
        return (C.c_float * self.num_scenes)(*a)

    def forward(self, pos0, vel0, nsteps: int, viscosity, gas_constant, collision: bool = True):
        self._check(pos0, "pos0")
        self._check(vel0, "vel0")
        pos_t = torch.empty_like(pos0)
        vel_t = torch.empty_like(pos0)
        st = torch.cuda.current_stream().cuda_stream
        N.check(N.lib().rtps_badj_forward(self._h, pos0.data_ptr(), vel0.data_ptr(), int(nsteps),
                                          self._scalars(viscosity, "viscosity"),
                                          self._scalars(gas_constant, "gas_constant"), int(collision),
                                          pos_t.data_ptr(), vel_t.data_ptr(), st))
        self.generation += 1
        return pos_t, vel_t

    def backward(self, grad_pos_t, grad_vel_t=None):
        self._check(grad_pos_t, "grad_pos_T")
        dv0 = torch.empty_like(grad_pos_t)
        dx0 = torch.empty_like(grad_pos_t)
        dmu = (C.c_double * self.num_scenes)()
        dK = (C.c_double * self.num_scenes)()
        st = torch.cuda.current_stream().cuda_stream
        gv = None
        if grad_vel_t is not None:
            self._check(grad_vel_t, "grad_vel_T")
            gv = grad_vel_t.data_ptr()
        N.check(N.lib().rtps_badj_backward(self._h, grad_pos_t.data_ptr(), gv, dv0.data_ptr(), dx0.data_ptr(),
                                           dmu, dK, st))
        return dv0, dx0, list(dmu), list(dK)


class _BatchedRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vel0, viscosity, gas_constant, sim, pos0, nsteps, collision):
        pos_t, _ = sim.forward(pos0.detach().contiguous(), vel0.detach().contiguous(), nsteps,
                               viscosity.detach().cpu(), gas_constant.detach().cpu(), collision)
        ctx.sim = sim
        ctx.generation = sim.generation
        ctx.nsteps, ctx.collision = nsteps, collision
        ctx.save_for_backward(viscosity, gas_constant, pos0, vel0)
        return pos_t

    @staticmethod
    def backward(ctx, grad_pos_t):
        visc, gas, pos0, vel0 = ctx.saved_tensors
        _replay_if_stale(ctx, visc, gas, pos0, vel0, lambda t: t.detach().cpu())
        g = grad_pos_t.detach().contiguous().clone()
        g[:, 3] = 0.0
        dv0, _dx0, dmu, dK = ctx.sim.backward(g)
        dv0[:, 3] = 0.0
        return (dv0, torch.tensor(dmu, dtype=visc.dtype, device=visc.device),
                torch.tensor(dK, dtype=gas.dtype, device=gas.device), None, None, None, None)


def batched_rollout(sim: BatchedAdjointSPH, pos0: torch.Tensor, vel0: torch.Tensor, viscosity: torch.Tensor,
                    gas_constant: torch.Tensor, nsteps: int, collision: bool = True) -> torch.Tensor:
    """Final positions of every scene after nsteps, differentiable in vel0 and
    in the per-scene viscosity and gas constant (tensors of num_scenes)."""
    return _BatchedRollout.apply(vel0, viscosity, gas_constant, sim, pos0, int(nsteps), bool(collision))
