"""Density-field gradients through the GPU adjoint, as a torch.autograd.Function.

The function maps (initial positions, friction per member, normal stiffness)
to the CIC grids at the end of a window of semi-implicit Euler steps with the
spring_history tangential model.  Its backward is the engine's discrete
adjoint (wd_adjoint_backward): the exact derivative of the recorded forward
with its contact sets and stick/slip branches frozen, recomputed from
checkpoints every K steps (K = ceil(sqrt(window)) by default).

The reference solver has no differentiation; its surrogate side
(surrogate/) consumes CIC grids written by the dataset driver, which is the
quantity differentiated here.  Initial positions on the GPU stay there (torch
must be imported before the first world_amd call in the process: torch ships
its own copy of the HIP runtime, and the library then binds to it rather
than loading a second one): the
state, the grids and the gradients cross the C ABI as device buffers of the
handle's GPU (wd_set_state_device, wd_batch_density_grid_device,
wd_adjoint_backward_device); only the friction and stiffness, scalars, go
through the host.  CPU tensors take host buffers.

A window handle holds one recorded trajectory at a time.  Each forward
through it is numbered; a backward whose forward is no longer the window's
latest re-records that forward from its saved inputs first (the run is
deterministic, so it is the same trajectory bit for bit), so that
L(f(x_a)) + L(f(x_b)) differentiates each term along its own trajectory.
"""
from __future__ import annotations

import math

import numpy as np

from .sim import Ensemble, ParticleDynamics


class DensityWindow:
    """A reusable window: one engine handle (single or batched members) whose
    forward runs `nsteps` steps from given initial positions."""

    def __init__(self, configs, nsteps: int, checkpoint_every: int | None = None):
        if not isinstance(configs, (list, tuple)):
            configs = [configs]
        self.configs = [c.copy() for c in configs]
        self.nsteps = int(nsteps)
        self.every = int(checkpoint_every or math.ceil(math.sqrt(self.nsteps)))
        self.sim = Ensemble(self.configs) if len(self.configs) > 1 else ParticleDynamics(self.configs[0])
        self.members = len(self.configs)
        st = self.sim.state()
        self.v0 = st[:, 3:].copy()
        self.x0 = st[:, :3].copy()
        self.nodes = self.sim.density_grid_nodes()
        self.generation = 0  # number of the forward the handle holds

    def close(self) -> None:
        self.sim.close()

    def run(self, x0: np.ndarray, mu, k_n: float) -> np.ndarray:
        """Final CIC grids [members, 4, nodes] from initial positions x0 [n, 3]."""
        mu = np.broadcast_to(np.asarray(mu, np.float64), (self.members,))
        st = np.concatenate([np.asarray(x0, np.float32).reshape(-1, 3), self.v0], axis=1)
        for m in range(self.members):
            self.sim.set_friction(float(mu[m]), m)
        self.sim.set_normal_stiffness(float(k_n))
        self.sim.set_state(st)  # also clears the contact history
        self.sim.adjoint_record(self.nsteps, self.every)
        self.generation += 1
        if self.members > 1:
            grids = self.sim.density_grids()
        else:
            self.sim.compute_density_grid()
            grids = self.sim.host_density_grid[None, :].copy()
        return grids.reshape(self.members, 4, self.nodes)

    def backward(self, grid_cotangent: np.ndarray):
        """dL/dx0 [n, 3], dL/dmu [members], dL/dk_n for the last run."""
        gs, gp = self.sim.adjoint_backward(np.asarray(grid_cotangent, np.float32).reshape(-1))
        return gs[:, :3].copy(), gp[:, 0].copy(), float(gp[:, 1].sum())

    def run_device(self, x0, mu, k_n: float):
        """run() with x0 a torch tensor [n, 3] on the GPU: the grids come back
        as a device tensor [members, 4, nodes]; nothing crosses the host but
        the scalars."""
        import torch

        if not torch.cuda.is_available():
            raise RuntimeError(
                "torch sees no GPU in this process: the engine library was loaded before torch, which "
                "bundles a second copy of the HIP runtime.  Import torch before the first world_amd call "
                "so that the library binds to torch's runtime.")
        mu = np.broadcast_to(np.asarray(mu, np.float64), (self.members,))
        if getattr(self, "_v0_dev", None) is None or self._v0_dev.device != x0.device:
            self._v0_dev = torch.from_numpy(self.v0).to(x0.device)
        st = torch.cat([x0.detach().to(torch.float32).reshape(-1, 3), self._v0_dev], dim=1).contiguous()
        for m in range(self.members):
            self.sim.set_friction(float(mu[m]), m)
        self.sim.set_normal_stiffness(float(k_n))
        self.sim.set_state_device(st)  # also clears the contact history
        self.sim.adjoint_record(self.nsteps, self.every)
        self.generation += 1
        grids = torch.empty(self.members * 4 * self.nodes, dtype=torch.float32, device=x0.device)
        self.sim.density_grids_device(grids)
        return grids.reshape(self.members, 4, self.nodes)

    def backward_device(self, grid_cotangent):
        """backward() with the cotangent a device tensor: dL/dx0 [n, 3] comes
        back on the device."""
        import torch

        g = grid_cotangent.detach().to(torch.float32).contiguous().reshape(-1)
        gs, gp = self.sim.adjoint_backward_device(g)
        return gs[:, :3], gp[:, 0].copy(), float(gp[:, 1].sum())


def density_window_function():
    """The torch.autograd.Function class (torch is imported on first use)."""
    import torch

    class DensityWindowFunction(torch.autograd.Function):
        @staticmethod
        def forward(ctx, x0, mu, k_n, window: DensityWindow):
            ctx.on_device = x0.is_cuda
            params = (mu.detach().cpu().numpy().copy(), float(k_n.detach().cpu()))
            if ctx.on_device:
                inputs = (x0.detach().clone(),) + params
                grids = window.run_device(*inputs)
            else:
                inputs = (x0.detach().cpu().numpy().copy(),) + params
                grids = torch.from_numpy(window.run(*inputs))
            ctx.window = window
            ctx.inputs = inputs
            ctx.generation = window.generation
            ctx.meta = [(t.dtype, t.shape, t.device) for t in (x0, mu, k_n)]
            return grids

        @staticmethod
        def backward(ctx, g):
            win = ctx.window
            if win.generation != ctx.generation:
                # another forward ran through the window since: record this
                # one's trajectory again from its inputs
                (win.run_device if ctx.on_device else win.run)(*ctx.inputs)
                ctx.generation = win.generation
            if ctx.on_device:
                gx, gmu, gk = win.backward_device(g)
            else:
                gx, gmu, gk = win.backward(g.detach().cpu().numpy())
            if ctx.meta[1][1].numel() == 1:
                gmu = np.asarray(gmu).sum()  # one friction broadcast to every member
            out = []
            for val, (dtype, shape, dev) in zip((gx, gmu, np.array(gk)), ctx.meta):
                val = val if isinstance(val, torch.Tensor) else torch.as_tensor(np.asarray(val))
                out.append(val.reshape(shape).to(device=dev, dtype=dtype))
            return out[0], out[1], out[2], None

    return DensityWindowFunction


def density_window(x0, mu, k_n, window: DensityWindow):
    """Differentiable final CIC grids [members, 4, nodes] of `window`."""
    return density_window_function().apply(x0, mu, k_n, window)
