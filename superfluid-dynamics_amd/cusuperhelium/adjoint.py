"""Differentiable RK4 trajectories of the gfx950 stepper (water with rho = 0, the helium film).

``evolve(Z0, Phi0, dt, steps, ...)`` is a ``torch.autograd.Function``: the forward pass advances
the state on the GPU (the same stepper as ``Ensemble.step``) and keeps a checkpoint every
``checkpoint_every`` steps; the backward pass re-runs each segment from its checkpoint to recover
the per-step states and propagates the cotangent back one RK4 step at a time with the hand-written
adjoint kernels (``sfh_step_vjp``: stage recomputation, four right-hand-side VJPs with a
transpose solve each).  Memory: one state per checkpoint plus one segment of states.

The reference has no adjoint; the forward map is AutonomousRungeKuttaStepper::runStep over
BaseBoundaryIntegralCalculator::runTimeStep (CuSuperHelium/AutonomousRungeKuttaStepper.cuh:338-375,
BaseBoundaryIntegrator.cuh:137-314).  Complex gradients follow torch's convention for real losses:
grad = dL/dRe + i dL/dIm.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from .stepper import BATCHED, FP64, Ensemble, Props

__all__ = ["evolve", "Trajectory"]


class Trajectory:
    """A GPU handle plus the checkpointing policy used by :func:`evolve`."""

    def __init__(self, sizes: Sequence[int] | int, props: Props | None = None, precision: int = FP64,
                 solver: int = BATCHED, tol: float = 1e-13, max_iter: int | None = None,
                 checkpoint_every: int = 20):
        # solver: BATCHED runs the reverse pass in one workgroup per surface (every N <= 1024), LU and
        # GMRES over many workgroups per surface (single large surfaces)
        if max_iter is None:
            max_iter = 48 if solver == BATCHED else 32
        self.ens = Ensemble(sizes, props or Props(), precision=precision, solver=solver, tol=tol,
                            max_iter=max_iter)
        self.checkpoint_every = int(checkpoint_every)
        if self.checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")

    def close(self):
        self.ens.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def forward(self, Z0: np.ndarray, P0: np.ndarray, dt: float, steps: int):
        ck = []
        self.ens.set_state(Z0, P0)
        done = 0
        while done < steps:
            ck.append(self.ens.get_state())
            n = min(self.checkpoint_every, steps - done)
            self.ens.step(dt, n)
            done += n
        Z, P = self.ens.get_state()
        return Z, P, ck

    def backward(self, ck, dt: float, steps: int, gZ: np.ndarray, gP: np.ndarray):
        gZ = np.array(gZ, dtype=np.complex128)
        gP = np.array(gP, dtype=np.float64)
        starts = list(range(0, steps, self.checkpoint_every))
        for seg in range(len(starts) - 1, -1, -1):
            n = min(self.checkpoint_every, steps - starts[seg])
            # states y_{s}, ..., y_{s+n-1} of the segment
            self.ens.set_state(*ck[seg])
            states = [ck[seg]]
            for _ in range(n - 1):
                self.ens.step(dt, 1)
                states.append(self.ens.get_state())
            for i in range(n - 1, -1, -1):
                self.ens.set_state(*states[i])
                gZ, gP = self.ens.step_vjp(dt, gZ, gP)
        return gZ, gP


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Z0, P0, traj, dt, steps):
        Zf, Pf, ck = traj.forward(Z0.detach().cpu().numpy(), P0.detach().cpu().numpy(), dt, steps)
        ctx.traj, ctx.dt, ctx.steps, ctx.ck = traj, dt, steps, ck
        return torch.from_numpy(Zf).to(Z0.device), torch.from_numpy(Pf).to(P0.device)

    @staticmethod
    def backward(ctx, gZ, gP):
        gZn = np.zeros(ctx.traj.ens.total, np.complex128) if gZ is None else gZ.detach().cpu().numpy()
        gPn = np.zeros(ctx.traj.ens.total) if gP is None else gP.detach().cpu().numpy()
        gZ0, gP0 = ctx.traj.backward(ctx.ck, ctx.dt, ctx.steps, gZn, gPn)
        dev = gZ.device if gZ is not None else gP.device
        return torch.from_numpy(gZ0).to(dev), torch.from_numpy(gP0).to(dev), None, None, None


def evolve(Z0: torch.Tensor, P0: torch.Tensor, traj: Trajectory, dt: float, steps: int):
    """(Z(T), Phi(T)) after `steps` RK4 steps of size dt, differentiable in (Z0, Phi0).

    Z0: complex128 tensor of all surfaces' nodes (concatenated), P0: float64 tensor."""
    if Z0.dtype != torch.complex128 or P0.dtype != torch.float64:
        raise TypeError("evolve takes complex128 Z0 and float64 Phi0")
    return _Evolve.apply(Z0, P0, traj, float(dt), int(steps))
