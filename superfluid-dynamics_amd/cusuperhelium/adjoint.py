"""Differentiable RK4 trajectories of the gfx950 stepper (water, the helium film, the autonomous drive).

``evolve(Z0, Phi0, dt, steps, ...)`` is a ``torch.autograd.Function``: the forward pass advances
the state on the GPU (the same stepper as ``Ensemble.step``) and keeps a device-resident checkpoint
every ``checkpoint_every`` steps (``sfh_traj_forward``); the backward pass (``sfh_traj_backward``)
re-runs each segment from its checkpoint into a device buffer and propagates the cotangent back one
RK4 step at a time with the hand-written adjoint kernels (stage recomputation, four right-hand-side
VJPs with a transpose solve each), the states and the cotangent staying in device memory: the host
sends the initial state and the final cotangent once and receives the final state and the gradient
once.  Device memory: one state per checkpoint plus one segment of states.

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


class _Record:
    """A trajectory record held by the library (device checkpoints); freed with this object."""

    def __init__(self, ens: Ensemble, rid: int):
        self.ens, self.id = ens, rid

    def release(self):
        if self.id is not None:
            rid, self.id = self.id, None
            self.ens.traj_release(rid)

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass


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
        """(Z(T), Phi(T), record): the record keeps the checkpoints in device memory."""
        self.ens.set_state(Z0, P0)
        rec = _Record(self.ens, self.ens.traj_forward(dt, steps, self.checkpoint_every))
        Z, P = self.ens.get_state()
        return Z, P, rec

    def backward(self, rec: _Record, gZ: np.ndarray, gP: np.ndarray):
        """Cotangent of (Z0, Phi0) from the cotangent of (Z(T), Phi(T))."""
        if rec.id is None:
            raise RuntimeError("the trajectory record was released")
        return self.ens.traj_backward(rec.id, gZ, gP)


class _Evolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Z0, P0, traj, dt, steps):
        if steps < 1:
            raise ValueError("evolve needs steps >= 1")
        Zf, Pf, rec = traj.forward(Z0.detach().cpu().numpy(), P0.detach().cpu().numpy(), dt, steps)
        ctx.traj, ctx.rec = traj, rec
        return torch.from_numpy(Zf).to(Z0.device), torch.from_numpy(Pf).to(P0.device)

    @staticmethod
    def backward(ctx, gZ, gP):
        gZn = np.zeros(ctx.traj.ens.total, np.complex128) if gZ is None else gZ.detach().cpu().numpy()
        gPn = np.zeros(ctx.traj.ens.total) if gP is None else gP.detach().cpu().numpy()
        gZ0, gP0 = ctx.traj.backward(ctx.rec, gZn, gPn)
        dev = gZ.device if gZ is not None else gP.device
        return torch.from_numpy(gZ0).to(dev), torch.from_numpy(gP0).to(dev), None, None, None


def evolve(Z0: torch.Tensor, P0: torch.Tensor, traj: Trajectory, dt: float, steps: int):
    """(Z(T), Phi(T)) after `steps` RK4 steps of size dt, differentiable in (Z0, Phi0).

    Z0: complex128 tensor of all surfaces' nodes (concatenated), P0: float64 tensor."""
    if Z0.dtype != torch.complex128 or P0.dtype != torch.float64:
        raise TypeError("evolve takes complex128 Z0 and float64 Phi0")
    return _Evolve.apply(Z0, P0, traj, float(dt), int(steps))
