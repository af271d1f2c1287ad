"""Python front end of the native handle API (``sfh_*`` in include/cusuperhelium.h).

``Ensemble`` holds S independent periodic surfaces (any mix of node counts) resident in HBM and
advances them with the fused gfx950 RK4 stepper.  It plays the role of the reference's
``BaseBoundaryIntegralCalculator`` + ``AutonomousRungeKuttaStepper`` pair
(CuSuperHelium/BaseBoundaryIntegrator.cuh:10-85, AutonomousRungeKuttaStepper.cuh:24-89):
``rhs()`` is ``run(state, rhs)`` and ``step(dt, n)`` is n calls of ``runStep``.
"""
from __future__ import annotations

from ctypes import POINTER, byref, c_double, c_int, c_void_p
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from ._lib import SfhDrive, SfhOptions, SfhProps, SfhStats, SfhTuning, check, lib

WATER, HELIUM = 0, 1
FP64, MIXED = 0, 1
LU, GMRES, BATCHED, AUTO = 0, 1, 2, 3


def _dp(a):
    return None if a is None else a.ctypes.data_as(POINTER(c_double))


@dataclass
class Props:
    """Nondimensional problem (ProblemProperties.hpp:5-32) plus kind and gravity scale g."""

    kind: int = WATER
    rho: float = 0.0
    kappa: float = 0.0
    depth: float = 1.0
    U: float = 0.0
    g: float = 1.0
    infinite_depth: bool = False
    use_expansions: bool = False
    expansion_order: int = 1
    rho_quirk: bool = True

    def c(self) -> SfhProps:
        return SfhProps(int(self.kind), int(self.infinite_depth), int(self.use_expansions),
                        int(self.expansion_order), int(self.rho_quirk), 0, float(self.rho),
                        float(self.kappa), float(self.depth), float(self.U), float(self.g))


class Ensemble:
    def __init__(self, sizes: Sequence[int] | int, props: Props | None = None, precision: int = FP64,
                 solver: int = LU, use_graph: bool = True, tol: float = 1e-12, max_iter: int = 32,
                 force_trim: int = 0, pair_rows: int = 0):
        self.sizes = [int(sizes)] if np.isscalar(sizes) else [int(n) for n in sizes]
        self.props = props or Props()
        self._lib = lib()
        opt = SfhOptions()
        self._lib.sfh_default_options(byref(opt))
        opt.precision, opt.solver, opt.use_graph = int(precision), int(solver), int(use_graph)
        opt.tol, opt.max_iter, opt.force_trim = float(tol), int(max_iter), int(force_trim)
        opt.pair_rows = int(pair_rows)
        arr = (c_int * len(self.sizes))(*self.sizes)
        h = c_void_p()
        cp = self.props.c()
        check(self._lib.sfh_create(byref(cp), arr, len(self.sizes), byref(opt), byref(h)), "sfh_create")
        self._h = h
        self.total = int(sum(self.sizes))
        self.augmented = False
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sfh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- state
    def set_state(self, Z, phi) -> None:
        Z = np.asarray(Z, dtype=np.complex128).ravel()
        x = np.ascontiguousarray(Z.real)
        y = np.ascontiguousarray(Z.imag)
        p = np.ascontiguousarray(phi, dtype=np.float64).ravel()
        if x.size != self.total or p.size != self.total:
            raise ValueError(f"state must have {self.total} nodes")
        check(self._lib.sfh_set_state(self._h, _dp(x), _dp(y), _dp(p)), "sfh_set_state")

    def get_state(self):
        x, y, p = (np.empty(self.total) for _ in range(3))
        check(self._lib.sfh_get_state(self._h, _dp(x), _dp(y), _dp(p)), "sfh_get_state")
        return x + 1j * y, p

    def split(self, v):
        return [v[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.sizes))]

    # ---- evaluation
    def rhs(self, diagnostics: bool = False):
        """(q = dZ/dt, dPhi/dt[, diag]) of the current state."""
        vx, vy, dp = (np.empty(self.total) for _ in range(3))
        mu = np.empty(self.total) if diagnostics else None
        zp = np.empty(2 * self.total) if diagnostics else None
        zpp = np.empty(2 * self.total) if diagnostics else None
        qup = np.empty(2 * self.total) if diagnostics else None
        check(self._lib.sfh_rhs(self._h, _dp(vx), _dp(vy), _dp(dp), _dp(mu), _dp(zp), _dp(zpp), _dp(qup)),
              "sfh_rhs")
        q = vx + 1j * vy
        if diagnostics:
            return q, dp, {"mu": mu, "Zp": zp.view(np.complex128), "Zpp": zpp.view(np.complex128),
                           "qup": qup.view(np.complex128)}
        return q, dp

    def step(self, dt: float, nsteps: int = 1) -> None:
        check(self._lib.sfh_step(self._h, float(dt), int(nsteps)), "sfh_step")

    # ---- optical drive (HeliumDrivenAutonomousProblem / AugmentedBoundaryIntegrator)
    def set_drive(self, augmented: bool = False, timed: bool = False, **kw) -> None:
        """Turns the optical drive on (keywords of sfh_drive: damping, strength, detuning, gamma, G,
        max_intensity, x0, sigma, beta, tau), or off with set_drive(None).  timed: the time-dependent
        delayed intensity (its clock starts at set_time's time)."""
        if augmented is None:
            check(self._lib.sfh_set_drive(self._h, None), "sfh_set_drive")
            self.augmented = False
            return
        d = SfhDrive()
        d.augmented = int(bool(augmented))
        d.timed = int(bool(timed))
        d.gamma, d.G, d.sigma, d.tau = 1.0, 1.0, 1.0, 1.0
        for k, v in kw.items():
            setattr(d, k, float(v))
        check(self._lib.sfh_set_drive(self._h, byref(d)), "sfh_set_drive")
        self.augmented = bool(augmented)

    def set_time(self, t: float) -> None:
        check(self._lib.sfh_set_time(self._h, float(t)), "sfh_set_time")

    def set_delayed(self, D) -> None:
        D = np.ascontiguousarray(D, dtype=np.float64).ravel()
        check(self._lib.sfh_set_delayed(self._h, _dp(D)), "sfh_set_delayed")

    def delayed(self) -> np.ndarray:
        out = np.empty(self.total)
        check(self._lib.sfh_get_delayed(self._h, _dp(out)), "sfh_get_delayed")
        return out

    def delayed_rhs(self) -> np.ndarray:
        out = np.empty(self.total)
        check(self._lib.sfh_get_delayed_rhs(self._h, _dp(out)), "sfh_get_delayed_rhs")
        return out

    def step_record(self, dt: float, nsteps: int, every: int = 1):
        """Advances nsteps steps and returns the states after every `every`-th one:
        (Z [frames, total] complex, phi [frames, total], times [frames])."""
        frames = int(nsteps) // int(every)
        fields = 4 if getattr(self, "augmented", False) else 3
        st = np.empty((max(frames, 1), fields * self.total))
        tm = np.empty(max(frames, 1))
        check(self._lib.sfh_step_record(self._h, float(dt), int(nsteps), int(every), _dp(st), _dp(tm)),
              "sfh_step_record")
        st, tm = st[:frames], tm[:frames]
        T = self.total
        out = (st[:, :T] + 1j * st[:, T:2 * T], st[:, 2 * T:3 * T], tm)
        return out + (st[:, 3 * T:],) if fields == 4 else out

    # ---- reverse mode (water, the helium film, the autonomous drive)
    def rhs_vjp(self, gq, gdphi, gdD=None):
        """Cotangent of (Z, Phi) from cotangents of (q = dZ/dt, dPhi/dt) at the current state; with
        the augmented drive gdD, the cotangent of dD/dt, is required and the cotangent of D is
        returned third.

        Complex cotangents follow dL = Re sum conj(g) dz, i.e. g = dL/dRe + i dL/dIm."""
        gq = np.asarray(gq, dtype=np.complex128).ravel()
        gvx, gvy = np.ascontiguousarray(gq.real), np.ascontiguousarray(gq.imag)
        gp = np.ascontiguousarray(gdphi, dtype=np.float64).ravel()
        gx, gy, gphi = (np.empty(self.total) for _ in range(3))
        if gdD is None:
            check(self._lib.sfh_rhs_vjp(self._h, _dp(gvx), _dp(gvy), _dp(gp), _dp(gx), _dp(gy), _dp(gphi)),
                  "sfh_rhs_vjp")
            return gx + 1j * gy, gphi
        gd = np.ascontiguousarray(gdD, dtype=np.float64).ravel()
        gD = np.empty(self.total)
        check(self._lib.sfh_rhs_vjp_augmented(self._h, _dp(gvx), _dp(gvy), _dp(gp), _dp(gd), _dp(gx), _dp(gy),
                                              _dp(gphi), _dp(gD)), "sfh_rhs_vjp_augmented")
        return gx + 1j * gy, gphi, gD

    def step_vjp(self, dt: float, gZ, gPhi, gD=None):
        """Cotangent of the current state from the cotangent of the state one RK4 step later (with the
        augmented drive also of D: gD in, the cotangent of the current D returned third)."""
        gZ = np.asarray(gZ, dtype=np.complex128).ravel()
        gxi, gyi = np.ascontiguousarray(gZ.real), np.ascontiguousarray(gZ.imag)
        gpi = np.ascontiguousarray(gPhi, dtype=np.float64).ravel()
        gx, gy, gp = (np.empty(self.total) for _ in range(3))
        if gD is None:
            check(self._lib.sfh_step_vjp(self._h, float(dt), _dp(gxi), _dp(gyi), _dp(gpi), _dp(gx), _dp(gy),
                                         _dp(gp)), "sfh_step_vjp")
            return gx + 1j * gy, gp
        gdi = np.ascontiguousarray(gD, dtype=np.float64).ravel()
        gdo = np.empty(self.total)
        check(self._lib.sfh_step_vjp_augmented(self._h, float(dt), _dp(gxi), _dp(gyi), _dp(gpi), _dp(gdi), _dp(gx),
                                               _dp(gy), _dp(gp), _dp(gdo)), "sfh_step_vjp_augmented")
        return gx + 1j * gy, gp, gdo

    def traj_forward(self, dt: float, steps: int, checkpoint_every: int) -> int:
        """Advance `steps` RK4 steps like step(), keeping a device-resident checkpoint every
        `checkpoint_every` steps; returns the record id for traj_backward."""
        rec = c_int(0)
        check(self._lib.sfh_traj_forward(self._h, float(dt), int(steps), int(checkpoint_every), byref(rec)),
              "sfh_traj_forward")
        return rec.value

    def traj_backward(self, record: int, gZ, gPhi):
        """Cotangent of a record's initial state from the cotangent of its final state (the whole
        checkpointed reverse sweep runs in the library, the cotangent in device memory)."""
        gZ = np.asarray(gZ, dtype=np.complex128).ravel()
        gxi, gyi = np.ascontiguousarray(gZ.real), np.ascontiguousarray(gZ.imag)
        gpi = np.ascontiguousarray(gPhi, dtype=np.float64).ravel()
        gx, gy, gp = (np.empty(self.total) for _ in range(3))
        check(self._lib.sfh_traj_backward(self._h, int(record), _dp(gxi), _dp(gyi), _dp(gpi), _dp(gx), _dp(gy),
                                          _dp(gp)), "sfh_traj_backward")
        return gx + 1j * gy, gp

    def traj_release(self, record: int):
        if self._h is not None:
            check(self._lib.sfh_traj_release(self._h, int(record)), "sfh_traj_release")

    def mu(self) -> np.ndarray:
        out = np.empty(self.total)
        check(self._lib.sfh_get_mu(self._h, _dp(out)), "sfh_get_mu")
        return out

    def track_energies(self, enable: bool = True) -> None:
        """Per-RHS energy terms on (EnergyContainer + VolumeFlux, Energies.cuh:136-204)."""
        check(self._lib.sfh_track_energies(self._h, int(bool(enable))), "sfh_track_energies")

    def energies(self) -> np.ndarray:
        """[surfaces, 4]: kinetic, potential, surface, volume flux of the last right-hand side."""
        out = np.empty(4 * len(self.sizes))
        check(self._lib.sfh_get_energies(self._h, _dp(out)), "sfh_get_energies")
        return out.reshape(len(self.sizes), 4)

    def sync(self) -> None:
        check(self._lib.sfh_sync(self._h), "sfh_sync")

    def stats(self) -> dict:
        st = SfhStats()
        check(self._lib.sfh_get_stats(self._h, byref(st)), "sfh_get_stats")
        return {k: getattr(st, k) for k, _ in SfhStats._fields_ if k != "reserved"}

    def kernel_timing(self, enable: bool) -> None:
        check(self._lib.sfh_kernel_timing(self._h, int(bool(enable))), "sfh_kernel_timing")

    def kernel_report(self, tag: str) -> dict:
        n = c_int()
        ms, fl, by = c_double(), c_double(), c_double()
        check(self._lib.sfh_kernel_report(self._h, tag.encode(), byref(n), byref(ms), byref(fl), byref(by)),
              "sfh_kernel_report")
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}

    def device_state(self):
        z, p = c_void_p(), c_void_p()
        check(self._lib.sfh_device_state(self._h, byref(z), byref(p)), "sfh_device_state")
        return z.value, p.value


def _tuning(t: SfhTuning) -> dict:
    return {"solver": t.solver, "precision": t.precision, "pair_rows": t.pair_rows, "from_cache": bool(t.from_cache),
            "ms_per_step": t.ms_per_step}


def autotune(n: int, props: Props | None = None, allow_mixed: bool = False, steps: int = 10) -> dict:
    """Times the candidate forms for one surface of n nodes and caches the fastest (sfh_autotune)."""
    t = SfhTuning()
    cp = (props or Props()).c()
    check(lib().sfh_autotune(byref(cp), int(n), int(bool(allow_mixed)), int(steps), byref(t)), "sfh_autotune")
    return _tuning(t)


def tuned(n: int, props: Props | None = None, allow_mixed: bool = False) -> dict:
    """The cached autotuner result for one surface of n nodes, tuning on a miss (sfh_tuned)."""
    t = SfhTuning()
    cp = (props or Props()).c()
    check(lib().sfh_tuned(byref(cp), int(n), int(bool(allow_mixed)), byref(t)), "sfh_tuned")
    return _tuning(t)
