"""ctypes front end of the CPU FP64 oracle (``oracle/libsfo.so``).

TEST INFRASTRUCTURE ONLY: the test-suite, the smoke check and the CPU-baseline leg of the
benchmark use this module as the checker.  The product (``superfluid-dynamics_amd``) never
imports it.

Every function mirrors one reference routine; the C source (``oracle/sfo.c``) cites the
reference file:line for each.  Arrays are numpy; complex arrays are ``complex128``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsfo.so")
_lib = None


class _Props(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("infinite_depth", ctypes.c_int),
        ("use_expansions", ctypes.c_int),
        ("expansion_order", ctypes.c_int),
        ("rho_quirk", ctypes.c_int),
        ("image_zp", ctypes.c_int),
        ("rho", ctypes.c_double),
        ("kappa", ctypes.c_double),
        ("depth", ctypes.c_double),
        ("U", ctypes.c_double),
        ("g", ctypes.c_double),
    ]


class _Drive(ctypes.Structure):
    _fields_ = [("augmented", ctypes.c_int), ("pad_", ctypes.c_int)] + [
        (k, ctypes.c_double) for k in ("damping", "strength", "detuning", "gamma", "G", "max_intensity", "x0",
                                       "sigma", "beta", "tau")]


class _Diag(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("mu", "Zp", "Zpp", "Phip", "mup", "qup")]


WATER = 0
HELIUM = 1


@dataclass
class Props:
    """Nondimensional problem description (ProblemProperties.hpp:5-32 plus the problem kind)."""

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
    image_zp: bool = False   # test-only: film image term of M with the Z'_k factor (not the reference)

    def c(self) -> _Props:
        return _Props(int(self.kind), int(self.infinite_depth), int(self.use_expansions),
                      int(self.expansion_order), int(self.rho_quirk), int(self.image_zp), float(self.rho),
                      float(self.kappa), float(self.depth), float(self.U), float(self.g))


@dataclass
class Drive:
    """Nondimensional optical drive (OptomechanicalVariables, OptomechanicalVariables.h:4-27, after
    adimensionalizeOptomechanicalVariables); strength = get_current_intensity_drive_strength."""

    augmented: bool = False
    damping: float = 0.0
    strength: float = 0.0
    detuning: float = 0.0
    gamma: float = 1.0
    G: float = 1.0
    max_intensity: float = 0.0
    x0: float = 0.0
    sigma: float = 1.0
    beta: float = 0.0
    tau: float = 1.0

    def c(self) -> _Drive:
        return _Drive(int(self.augmented), 0, float(self.damping), float(self.strength), float(self.detuning),
                      float(self.gamma), float(self.G), float(self.max_intensity), float(self.x0),
                      float(self.sigma), float(self.beta), float(self.tau))


def build(force: bool = False) -> str:
    """Compile ``libsfo.so`` with the committed Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.sfo_cot.argtypes = [ctypes.c_double, ctypes.c_double, dp]
        L.sfo_fft.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.sfo_fft_derivative.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.sfo_zphi_derivative.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, ctypes.c_int,
                                          dp, dp, dp, dp, dp]
        L.sfo_assemble.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.sfo_lu_factor.argtypes = [ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
        L.sfo_lu_factor.restype = ctypes.c_int
        L.sfo_lu_solve.argtypes = [ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int), dp]
        L.sfo_velocity.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                   dp, dp, ctypes.c_int, dp]
        L.sfo_rhs_phi.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.sfo_rhs.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, ctypes.c_int, dp, dp,
                              ctypes.POINTER(_Diag)]
        L.sfo_rhs.restype = ctypes.c_int
        L.sfo_rk4.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, ctypes.c_int, ctypes.c_double,
                              ctypes.c_int, dp, dp]
        L.sfo_rk4.restype = ctypes.c_int
        L.sfo_rk4_rotation.argtypes = [ctypes.c_size_t, ctypes.c_double, ctypes.c_int, dp]
        L.sfo_adimensionalize.argtypes = [ctypes.c_double] * 4 + [dp]
        L.sfo_intensity.argtypes = [ctypes.POINTER(_Drive), ctypes.c_double, ctypes.c_double]
        L.sfo_intensity.restype = ctypes.c_double
        L.sfo_rhs_driven.argtypes = [ctypes.POINTER(_Props), ctypes.POINTER(_Drive), ctypes.c_int, ctypes.c_int,
                                     dp, dp]
        L.sfo_rhs_driven.restype = ctypes.c_int
        L.sfo_rk4_driven.argtypes = [ctypes.POINTER(_Props), ctypes.POINTER(_Drive), ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int, dp, dp]
        L.sfo_rk4_driven.restype = ctypes.c_int
        L.sfo_rk4_timed.argtypes = [ctypes.POINTER(_Props), ctypes.POINTER(_Drive), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, dp, dp, dp]
        L.sfo_rk4_timed.restype = ctypes.c_int
        L.sfo_jacobian.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, dp, ctypes.c_double, dp]
        L.sfo_jacobian.restype = ctypes.c_int
        L.sfo_gl2.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                              ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_int]
        L.sfo_gl2.restype = ctypes.c_int
        L.sfo_gl2_rotation.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.sfo_gl2_rotation.restype = ctypes.c_int
        L.sfo_energies.argtypes = [ctypes.POINTER(_Props), ctypes.c_int, dp, dp, dp, dp, dp]
        L.sfo_num_threads.restype = ctypes.c_int
        L.sfo_set_num_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return lib().sfo_num_threads()


def set_num_threads(n: int) -> None:
    lib().sfo_set_num_threads(int(n))


def cot(z: complex) -> complex:
    out = np.zeros(2)
    lib().sfo_cot(float(np.real(z)), float(np.imag(z)), _p(out))
    return complex(out[0], out[1])


def fft(a, sign: int = -1) -> np.ndarray:
    x = _c128(a).copy()
    lib().sfo_fft(_p(x.view(np.float64)), x.size, int(sign))
    return x


def fft_derivative(a, second: bool = False, scale: float = 1.0) -> np.ndarray:
    x = _c128(a)
    out = np.empty_like(x)
    lib().sfo_fft_derivative(_p(x.view(np.float64)), _p(out.view(np.float64)), x.size,
                             int(second), float(scale))
    return out


def zphi_derivative(props: Props, Z, Phi, batch: int = 1):
    Z = _c128(Z)
    Phi = _c128(Phi)
    n = Z.size // batch
    Zp, Zpp, Php = (np.empty_like(Z) for _ in range(3))
    cp = props.c()
    lib().sfo_zphi_derivative(ctypes.byref(cp), n, batch, _p(Z.view(np.float64)),
                              _p(Phi.view(np.float64)), _p(Zp.view(np.float64)),
                              _p(Zpp.view(np.float64)), _p(Php.view(np.float64)))
    return Zp, Zpp, Php


def assemble(props: Props, Z, Zp, Zpp) -> np.ndarray:
    Z, Zp, Zpp = _c128(Z), _c128(Zp), _c128(Zpp)
    n = Z.size
    M = np.empty((n, n))
    cp = props.c()
    lib().sfo_assemble(ctypes.byref(cp), n, 1, _p(Z.view(np.float64)), _p(Zp.view(np.float64)),
                       _p(Zpp.view(np.float64)), _p(M))
    return M


def lu_factor(M):
    A = _f64(M).copy()
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib().sfo_lu_factor(n, _p(A), piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return A, piv, info


def lu_solve(LU, piv, b) -> np.ndarray:
    x = _f64(b).copy()
    LU = _f64(LU)
    piv = np.ascontiguousarray(piv, dtype=np.int32)
    lib().sfo_lu_solve(LU.shape[0], _p(LU), piv.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _p(x))
    return x


def velocity(props: Props, Z, Zp, Zpp, mu, mup, lower: bool = True) -> np.ndarray:
    Z, Zp, Zpp, mup = _c128(Z), _c128(Zp), _c128(Zpp), _c128(mup)
    mu = _f64(mu)
    q = np.empty_like(Z)
    cp = props.c()
    lib().sfo_velocity(ctypes.byref(cp), Z.size, 1, _p(Z.view(np.float64)), _p(Zp.view(np.float64)),
                       _p(Zpp.view(np.float64)), _p(mu), _p(mup.view(np.float64)), int(lower),
                       _p(q.view(np.float64)))
    return q


def rhs_phi(props: Props, Z, Zp, Zpp, qlow, qup) -> np.ndarray:
    Z, Zp, Zpp, qlow, qup = (_c128(v) for v in (Z, Zp, Zpp, qlow, qup))
    out = np.empty(Z.size)
    cp = props.c()
    lib().sfo_rhs_phi(ctypes.byref(cp), Z.size, _p(Z.view(np.float64)), _p(Zp.view(np.float64)),
                      _p(Zpp.view(np.float64)), _p(qlow.view(np.float64)), _p(qup.view(np.float64)),
                      _p(out))
    return out


def pack_state(Z, Phi) -> np.ndarray:
    """Reference complex state [Z, Phi + 0i] (BaseBoundaryIntegrator.cuh:141-146)."""
    return np.concatenate([_c128(Z).ravel(), _c128(np.asarray(Phi, dtype=np.float64)).ravel()])


def rhs(props: Props, Z, Phi, batch: int = 1, diagnostics: bool = False):
    """Returns (q = dZ/dt complex, dPhi/dt real[, diag dict])."""
    state = pack_state(Z, Phi)
    tot = state.size // 2
    n = tot // batch
    out = np.empty_like(state)
    cp = props.c()
    d = _Diag()
    arrs = {}
    if diagnostics:
        arrs = {"mu": np.empty(tot), "Zp": np.empty(tot, np.complex128),
                "Zpp": np.empty(tot, np.complex128), "Phip": np.empty(tot, np.complex128),
                "mup": np.empty(tot, np.complex128), "qup": np.empty(tot, np.complex128)}
        for k, v in arrs.items():
            setattr(d, k, v.ctypes.data)
    info = lib().sfo_rhs(ctypes.byref(cp), n, batch, _p(state.view(np.float64)),
                         _p(out.view(np.float64)), ctypes.byref(d) if diagnostics else None)
    q = out[:tot].copy()
    dphi = out[tot:].real.copy()
    if diagnostics:
        arrs["info"] = info
        return q, dphi, arrs
    return q, dphi


def rk4(props: Props, Z, Phi, dt: float, steps: int, batch: int = 1, trajectory: bool = False):
    """Advances (Z, Phi) by `steps` classical RK4 steps; returns (Z, Phi[, traj])."""
    state = pack_state(Z, Phi)
    tot = state.size // 2
    n = tot // batch
    traj = np.empty((steps, state.size), np.complex128) if trajectory else None
    cp = props.c()
    lib().sfo_rk4(ctypes.byref(cp), n, batch, float(dt), int(steps), _p(state.view(np.float64)),
                  _p(traj.view(np.float64)) if trajectory else None)
    Zo = state[:tot].copy()
    Po = state[tot:].real.copy()
    if trajectory:
        return Zo, Po, traj
    return Zo, Po


def rk4_rotation(z0, dt: float, steps: int) -> np.ndarray:
    y = _c128(z0).copy()
    lib().sfo_rk4_rotation(y.size, float(dt), int(steps), _p(y.view(np.float64)))
    return y


def intensity(drive: Drive, y: float, x: float) -> float:
    cd = drive.c()
    return lib().sfo_intensity(ctypes.byref(cd), float(y), float(x))


def _driven_state(drive, Z, Phi, D):
    parts = [_c128(Z).ravel(), _c128(np.asarray(Phi, dtype=np.float64)).ravel()]
    if drive.augmented:
        parts.append(_c128(np.asarray(D, dtype=np.float64)).ravel())
    return np.concatenate(parts)


def rhs_driven(props: Props, drive: Drive, Z, Phi, D=None, batch: int = 1):
    """(q, dPhi/dt, dD/dt or None) of the driven (augmented) helium system."""
    state = _driven_state(drive, Z, Phi, D)
    tot = _c128(Z).size
    out = np.empty_like(state)
    cp, cd = props.c(), drive.c()
    lib().sfo_rhs_driven(ctypes.byref(cp), ctypes.byref(cd), tot // batch, batch, _p(state.view(np.float64)),
                         _p(out.view(np.float64)))
    return out[:tot].copy(), out[tot:2 * tot].real.copy(), (out[2 * tot:].real.copy() if drive.augmented else None)


def rk4_driven(props: Props, drive: Drive, Z, Phi, D, dt: float, steps: int, batch: int = 1):
    """(Z, Phi, D or None) after `steps` RK4 steps of the driven (augmented) system."""
    state = _driven_state(drive, Z, Phi, D)
    tot = _c128(Z).size
    cp, cd = props.c(), drive.c()
    lib().sfo_rk4_driven(ctypes.byref(cp), ctypes.byref(cd), tot // batch, batch, float(dt), int(steps),
                         _p(state.view(np.float64)), None)
    return state[:tot].copy(), state[tot:2 * tot].real.copy(), (state[2 * tot:].real.copy() if drive.augmented else None)


def rk4_timed(props: Props, drive: Drive, Z, Phi, dt: float, t0: float, steps: int, Dsav=None, batch: int = 1):
    """(Z, Phi, Dsav) after `steps` RK4 steps of the time-dependent drive from time t0; Dsav is the
    saved delayed intensity (read only after the first step, whose first stage stores it)."""
    Zc = _c128(Z)
    tot = Zc.size
    state = np.concatenate([Zc, _c128(_f64(Phi))])
    ds = np.zeros(tot) if Dsav is None else _f64(Dsav).copy()
    cp, cd = props.c(), drive.c()
    lib().sfo_rk4_timed(ctypes.byref(cp), ctypes.byref(cd), tot // batch, batch, float(dt), float(t0), int(steps),
                        _p(state.view(np.float64)), _p(ds), None)
    return state[:tot].copy(), state[tot:].real.copy(), ds


def jacobian(props: Props, state, eps: float = 1e-6) -> np.ndarray:
    """Central-difference Jacobian of [vx, vy, dphi] w.r.t. the real state [x, y, phi] (3n), as the
    (3n x 3n) matrix J[r, k] (the reference stores it column-major: column k = perturbed variable)."""
    s = _f64(state).copy()
    m = s.size
    J = np.empty(m * m)
    lib().sfo_jacobian(ctypes.byref(props.c()), m // 3, _p(s), float(eps), _p(J))
    return J.reshape(m, m).T.copy()


def gl2(props: Props, state, t0: float, t1: float, h: float, tol: float = 1e-10, max_newton: int = 20,
        eps: float = 1e-6, max_frames: int = 10000):
    """Gauss-Legendre (2 stages) trajectory of the real state [x, y, phi]: (frames [F, 3n], times [F])."""
    s = _f64(state).copy()
    traj = np.empty((max_frames, s.size))
    times = np.empty(max_frames)
    F = lib().sfo_gl2(ctypes.byref(props.c()), s.size // 3, _p(s), float(t0), float(t1), float(h), float(tol),
                      int(max_newton), float(eps), _p(traj), _p(times), int(max_frames))
    if F < 0:
        raise RuntimeError("Gauss-Legendre step failed to converge")
    return traj[:F].copy(), times[:F].copy()


def gl2_rotation(y0, t1: float, h: float, tol: float = 1e-12) -> np.ndarray:
    y = _f64(y0).copy()
    lib().sfo_gl2_rotation(y.size // 2, _p(y), float(t1), float(h), float(tol))
    return y


def energies(props: Props, Z, Phi, Zp, q) -> np.ndarray:
    """[kinetic, potential, surface, volume flux] of one surface (Energies.cuh:56-204)."""
    Z, Zp, q = _c128(Z), _c128(Zp), _c128(q)
    Phi = _f64(Phi)
    out = np.zeros(4)
    cp = props.c()
    lib().sfo_energies(ctypes.byref(cp), Z.size, _p(Z.view(np.float64)), _p(Phi), _p(Zp.view(np.float64)),
                       _p(q.view(np.float64)), _p(out))
    return out


def adimensionalize(L: float, rho: float, kappa: float, depth: float) -> dict:
    out = np.zeros(5)
    lib().sfo_adimensionalize(L, rho, kappa, depth, _p(out))
    return {"rho": out[0], "kappa": out[1], "depth": out[2], "base_length": out[3],
            "base_time": out[4]}
