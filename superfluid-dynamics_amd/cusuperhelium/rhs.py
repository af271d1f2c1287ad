"""Linux counterpart of the reference ctypes layer ``SimulationManager``
(CuSuperHelium/Python/integration/rhs.py:68-358) for the RK4 / RHS path.

Same method names, argument meaning and return tuples ``(res, vx, vy, dphi)`` /
``(res, times, states)``; the library is the in-tree gfx950 build instead of the Windows DLL.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_size_t

import numpy as np

from ._lib import CRK4Options, CSimulationProperties, lib

__all__ = ["SimulationManager", "CSimulationProperties", "CRK4Options"]


def _dp(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class SimulationManager:
    def __init__(self, dll_path: str | None = None):
        # dll_path is accepted for signature compatibility; the in-tree build is always used.
        self.lib = lib()

    def _rhs_call(self, fn, x, y, phi, *args):
        x, y, phi = _f64(x), _f64(y), _f64(phi)
        if y.shape[0] != x.shape[0] or phi.shape[0] != x.shape[0]:
            raise ValueError("Input arrays must have the same length")
        vx, vy, dphi = np.empty_like(x), np.empty_like(x), np.empty_like(x)
        res = fn(_dp(x), _dp(y), _dp(phi), _dp(vx), _dp(vy), _dp(dphi), *args)
        return res, vx, vy, dphi

    def calculate_rhs256_from_vectors(self, x, y, phi, L: float, rho: float, kappa: float, depth: float):
        return self._rhs_call(self.lib.calculateRHS256FromVectors, x, y, phi, L, rho, kappa, depth)

    def calculate_rhs_from_vectors(self, x, y, phi, L: float, rho: float, kappa: float, depth: float):
        return self._rhs_call(self.lib.calculateRHSFromVectors, x, y, phi, L, rho, kappa, depth,
                              np.asarray(x).shape[0])

    def calculate_rhs2048_from_vectors(self, x, y, phi, L: float, rho: float, kappa: float, depth: float):
        return self._rhs_call(self.lib.calculateRHS2048FromVectors, x, y, phi, L, rho, kappa, depth)

    def calculate_rhs256_from_vectors_batched(self, x, y, phi, L: float, rho: float, kappa: float,
                                              depth: float, batch_size: int):
        if np.asarray(phi).shape[0] != 256 * batch_size:
            raise ValueError("Input arrays must have the same length equal to 256 * batch_size")
        return self._rhs_call(self.lib.calculateRHS256FromVectorsBatched, x, y, phi, L, rho, kappa, depth,
                              batch_size)

    def calculate_vorticities256_from_vectors(self, Z, phi, L: float, rho: float, kappa: float, depth: float):
        Z = np.ascontiguousarray(Z, dtype=np.complex128)
        phi = np.ascontiguousarray(phi, dtype=np.complex128)
        if phi.shape[0] != Z.shape[0]:
            raise ValueError("Input arrays must have the same length")
        n = Z.shape[0]
        vort = np.empty(n, dtype=np.float64)
        Zp = np.empty(n, dtype=np.complex128)
        Zpp = np.empty(n, dtype=np.complex128)
        res = self.lib.calculateVorticities256FromVectors(Z.ctypes.data, phi.ctypes.data, _dp(vort),
                                                          Zp.ctypes.data, Zpp.ctypes.data, L, rho, kappa,
                                                          depth)
        return res, vort, Zp, Zpp

    def calculate_derivativeFFT256(self, input) -> tuple[int, np.ndarray]:
        inp = np.ascontiguousarray(input, dtype=np.complex128)
        out = np.empty(inp.shape[0], dtype=np.complex128)
        return self.lib.calculateDerivativeFFT256(inp.ctypes.data, out.ctypes.data), out

    def integrate_rk4(self, y0, sim_props: CSimulationProperties, rk4_props: CRK4Options):
        """integrateSimulationRK4: returns (res, times, states[steps, 3N])."""
        y0 = _f64(y0)
        n = y0.shape[0] // 3
        states_ptr, times_ptr = POINTER(c_double)(), POINTER(c_double)()
        states_len, times_len = c_size_t(), c_size_t()
        res = self.lib.integrateSimulationRK4(_dp(y0), byref(states_ptr), byref(states_len),
                                              byref(times_ptr), byref(times_len), byref(sim_props),
                                              byref(rk4_props), n)
        if res != 0:
            return res, None, None
        try:
            states = np.array(np.ctypeslib.as_array(states_ptr, shape=(states_len.value, 3 * n)), copy=True)
            times = (np.array(np.ctypeslib.as_array(times_ptr, shape=(times_len.value,)), copy=True)
                     if times_len.value else np.empty(0))
            return res, times, states
        finally:
            self.lib.integrateSimulationRK4_freeMemory(states_ptr, times_ptr)
