"""Linux counterpart of the reference ctypes layer ``SimulationManager``
(CuSuperHelium/Python/integration/rhs.py:68-358).

Same method names, argument meaning and return tuples ``(res, vx, vy, dphi)`` /
``(res, times, states)`` / ``(res, J)``; the library is the in-tree gfx950 build instead of the Windows
DLL.  The option structs go by reference, as the C signatures declare them (the reference passes them
by value, which the Windows x64 calling convention happens to turn into a pointer for structs this
size; the Linux one does not).
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_double, c_size_t

import numpy as np

from ._lib import CGaussLegendreOptions, COptomechanicalVariables, CRK4Options, CSimulationProperties, lib

# the reference's names for the same structs (rhs.py:31-66)
CGaussLegendreProperties = CGaussLegendreOptions
COptomechanicalProperties = COptomechanicalVariables

__all__ = ["SimulationManager", "CSimulationProperties", "CRK4Options", "CGaussLegendreProperties",
           "COptomechanicalProperties"]


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

    def _trajectory(self, fn, free, y0, width, *structs):
        y0 = _f64(y0)
        n = y0.shape[0] // width
        states_ptr, times_ptr = POINTER(c_double)(), POINTER(c_double)()
        states_len, times_len = c_size_t(), c_size_t()
        res = fn(_dp(y0), byref(states_ptr), byref(states_len), byref(times_ptr), byref(times_len),
                 *[byref(s) for s in structs], n)
        if res != 0:
            return res, None, None
        try:
            states = np.array(np.ctypeslib.as_array(states_ptr, shape=(states_len.value, y0.shape[0])), copy=True) \
                if states_len.value else np.empty((0, y0.shape[0]))
            times = (np.array(np.ctypeslib.as_array(times_ptr, shape=(times_len.value,)), copy=True)
                     if times_len.value else np.empty(0))
            return res, times, states
        finally:
            free(states_ptr, times_ptr)

    def integrate(self, y0, sim_props: CSimulationProperties, gl_props: CGaussLegendreOptions):
        """integrateSimulationGL2 (rhs.py:224-262): returns (res, times, states[frames, 3N])."""
        return self._trajectory(self.lib.integrateSimulationGL2, self.lib.integrateSimulationGL2_freeMemory, y0, 3,
                                sim_props, gl_props)

    def integrate_optomechanical_problem(self, y0, sim_props: CSimulationProperties,
                                         opt_props: COptomechanicalVariables, rk4_props: CRK4Options):
        """integrateOptomechanicalSimulationRK4 (rhs.py:264-304): returns (res, times, states[steps, 3N])."""
        return self._trajectory(self.lib.integrateOptomechanicalSimulationRK4,
                                self.lib.integrateOptomechanicalSimulationRK4_freeMemory, y0, 3, sim_props, rk4_props,
                                opt_props)

    def integrate_augmented_optomechanical_problem(self, y0, sim_props: CSimulationProperties,
                                                   opt_props: COptomechanicalVariables, rk4_props: CRK4Options):
        """integrateAugmentedOptomechanicalSimulationRK4 (rhs.py:306-346): (res, times, states[steps, 4N])."""
        return self._trajectory(self.lib.integrateAugmentedOptomechanicalSimulationRK4,
                                self.lib.integrateAugmentedOptomechanicalSimulationRK4_freeMemory, y0, 4, sim_props,
                                rk4_props, opt_props)

    def calculate_augmented_rhs(self, y0, sim_props: CSimulationProperties, opt_props: COptomechanicalVariables):
        """calculateRhsAugmentedOptomechanical (rhs.py:348-358): the 4N right-hand side; raises on error."""
        y0 = _f64(y0)
        rhs = np.empty_like(y0)
        res = self.lib.calculateRhsAugmentedOptomechanical(_dp(y0), _dp(rhs), byref(sim_props), byref(opt_props),
                                                           y0.shape[0] // 4)
        if res != 0:
            raise Exception("Error in calculating augmented RHS")
        return rhs

    def calculate_perturbed_states256(self, x, y, phi, L: float, rho: float, kappa: float, depth: float,
                                      epsilon: float = 1e-6):
        """calculatePerturbedStates256 (rhs.py:197-211): returns (res, Zperturbed[6 n^2] complex)."""
        x, y, phi = _f64(x), _f64(y), _f64(phi)
        n = x.shape[0]
        if y.shape[0] != n or phi.shape[0] != n:
            raise ValueError("Input arrays must have the same length")
        Zp = np.empty(6 * n * n, dtype=np.complex128)
        return self.lib.calculatePerturbedStates256(_dp(x), _dp(y), _dp(phi), Zp.ctypes.data, L, rho, kappa, depth,
                                                    epsilon), Zp

    def calculate_jacobian(self, x, y, phi, L: float, rho: float, kappa: float, depth: float, epsilon: float = 1e-6):
        """calculateJacobian (rhs.py:213-222): returns (res, J) with J the (3n, 3n) Fortran-ordered Jacobian."""
        x, y, phi = _f64(x), _f64(y), _f64(phi)
        n = x.shape[0]
        if y.shape[0] != n or phi.shape[0] != n:
            raise ValueError("Input arrays must have the same length")
        J = np.ones((3 * n, 3 * n), dtype=np.float64, order="F") * 100
        s = np.ascontiguousarray(np.concatenate((x, y, phi)))
        res = self.lib.calculateJacobian(_dp(s), J.ctypes.data_as(POINTER(c_double)), L, rho, kappa, depth, epsilon, n)
        return res, J
