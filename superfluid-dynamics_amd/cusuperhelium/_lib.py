"""Loader for ``lib/libcusuperhelium.so`` (gfx950 kernels behind the C ABI of
``include/cusuperhelium.h``).

The shared library is built in-tree by ``make -C superfluid-dynamics_amd``; there is no
fallback implementation: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_bool, c_double, c_int, c_longlong, c_size_t, c_void_p

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libcusuperhelium.so")
_lib = None


class CSimulationProperties(Structure):
    """SimProperties (ExportTypes.cuh:8-20), same field order as the reference's ctypes mirror
    (CuSuperHelium/Python/integration/rhs.py:25-33)."""

    _fields_ = [("L", c_double), ("rho", c_double), ("kappa", c_double), ("depth", c_double),
                ("use_expansions", c_bool), ("expansion_order", c_int), ("infinite_depth", c_bool)]


class CRK4Options(Structure):
    """RK4SolverOptions (ExportTypes.cuh:35-40), mirror of rhs.py:48-53."""

    _fields_ = [("timeStep", c_double), ("t0", c_double), ("t1", c_double), ("returnTrajectory", c_bool)]


class SfhProps(Structure):
    _fields_ = [("kind", c_int), ("infinite_depth", c_int), ("use_expansions", c_int),
                ("expansion_order", c_int), ("rho_quirk", c_int), ("reserved", c_int),
                ("rho", c_double), ("kappa", c_double), ("depth", c_double), ("U", c_double),
                ("g", c_double)]


class SfhOptions(Structure):
    _fields_ = [("precision", c_int), ("solver", c_int), ("use_graph", c_int), ("max_iter", c_int),
                ("restart", c_int), ("force_trim", c_int), ("tol", c_double), ("pair_rows", c_int),
                ("reserved", c_int)]


class SfhTuning(Structure):
    _fields_ = [("solver", c_int), ("precision", c_int), ("pair_rows", c_int), ("from_cache", c_int),
                ("ms_per_step", c_double)]


class SfhDrive(Structure):
    _fields_ = [("augmented", c_int), ("timed", c_int)] + [
        (k, c_double) for k in ("damping", "strength", "detuning", "gamma", "G", "max_intensity", "x0", "sigma",
                                "beta", "tau")]


class COptomechanicalVariables(Structure):
    """COptomechanicalVariables (ExportTypes.cuh:41-64), SI units."""

    _fields_ = [(k, c_double) for k in ("detuning", "gamma", "G", "tau", "max_intensity", "initial_time",
                                        "location_x0_mode", "sigma_optical_mode", "beta", "damping_strength")]


class CGaussLegendreOptions(Structure):
    """GaussLegendreOptions (ExportTypes.cuh:21-34)."""

    _fields_ = [("t0", c_double), ("t1", c_double), ("stepSize", c_double), ("newtonTolerance", c_double),
                ("maxNewtonIterations", c_size_t), ("allowSimplifiedFallback", c_bool),
                ("returnTrajectory", c_bool), ("armijo_c", c_double), ("backtrack", c_double),
                ("minAlpha", c_double), ("maxStepsHalves", c_size_t)]


class SfhStats(Structure):
    _fields_ = [("rhs_evals", c_longlong), ("steps", c_longlong), ("fallback_solves", c_int),
                ("spin_timeouts", c_int), ("last_iterations", c_int), ("max_iterations", c_int),
                ("unconverged", c_int), ("refinements", c_int), ("redone_steps", c_int),
                ("sequence_length", c_int), ("solver", c_int), ("precision", c_int), ("pair_rows", c_int),
                ("reserved", c_int)]


DP = POINTER(c_double)

# name -> (restype, argtypes); every symbol declared in include/cusuperhelium.h
SIGNATURES = {
    "calculateRHSFromVectors": (c_int, [DP, DP, DP, DP, DP, DP, c_double, c_double, c_double, c_double, c_size_t]),
    "calculateRHS256FromVectors": (c_int, [DP, DP, DP, DP, DP, DP, c_double, c_double, c_double, c_double]),
    "calculateRHS2048FromVectors": (c_int, [DP, DP, DP, DP, DP, DP, c_double, c_double, c_double, c_double]),
    "calculateRHS256FromVectorsBatched": (c_int, [DP, DP, DP, DP, DP, DP, c_double, c_double, c_double,
                                                  c_double, c_int]),
    "calculateVorticities256FromVectors": (c_int, [c_void_p, c_void_p, DP, c_void_p, c_void_p, c_double,
                                                   c_double, c_double, c_double]),
    "calculateDerivativeFFT256": (c_int, [c_void_p, c_void_p]),
    "integrateSimulationRK4": (c_int, [DP, POINTER(DP), POINTER(c_size_t), POINTER(DP), POINTER(c_size_t),
                                       POINTER(CSimulationProperties), POINTER(CRK4Options), c_size_t]),
    "integrateSimulationRK4_freeMemory": (c_int, [DP, DP]),
    "integrateAugmentedOptomechanicalSimulationRK4": (c_int, [DP, POINTER(DP), POINTER(c_size_t), POINTER(DP),
                                                              POINTER(c_size_t), POINTER(CSimulationProperties),
                                                              POINTER(CRK4Options), POINTER(COptomechanicalVariables),
                                                              c_size_t]),
    "integrateOptomechanicalSimulationRK4": (c_int, [DP, POINTER(DP), POINTER(c_size_t), POINTER(DP),
                                                     POINTER(c_size_t), POINTER(CSimulationProperties),
                                                     POINTER(CRK4Options), POINTER(COptomechanicalVariables),
                                                     c_size_t]),
    "integrateOptomechanicalSimulationRK4_freeMemory": (c_int, [DP, DP]),
    "calculateRhsAugmentedOptomechanical": (c_int, [DP, DP, POINTER(CSimulationProperties),
                                                    POINTER(COptomechanicalVariables), c_size_t]),
    "calculatePerturbedStates256": (c_int, [DP, DP, DP, c_void_p, c_double, c_double, c_double, c_double,
                                            c_double]),
    "integrateAugmentedOptomechanicalSimulationRK4_freeMemory": (c_int, [DP, DP]),
    "calculateJacobian": (c_int, [DP, DP, c_double, c_double, c_double, c_double, c_double, c_size_t]),
    "integrateSimulationGL2": (c_int, [DP, POINTER(DP), POINTER(c_size_t), POINTER(DP), POINTER(c_size_t),
                                       POINTER(CSimulationProperties), POINTER(CGaussLegendreOptions), c_size_t]),
    "integrateSimulationGL2_freeMemory": (c_int, [DP, DP]),
    "sfh_jacobian": (c_int, [POINTER(SfhProps), c_int, DP, c_double, DP]),
    "sfh_autotune": (c_int, [POINTER(SfhProps), c_int, c_int, c_int, POINTER(SfhTuning)]),
    "sfh_tuned": (c_int, [POINTER(SfhProps), c_int, c_int, POINTER(SfhTuning)]),
    "sfh_set_drive": (c_int, [c_void_p, POINTER(SfhDrive)]),
    "sfh_set_time": (c_int, [c_void_p, c_double]),
    "sfh_set_delayed": (c_int, [c_void_p, DP]),
    "sfh_get_delayed": (c_int, [c_void_p, DP]),
    "sfh_get_delayed_rhs": (c_int, [c_void_p, DP]),
    "sfh_default_options": (None, [POINTER(SfhOptions)]),
    "sfh_default_props": (None, [POINTER(SfhProps)]),
    "sfh_create": (c_int, [POINTER(SfhProps), POINTER(c_int), c_int, POINTER(SfhOptions), POINTER(c_void_p)]),
    "sfh_destroy": (c_int, [c_void_p]),
    "sfh_total_nodes": (c_int, [c_void_p]),
    "sfh_set_state": (c_int, [c_void_p, DP, DP, DP]),
    "sfh_get_state": (c_int, [c_void_p, DP, DP, DP]),
    "sfh_rhs": (c_int, [c_void_p, DP, DP, DP, DP, DP, DP, DP]),
    "sfh_step": (c_int, [c_void_p, c_double, c_int]),
    "sfh_step_record": (c_int, [c_void_p, c_double, c_int, c_int, DP, DP]),
    "sfh_set_state_device": (c_int, [c_void_p, c_void_p]),
    "sfh_get_state_device": (c_int, [c_void_p, c_void_p]),
    "sfh_rhs_device": (c_int, [c_void_p, c_void_p, c_void_p]),
    "sfh_rhs_vjp": (c_int, [c_void_p, DP, DP, DP, DP, DP, DP]),
    "sfh_step_vjp": (c_int, [c_void_p, c_double, DP, DP, DP, DP, DP, DP]),
    "sfh_rhs_vjp_augmented": (c_int, [c_void_p, DP, DP, DP, DP, DP, DP, DP, DP]),
    "sfh_step_vjp_augmented": (c_int, [c_void_p, c_double, DP, DP, DP, DP, DP, DP, DP, DP]),
    "sfh_traj_forward": (c_int, [c_void_p, c_double, c_int, c_int, POINTER(c_int)]),
    "sfh_traj_backward": (c_int, [c_void_p, c_int, DP, DP, DP, DP, DP, DP]),
    "sfh_traj_release": (c_int, [c_void_p, c_int]),
    "sfh_get_mu": (c_int, [c_void_p, DP]),
    "sfh_track_energies": (c_int, [c_void_p, c_int]),
    "sfh_get_energies": (c_int, [c_void_p, DP]),
    "sfh_sync": (c_int, [c_void_p]),
    "sfh_get_stats": (c_int, [c_void_p, POINTER(SfhStats)]),
    "sfh_device_state": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p)]),
    "sfh_last_error": (ctypes.c_char_p, []),
    "sfh_rhs_noalloc": (c_int, [POINTER(SfhProps), c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_double, c_int, c_void_p]),
    "sfh_rhs_noalloc_scratch_bytes": (c_size_t, []),
    "sfh_debug_lu_solve": (c_int, [c_int, DP, DP, DP, POINTER(c_int)]),
    "sfh_kernel_timing": (c_int, [c_void_p, c_int]),
    "sfh_kernel_report": (c_int, [c_void_p, ctypes.c_char_p, POINTER(c_int), DP, DP, DP]),
}


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Returns the loaded library, raising if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing(f"{LIB_PATH} not built: run `make -C superfluid-dynamics_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    msg = lib().sfh_last_error()
    return msg.decode() if msg else ""


class SolverError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"{where} failed with code {code}: {last_error()}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != 0:
        raise SolverError(code, where)
