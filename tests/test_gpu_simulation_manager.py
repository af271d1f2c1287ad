"""The Linux SimulationManager (cusuperhelium.rhs, mirroring CuSuperHelium/Python/integration/rhs.py)
for the entry points beyond the RHS/RK4 path: calculate_jacobian, integrate (Gauss-Legendre),
integrate_optomechanical_problem, integrate_augmented_optomechanical_problem, calculate_augmented_rhs,
calculate_perturbed_states256.  Each is compared with the same entry point called through the C ABI
directly (the C-ABI tests compare those with the oracle), so that the method names, argument order and
return tuples are the reference's and the results the library's.
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


DP = ctypes.POINTER(ctypes.c_double)


def _opto(csh):
    return csh.rhs.COptomechanicalProperties(detuning=2e3, gamma=5e3, G=1e11, tau=2e-3, max_intensity=1e18,
                                             initial_time=0.0, location_x0_mode=0.5e-3, sigma_optical_mode=1.25e-4,
                                             beta=2e3, damping_strength=0.02)


def test_manager_jacobian_and_gauss_legendre(csh):
    sm = csh.SimulationManager()
    n = 32
    x = 2 * np.pi * np.arange(n) / n
    y, phi = 0.02 * np.cos(x), 0.01 * np.sin(x)
    L, depth = 2 * np.pi * 1e-7, 1e-7
    res, J = sm.calculate_jacobian(x, y, phi, L, 150.0, 0.0, depth, 1e-6)
    assert res == 0 and J.shape == (3 * n, 3 * n) and J.flags.f_contiguous
    ref = np.empty(9 * n * n)
    s = np.concatenate([x, y, phi])
    assert csh.lib().calculateJacobian(s.ctypes.data_as(DP), ref.ctypes.data_as(DP), L, 150.0, 0.0, depth, 1e-6, n) == 0
    assert np.array_equal(J.ravel(order="F"), ref)
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gl2_helium_n32.npz")) as f:
        g = {k: f[k] for k in f.files}
    sp = csh.CSimulationProperties(L, 150.0, 0.0, depth, False, 1, False)
    go = csh.rhs.CGaussLegendreProperties(0.0, 0.2, 0.05, 1e-12, 20, False, True, 1e-4, 0.5, 1e-6, 6)
    res, times, states = sm.integrate(g["s0"], sp, go)
    assert res == 0 and states.shape == g["frames"].shape
    assert np.abs(states - g["frames"]).max() < 1e-10 and np.abs(times - g["times"]).max() < 1e-15


def test_manager_optomechanical_entry_points(csh):
    sm = csh.SimulationManager()
    N, L, depth, rho = 64, 1e-3, 5e-8, 145.0
    sp = csh.CSimulationProperties(L, rho, 0.0, depth, False, 1, False)
    ov = _opto(csh)
    tb = np.sqrt((L / (2 * np.pi)) / (3 * 3.5e-24 / depth ** 4))
    rk = csh.CRK4Options(0.002 * tb, 0.0, 0.002 * tb * 6.000001, True)
    x = 2 * np.pi * np.arange(N) / N
    y0 = np.concatenate([x, 1e-4 * np.cos(x), 1e-6 * np.sin(2 * x)])
    res, times, states = sm.integrate_optomechanical_problem(y0, sp, ov, rk)
    assert res == 0 and states.shape == (6, 3 * N) and times.shape == (6,)
    sd, td = DP(), DP()
    ns, nt = ctypes.c_size_t(), ctypes.c_size_t()
    y = y0.copy()
    assert csh.lib().integrateOptomechanicalSimulationRK4(y.ctypes.data_as(DP), ctypes.byref(sd), ctypes.byref(ns),
                                                          ctypes.byref(td), ctypes.byref(nt), ctypes.byref(sp),
                                                          ctypes.byref(rk), ctypes.byref(ov), N) == 0
    direct = np.ctypeslib.as_array(sd, shape=(ns.value, 3 * N)).copy()
    csh.lib().integrateOptomechanicalSimulationRK4_freeMemory(sd, td)
    assert np.array_equal(states, direct)
    ya = np.concatenate([y0, 1e-3 * np.cos(x)])
    res, times, states = sm.integrate_augmented_optomechanical_problem(ya, sp, ov, rk)
    assert res == 0 and states.shape == (6, 4 * N)
    rhs = sm.calculate_augmented_rhs(ya, sp, ov)
    ref = np.empty(4 * N)
    yy = ya.copy()
    assert csh.lib().calculateRhsAugmentedOptomechanical(yy.ctypes.data_as(DP), ref.ctypes.data_as(DP),
                                                         ctypes.byref(sp), ctypes.byref(ov), N) == 0
    assert np.array_equal(rhs, ref)
    with pytest.raises(Exception, match="augmented RHS"):
        sm.calculate_augmented_rhs(ya[:4 * 48], sp, ov)


def test_manager_perturbed_states(csh):
    sm = csh.SimulationManager()
    n = 256
    x = 2 * np.pi * np.arange(n) / n
    res, Zp = sm.calculate_perturbed_states256(x, 0.01 * np.cos(x), np.zeros(n), 1e-6, 0.0, 0.0, 1e-8, 1e-5)
    assert res == 0 and Zp.shape == (6 * n * n,)
    assert Zp[5 * n + 5] == x[5] + 1e-5 + 0.01j * np.cos(x[5])
    assert Zp[3 * n * n + (2 * n + 7) * n + 7] == 1e-5
