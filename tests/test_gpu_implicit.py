"""Central-difference Jacobian (JacobianCalculator, JacobianCalculator.cuh:223-284) and two-stage
Gauss-Legendre integration (GaussLegendre2, GaussLegendre.cuh:190-560) on the GPU against the CPU
FP64 oracle's sfo_jacobian / sfo_gl2, through the C ABI.

Bars: Jacobian entries to 1e-7 of max|J| (a central difference with eps = 1e-6 carries the rounding
of the two right-hand sides divided by 2 eps, ~1e-10 relative, and the GPU's LU and the oracle's GEPP
round differently); Gauss-Legendre states to 1e-10 (both Newton iterations stop at tol 1e-12).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


def _state(n, amp=0.03, seed=2):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    y = sum(rng.uniform(0, amp / k) * np.cos(k * x + rng.uniform(0, 2 * np.pi)) for k in range(1, 6))
    phi = 0.01 * np.sin(x + 0.4)
    return np.concatenate([x, y, phi])


@pytest.mark.parametrize("n,kw", [(32, {}), (64, dict(kind=1, depth=1.0))])
def test_sfh_jacobian_matches_oracle(oracle, csh, n, kw):
    s = _state(n)
    m = 3 * n
    jac = np.empty(m * m)
    cp = csh.Props(**kw).c()
    rc = csh.lib().sfh_jacobian(ctypes.byref(cp), n, s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1e-6,
                                jac.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, csh._lib.last_error()
    J = jac.reshape(m, m).T          # column-major: J[r, k] = d f_r / d s_k
    Jo = oracle.jacobian(oracle.Props(**kw), s)
    err = np.abs(J - Jo).max() / np.abs(Jo).max()
    print(f"jacobian n={n} {kw}: max |J| {np.abs(Jo).max():.3e}, rel err {err:.2e}", flush=True)
    assert err < 1e-7


def _si_props(csh):
    # SI inputs whose nondimensional film depth is 1: L0 = L / 2 pi = depth
    L, depth = 2 * np.pi * 1e-7, 1e-7
    return csh.CSimulationProperties(L, 150.0, 0.0, depth, False, 1, False)


def test_dropin_calculate_jacobian(oracle, csh):
    n = 32
    s = _state(n)
    m = 3 * n
    jac = np.empty(m * m)
    sp = _si_props(csh)
    rc = csh.lib().calculateJacobian(s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     jac.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     sp.L, sp.rho, sp.kappa, sp.depth, 1e-6, n)
    assert rc == 0, csh._lib.last_error()
    d = oracle.adimensionalize(sp.L, sp.rho, sp.kappa, sp.depth)
    Jo = oracle.jacobian(oracle.Props(kind=oracle.HELIUM, depth=d["depth"], rho=d["rho"]), s)
    err = np.abs(jac.reshape(m, m).T - Jo).max() / np.abs(Jo).max()
    print(f"drop-in jacobian: depth {d['depth']:.6f} rel err {err:.2e}", flush=True)
    assert err < 1e-7
    assert csh.lib().calculateJacobian(s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       jac.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       sp.L, sp.rho, sp.kappa, sp.depth, 1e-6, 48) == -2


def test_dropin_gauss_legendre(csh):
    # against the oracle's Gauss-Legendre frames of the same film (tests/golden/gl2_helium_n32.npz)
    import os
    with np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gl2_helium_n32.npz")) as f:
        g = {k: f[k] for k in f.files}
    n, h, t1 = 32, 0.05, 0.2
    s0 = g["s0"]
    sp = _si_props(csh)
    go = csh.CGaussLegendreOptions(0.0, t1, h, 1e-12, 20, False, True, 1e-4, 0.5, 1e-6, 6)
    states, times = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    ns, nt = ctypes.c_size_t(), ctypes.c_size_t()
    rc = csh.lib().integrateSimulationGL2(s0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(states),
                                          ctypes.byref(ns), ctypes.byref(times), ctypes.byref(nt), ctypes.byref(sp),
                                          ctypes.byref(go), n)
    assert rc == 0, csh._lib.last_error()
    F = ns.value
    out = np.ctypeslib.as_array(states, shape=(F, 3 * n)).copy()
    tt = np.ctypeslib.as_array(times, shape=(F,)).copy()
    csh.lib().integrateSimulationGL2_freeMemory(states, times)
    print(f"gauss-legendre: frames {F} (oracle {len(g['times'])}), |ds| {np.abs(out - g['frames']).max():.2e}",
          flush=True)
    assert F == len(g["times"]) == 5
    assert np.array_equal(out[0], s0)
    assert np.abs(tt - g["times"]).max() < 1e-15
    assert np.abs(out - g["frames"]).max() < 1e-10
    # without a trajectory the reference returns nothing (GaussLegendre.cuh copyStatesToHost)
    go.returnTrajectory = False
    rc = csh.lib().integrateSimulationGL2(s0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(states),
                                          ctypes.byref(ns), ctypes.byref(times), ctypes.byref(nt), ctypes.byref(sp),
                                          ctypes.byref(go), n)
    assert rc == 0 and ns.value == 0 and nt.value == 0
