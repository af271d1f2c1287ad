"""The optically driven helium film (HeliumDrivenAutonomousProblem) and its augmented form with the
delayed intensity D as a third state field (AugmentedBoundaryIntegrator), on the GPU against the
CPU FP64 oracle's sfo_rhs_driven / sfo_rk4_driven, through the C ABI.

Tolerances: right-hand sides 1e-12 relative to 1 + max|reference|, 20-step trajectories 1e-11.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DRIVE = dict(damping=0.05, strength=0.3, detuning=0.4, gamma=0.8, G=2.0, max_intensity=0.6, x0=np.pi,
             sigma=0.7, beta=0.9, tau=1.5)


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


def _surface(n, seed, amp=0.03):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, 9))
    th = rng.uniform(0, 2 * np.pi, 8)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(8))


def _close(a, b, tol):
    err = np.abs(np.asarray(a) - np.asarray(b)).max()
    assert err <= tol * (1.0 + np.abs(b).max()), f"max err {err:.3e}"
    return err


@pytest.mark.parametrize("augmented", [False, True])
@pytest.mark.parametrize("solver", ["lu", "gmres", "batched"])
def test_driven_rhs_matches_oracle(oracle, csh, solver, augmented):
    n = 256
    Z = _surface(n, 7)
    x = 2 * np.pi * np.arange(n) / n
    P = 0.01 * np.sin(x + 0.3)
    D = 0.02 * np.cos(2 * x) if augmented else None
    kw = dict(kind=1, depth=1.0)
    sv = {"lu": csh.LU, "gmres": csh.GMRES, "batched": csh.BATCHED}[solver]
    with csh.Ensemble([n], csh.Props(**kw), solver=sv, tol=1e-13) as e:
        e.set_drive(augmented=augmented, **DRIVE)
        e.set_state(Z, P)
        if augmented:
            e.set_delayed(D)
        q, dp = e.rhs()
        dD = e.delayed_rhs() if augmented else None
    qo, dpo, dDo = oracle.rhs_driven(oracle.Props(**kw), oracle.Drive(augmented=augmented, **DRIVE), Z, P, D)
    eq, ep = _close(q, qo, 1e-12), _close(dp, dpo, 1e-12)
    ed = _close(dD, dDo, 1e-12) if augmented else 0.0
    print(f"driven rhs {solver} augmented={augmented}: |dq| {eq:.2e} |ddphi| {ep:.2e} |ddD| {ed:.2e}", flush=True)


@pytest.mark.parametrize("solver,force_trim", [("lu", 0), ("gmres", 0), ("gmres", 2), ("batched", 0)])
def test_augmented_rk4_matches_oracle(oracle, csh, solver, force_trim):
    # 20 steps of the augmented film; force_trim = 2 makes the trimmed GMRES chunks run out of launches,
    # so the chunk checkpoint (which holds D as well) is restored and the chunk run again
    n, dt, steps = 256, 0.005, 20
    Z = _surface(n, 8)
    x = 2 * np.pi * np.arange(n) / n
    P = 0.005 * np.sin(2 * x)
    D = 0.01 * np.cos(x)
    kw = dict(kind=1, depth=1.0)
    sv = {"lu": csh.LU, "gmres": csh.GMRES, "batched": csh.BATCHED}[solver]
    with csh.Ensemble([n], csh.Props(**kw), solver=sv, tol=1e-13, force_trim=force_trim) as e:
        e.set_drive(augmented=True, **DRIVE)
        e.set_state(Z, P)
        e.set_delayed(D)
        for k in (3, 17):
            e.step(dt, k)
        Zg, Pg = e.get_state()
        Dg = e.delayed()
        st = e.stats()
    Zo, Po, Do = oracle.rk4_driven(oracle.Props(**kw), oracle.Drive(augmented=True, **DRIVE), Z, P, D, dt, steps)
    ez, ep, ed = _close(Zg, Zo, 1e-11), _close(Pg, Po, 1e-11), _close(Dg, Do, 1e-11)
    print(f"augmented rk4 {solver} trim={force_trim}: |dz| {ez:.2e} |dphi| {ep:.2e} |dD| {ed:.2e} redone "
          f"{st['redone_steps']}", flush=True)
    assert st["unconverged"] == 0
    if force_trim:
        assert st["redone_steps"] > 0


def test_augmented_step_record_frames_hold_D(csh):
    n = 128
    Z = _surface(n, 9)
    with csh.Ensemble([n], csh.Props(kind=1, depth=1.0), solver=csh.LU) as e:
        e.set_drive(augmented=True, **DRIVE)
        e.set_state(Z, np.zeros(n))
        Zf, Pf, tf, Df = e.step_record(0.005, 6, 2)
        D = e.delayed()
        Zl, Pl = e.get_state()
    assert Df.shape == (3, n)
    assert np.array_equal(Df[-1], D) and np.array_equal(Zf[-1], Zl) and np.array_equal(Pf[-1], Pl)
    assert np.abs(Df[0]).max() > 0


def test_dropin_augmented_optomechanical_integration(oracle, csh):
    # integrateAugmentedOptomechanicalSimulationRK4 (Export.cu:981-1093) in SI units against the
    # oracle run on the nondimensional problem the reference derives (adimensionalizeProperties,
    # adimensionalizeOptomechanicalVariables, get_current_intensity_drive_strength)
    lib = csh.lib()
    N, L, depth, rho = 64, 1e-3, 5e-8, 145.0
    sp = csh.CSimulationProperties(L, rho, 0.0, depth, False, 1, False)
    alpha, rho_he, hbar = 3.5e-24, 150.0, 1.054571817e-34
    L0 = L / (2 * np.pi)
    t0 = np.sqrt(L0 / (3 * alpha / depth ** 4))
    E0 = 3 * rho_he * alpha * L0 ** 4 / depth ** 4
    opt = csh.COptomechanicalVariables(detuning=2e3, gamma=5e3, G=1e11, tau=2e-3, max_intensity=1e18,
                                       initial_time=0.0, location_x0_mode=L / 2, sigma_optical_mode=L / 8,
                                       beta=2e3, damping_strength=0.02)
    dt_si = 0.002 * t0
    rk = csh.CRK4Options(dt_si, 0.0, 10 * dt_si + 1e-6 * dt_si, True)
    x = 2 * np.pi * np.arange(N) / N
    y = 1e-4 * np.cos(x)       # the film is h = depth / L0 = 3.1e-4 thick
    phi = np.zeros(N)
    D0 = 0.001 * np.sin(x)
    init = np.concatenate([x, y, phi, D0])
    states, times = ctypes.POINTER(ctypes.c_double)(), ctypes.POINTER(ctypes.c_double)()
    ns, nt = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib.integrateAugmentedOptomechanicalSimulationRK4(
        init.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(states), ctypes.byref(ns),
        ctypes.byref(times), ctypes.byref(nt), ctypes.byref(sp), ctypes.byref(rk), ctypes.byref(opt), N)
    assert rc == 0, csh._lib.last_error()
    assert ns.value == 10 and nt.value == 10
    out = np.ctypeslib.as_array(states, shape=(10, 4 * N)).copy()
    tt = np.ctypeslib.as_array(times, shape=(10,)).copy()
    lib.integrateAugmentedOptomechanicalSimulationRK4_freeMemory(states, times)
    rho_a = rho / rho_he
    G = 1e11 * t0 * L0
    tau = 2e-3 / t0
    sig = (L / 8) / L0
    drive = oracle.Drive(augmented=True, damping=0.02, strength=hbar / (E0 * t0 * rho_a) * G / sig ** 2,
                         detuning=2e3 * t0, gamma=5e3 * t0, G=G, max_intensity=1e18, x0=(L / 2) / L0, sigma=sig,
                         beta=2e3 * (hbar / E0 / t0) * G / tau / (sig ** 2 * rho_a), tau=tau)
    props = oracle.Props(kind=oracle.HELIUM, depth=depth / L0, rho=rho_a, rho_quirk=True)
    Zo, Po, Do = oracle.rk4_driven(props, drive, x + 1j * y, phi, D0, dt_si / t0, 10)
    print(f"drop-in augmented: drive strength {drive.strength:.3e} beta {drive.beta:.3e}", flush=True)
    # the converted drive moves the surface: without it the trajectory differs well above the bound
    Zn, _, _ = oracle.rk4_driven(props, oracle.Drive(augmented=True, damping=0.02, tau=tau), x + 1j * y, phi,
                                 D0, dt_si / t0, 10)
    assert np.abs(Zn - Zo).max() > 1e-8
    _close(out[-1, :N] + 1j * out[-1, N:2 * N], Zo, 1e-11)
    _close(out[-1, 2 * N:3 * N], Po, 1e-11)
    assert np.array_equal(out[:, 3 * N:], out[:, 2 * N:3 * N])   # the reference's fourth block is phi
    assert np.allclose(tt, (np.arange(10) + 1) * (dt_si / t0), rtol=1e-13, atol=0)
