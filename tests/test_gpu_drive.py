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


@pytest.mark.parametrize("solver,force_trim", [("lu", 0), ("gmres", 0), ("gmres", 2), ("batched", 0)])
def test_timed_drive_rk4_matches_oracle(oracle, csh, solver, force_trim):
    # the time-dependent delayed intensity (RK4_Time_Dependent.cuh:170-412): 20 steps from t0 = 0.7,
    # in two calls; force_trim = 2 makes trimmed GMRES chunks run out of launches, so a chunk is
    # restored (state, saved D and the drive's clock) and run again
    n, dt, t0 = 256, 0.005, 0.7
    Z = _surface(n, 12)
    x = 2 * np.pi * np.arange(n) / n
    P = 0.004 * np.cos(3 * x)
    kw = dict(kind=1, depth=1.0)
    sv = {"lu": csh.LU, "gmres": csh.GMRES, "batched": csh.BATCHED}[solver]
    with csh.Ensemble([n], csh.Props(**kw), solver=sv, tol=1e-13, force_trim=force_trim) as e:
        e.set_drive(timed=True, **DRIVE)
        e.set_state(Z, P)
        e.set_time(t0)
        for k in (7, 13):
            e.step(dt, k)
        Zg, Pg = e.get_state()
        Dg = e.delayed()
        st = e.stats()
    Zo, Po, Do = oracle.rk4_timed(oracle.Props(**kw), oracle.Drive(**DRIVE), Z, P, dt, t0, 20)
    ez, ep, ed = _close(Zg, Zo, 1e-11), _close(Pg, Po, 1e-11), _close(Dg, Do, 1e-11)
    print(f"timed rk4 {solver} trim={force_trim}: |dz| {ez:.2e} |dphi| {ep:.2e} |dD| {ed:.2e} redone "
          f"{st['redone_steps']}", flush=True)
    assert st["unconverged"] == 0
    if force_trim:
        assert st["redone_steps"] > 0
    # the drive acts: without it the trajectory differs well above the bound
    Zn, Pn = oracle.rk4(oracle.Props(**kw), Z, P, dt, 20)
    assert np.abs(Pn - Po).max() > 1e-6


def _opto(csh, **kw):
    base = dict(detuning=2e3, gamma=5e3, G=1e11, tau=2e-3, max_intensity=1e18, initial_time=0.0,
                location_x0_mode=0.5e-3, sigma_optical_mode=1.25e-4, beta=2e3, damping_strength=0.02)
    base.update(kw)
    return csh.COptomechanicalVariables(**base)


def _si_drive(oracle, N, L, depth, rho, ov, augmented, timed):
    # the nondimensional drive the reference derives (adimensionalizeProperties,
    # adimensionalizeOptomechanicalVariables, get_current_intensity_drive_strength)
    alpha, rho_he, hbar = 3.5e-24, 150.0, 1.054571817e-34
    L0 = L / (2 * np.pi)
    t0 = np.sqrt(L0 / (3 * alpha / depth ** 4))
    E0 = 3 * rho_he * alpha * L0 ** 4 / depth ** 4
    rho_a = rho / rho_he
    G = ov.G * t0 * L0
    tau = ov.tau / t0
    sig = ov.sigma_optical_mode / L0
    d = oracle.Drive(augmented=augmented, damping=ov.damping_strength, strength=hbar / (E0 * t0 * rho_a) * G / sig ** 2,
                     detuning=ov.detuning * t0, gamma=ov.gamma * t0, G=G, max_intensity=ov.max_intensity,
                     x0=ov.location_x0_mode / L0, sigma=sig, beta=ov.beta * (hbar / E0 / t0) * G / tau / (sig ** 2 * rho_a),
                     tau=tau)
    props = oracle.Props(kind=oracle.HELIUM, depth=depth / L0, rho=rho_a, rho_quirk=True)
    return d, props, t0


def test_dropin_time_dependent_optomechanical_integration(oracle, csh):
    # integrateOptomechanicalSimulationRK4 (Export.cu:780-968) in SI units against the oracle's
    # time-dependent drive on the nondimensional problem; the logged times are those of the start
    # of each step (RK4_Time_Dependent.cuh:402-408)
    lib = csh.lib()
    N, L, depth, rho = 64, 1e-3, 5e-8, 145.0
    sp = csh.CSimulationProperties(L, rho, 0.0, depth, False, 1, False)
    ov = _opto(csh)
    d, props, tb = _si_drive(oracle, N, L, depth, rho, ov, False, True)
    dt_si, t0_si = 0.002 * tb, 0.3 * tb
    rk = csh.CRK4Options(dt_si, t0_si, t0_si + 10 * dt_si + 1e-6 * dt_si, True)
    x = 2 * np.pi * np.arange(N) / N
    y = 1e-4 * np.cos(x)
    phi = 1e-6 * np.sin(2 * x)
    init = np.concatenate([x, y, phi])
    dp = ctypes.POINTER(ctypes.c_double)
    states, times = dp(), dp()
    ns, nt = ctypes.c_size_t(), ctypes.c_size_t()
    rc = lib.integrateOptomechanicalSimulationRK4(init.ctypes.data_as(dp), ctypes.byref(states), ctypes.byref(ns),
                                                  ctypes.byref(times), ctypes.byref(nt), ctypes.byref(sp),
                                                  ctypes.byref(rk), ctypes.byref(ov), N)
    assert rc == 0, csh._lib.last_error()
    assert ns.value == 10 and nt.value == 10
    out = np.ctypeslib.as_array(states, shape=(10, 3 * N)).copy()
    tt = np.ctypeslib.as_array(times, shape=(10,)).copy()
    lib.integrateOptomechanicalSimulationRK4_freeMemory(states, times)
    Zo, Po, _ = oracle.rk4_timed(props, d, x + 1j * y, phi, dt_si / tb, t0_si / tb, 10)
    Zn, _ = oracle.rk4(props, x + 1j * y, phi, dt_si / tb, 10)
    print(f"drop-in timed drive: strength {d.strength:.3e} beta {d.beta:.3e}, effect {np.abs(Zn - Zo).max():.2e}",
          flush=True)
    assert np.abs(Zn - Zo).max() > 1e-8
    _close(out[-1, :N] + 1j * out[-1, N:2 * N], Zo, 1e-11)
    _close(out[-1, 2 * N:], Po, 1e-11)
    assert np.allclose(tt, t0_si / tb + np.arange(10) * (dt_si / tb), rtol=1e-13, atol=0)


def test_dropin_augmented_rhs(oracle, csh):
    # calculateRhsAugmentedOptomechanical (Export.cu:1108-1209): [vx, vy, dphi, dD] of [x, y, phi, D]
    lib = csh.lib()
    N, L, depth, rho = 128, 1e-3, 5e-8, 145.0
    sp = csh.CSimulationProperties(L, rho, 0.0, depth, False, 1, False)
    ov = _opto(csh)
    d, props, _ = _si_drive(oracle, N, L, depth, rho, ov, True, False)
    x = 2 * np.pi * np.arange(N) / N
    y = 1e-4 * np.cos(x) + 2e-5 * np.sin(3 * x)
    phi = 1e-6 * np.sin(2 * x)
    D = 1e-3 * np.cos(x)
    state = np.concatenate([x, y, phi, D])
    rhs = np.zeros(4 * N)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.calculateRhsAugmentedOptomechanical(state.ctypes.data_as(dp), rhs.ctypes.data_as(dp), ctypes.byref(sp),
                                                 ctypes.byref(ov), N)
    assert rc == 0, csh._lib.last_error()
    q, dphi, dD = oracle.rhs_driven(props, d, x + 1j * y, phi, D)
    _close(rhs[:N] + 1j * rhs[N:2 * N], q, 1e-12)
    _close(rhs[2 * N:3 * N], dphi, 1e-12)
    _close(rhs[3 * N:], dD, 1e-12)
    assert lib.calculateRhsAugmentedOptomechanical(state.ctypes.data_as(dp), rhs.ctypes.data_as(dp),
                                                   ctypes.byref(sp), ctypes.byref(ov), 100) == -2


def test_dropin_perturbed_states_layout(csh):
    # calculatePerturbedStates256 (Export.cu:601-660, createInitialBatchedZ JacobianCalculator.cuh:11-76)
    lib = csh.lib()
    n, eps = 256, 1e-3
    rng = np.random.default_rng(3)
    x, y, phi = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    out = np.zeros(6 * n * n, dtype=np.complex128)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.calculatePerturbedStates256(x.ctypes.data_as(dp), y.ctypes.data_as(dp), phi.ctypes.data_as(dp),
                                         out.ctypes.data, 1e-6, 0.0, 0.0, 1e-8, eps)
    assert rc == 0, csh._lib.last_error()
    Zs, Ps = out[:3 * n * n].reshape(3 * n, n), out[3 * n * n:].reshape(3 * n, n)
    Z = x + 1j * y
    for b in (0, 17, n + 5, 2 * n + 200, 3 * n - 1):
        c, p = divmod(b, n)
        ez, ep = Z.copy(), phi.astype(np.complex128)
        if c == 0:
            ez[p] += eps
        elif c == 1:
            ez[p] += 1j * eps
        else:
            ep[p] += eps
        assert np.array_equal(Zs[b], ez) and np.array_equal(Ps[b], ep), b
