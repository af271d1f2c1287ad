"""GPU parity of the gfx950 stepper against the CPU FP64 oracle, through the C ABI.

Tolerances (FP64 path): the two implementations use different FFTs, cot formulas, LU
summation orders and velocity summation orders, so they agree to rounding, not bitwise:
  right-hand side, mu, Z', Z''  : <= 1e-12 * (1 + max|reference|)
  RK4 trajectories              : <= 1e-11 after the step counts below
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


def _surface(n, seed, amp=0.02, modes=8):
    # 8-mode seeded surface: a_k ~ U(0, amp/k), theta_k ~ U(0, 2 pi)
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, modes + 1))
    th = rng.uniform(0, 2 * np.pi, modes)
    y = sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(modes))
    return x + 1j * y


def _props(O, c, **kw):
    return O.Props(**kw), c.Props(**kw)


def _close(a, b, tol):
    scale = 1.0 + np.abs(b).max()
    err = np.abs(a - b).max()
    assert err <= tol * scale, f"max err {err:.3e} > {tol:.1e} * {scale:.3e}"


CASES = [
    dict(n=1, kind=0), dict(n=2, kind=0), dict(n=4, kind=0), dict(n=16, kind=0),
    dict(n=64, kind=0), dict(n=96, kind=0), dict(n=256, kind=0), dict(n=768, kind=0),
    dict(n=2048, kind=0),
    dict(n=256, kind=1, depth=1.0), dict(n=1024, kind=1, depth=1.0),
    dict(n=128, kind=1, depth=0.7, infinite_depth=True),
    dict(n=64, kind=1, depth=1.0, kappa=0.01),
    dict(n=64, kind=1, depth=1.0, use_expansions=True, expansion_order=3),
    dict(n=64, kind=0, rho=0.3),
    dict(n=64, kind=0, g=0.15, U=0.2),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}{v}" for k, v in c.items()))
def test_rhs_matches_oracle(oracle, csh, case):
    case = dict(case)
    n = case.pop("n")
    po, pc = _props(oracle, csh, **case)
    print(f"rhs parity n={n} {case}", flush=True)
    Z = _surface(n, 20261202 + n, amp=0.05)
    rng = np.random.default_rng(n)
    Phi = 0.01 * np.sin(2 * np.pi * np.arange(n) / n + rng.uniform())
    q, dp, d = oracle.rhs(po, Z, Phi, diagnostics=True)
    with csh.Ensemble([n], pc) as e:
        e.set_state(Z, Phi)
        qg, dpg, dg = e.rhs(diagnostics=True)
        st = e.stats()
    _close(dg["Zp"], d["Zp"], 1e-12)
    _close(dg["Zpp"], d["Zpp"], 1e-12)
    _close(dg["mu"], d["mu"], 1e-12)
    _close(qg, q, 1e-12)
    _close(dpg, dp, 1e-12)
    assert st["fallback_solves"] == 0


def test_flat_surface_is_exact_potential_flow(csh):
    n, m, eps = 128, 3, 1e-3
    x = 2 * np.pi * np.arange(n) / n
    with csh.Ensemble([n]) as e:
        e.set_state(x + 0j, eps * np.sin(m * x))
        q, dp = e.rhs()
    assert np.abs(q - eps * m * (np.cos(m * x) + 1j * np.sin(m * x))).max() < 1e-15
    assert np.abs(dp - 0.5 * eps * eps * m * m).max() < 1e-18


def test_rk4_config1_matches_oracle(oracle, csh):
    # N = 256 nodes, y = 0.1 cos x, phi = 0, water g = 1, dt = 0.01, 100 steps
    n = 256
    x = 2 * np.pi * np.arange(n) / n
    Z0, P0 = x + 0.1j * np.cos(x), np.zeros(n)
    Zo, Po = oracle.rk4(oracle.Props(), Z0, P0, 0.01, 100)
    with csh.Ensemble([n]) as e:
        e.set_state(Z0, P0)
        e.step(0.01, 100)
        Zg, Pg = e.get_state()
    print("config1 max |dZ|", np.abs(Zg - Zo).max(), "max |dphi|", np.abs(Pg - Po).max())
    assert np.abs(Zg - Zo).max() < 1e-11
    assert np.abs(Pg - Po).max() < 1e-11


def test_rk4_n2048_short_run_matches_oracle(oracle, csh):
    n = 2048
    Z0 = _surface(n, 20261202)
    P0 = np.zeros(n)
    Zo, Po = oracle.rk4(oracle.Props(), Z0, P0, 0.002, 3)
    with csh.Ensemble([n]) as e:
        e.set_state(Z0, P0)
        e.step(0.002, 3)
        Zg, Pg = e.get_state()
    assert np.abs(Zg - Zo).max() < 1e-12
    assert np.abs(Pg - Po).max() < 1e-12


def test_graph_replay_matches_direct_launches(csh):
    n = 256
    Z0 = _surface(n, 7, amp=0.05)
    out = []
    for graph in (True, False):
        with csh.Ensemble([n], use_graph=graph) as e:
            e.set_state(Z0, np.zeros(n))
            e.step(0.005, 3)
            e.step(0.005, 2)
            out.append(e.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_uneven_ensemble_matches_single_surfaces(oracle, csh):
    sizes = [64, 128, 96, 256, 32]
    Zs = [_surface(n, 100 + i, amp=0.05) for i, n in enumerate(sizes)]
    Ps = [0.01 * np.cos(2 * np.pi * np.arange(n) / n) for n in sizes]
    with csh.Ensemble(sizes) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(0.01, 5)
        Zg, Pg = e.get_state()
    for i, n in enumerate(sizes):
        Zo, Po = oracle.rk4(oracle.Props(), Zs[i], Ps[i], 0.01, 5)
        sl = slice(int(sum(sizes[:i])), int(sum(sizes[:i + 1])))
        assert np.abs(Zg[sl] - Zo).max() < 1e-12
        assert np.abs(Pg[sl] - Po).max() < 1e-12


def test_film_rk4_matches_oracle(oracle, csh):
    n = 256
    Z0 = _surface(n, 5, amp=0.05)
    kw = dict(kind=1, depth=1.0)
    Zo, Po = oracle.rk4(oracle.Props(**kw), Z0, np.zeros(n), 0.002, 10)
    with csh.Ensemble([n], csh.Props(**kw)) as e:
        e.set_state(Z0, np.zeros(n))
        e.step(0.002, 10)
        Zg, Pg = e.get_state()
    assert np.abs(Zg - Zo).max() < 1e-12
    assert np.abs(Pg - Po).max() < 1e-12


# ---------------------------------------------------------------- LU solver on its own
@pytest.mark.parametrize("n,kind", [(5, "dominant"), (64, "dominant"), (200, "dominant"),
                                    (130, "pivoting"), (64, "pivoting"), (1, "pivoting")])
def test_lu_solve_matches_lapack(csh, n, kind):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    if kind == "dominant":
        A = A * 0.01 / n + np.eye(n) * 0.5
    b = rng.standard_normal(n)
    x = np.empty(n)
    flag = ctypes.c_int(-1)
    dp = ctypes.POINTER(ctypes.c_double)
    A = np.ascontiguousarray(A)
    rc = csh.lib().sfh_debug_lu_solve(n, A.ctypes.data_as(dp), b.ctypes.data_as(dp), x.ctypes.data_as(dp),
                                      ctypes.byref(flag))
    assert rc == 0
    xr = np.linalg.solve(A, b)
    assert np.abs(x - xr).max() <= 1e-13 * np.linalg.cond(A) * max(1.0, np.abs(xr).max()) * n
    assert flag.value == (1 if (kind == "pivoting" and n > 1) else 0)


# ---------------------------------------------------------------- drop-in C ABI
def test_dropin_rhs_matches_oracle_helium(oracle, csh):
    n = 256
    L, rho, kappa, depth = 1e-6, 145.0, 0.0, 15e-9
    u = oracle.adimensionalize(L, rho, kappa, depth)
    Z = _surface(n, 11, amp=0.02)
    phi = 0.001 * np.sin(np.real(Z))
    sm = csh.SimulationManager()
    res, vx, vy, dphi = sm.calculate_rhs_from_vectors(Z.real.copy(), Z.imag.copy(), phi, L, rho, kappa, depth)
    assert res == 0
    q, dp = oracle.rhs(oracle.Props(kind=1, rho=u["rho"], kappa=u["kappa"], depth=u["depth"]), Z, phi)
    _close(vx + 1j * vy, q, 1e-12)
    _close(dphi, dp, 1e-12)
    res2, vx2, vy2, dphi2 = sm.calculate_rhs256_from_vectors(Z.real.copy(), Z.imag.copy(), phi, L, rho, kappa,
                                                             depth)
    assert res2 == 0 and np.array_equal(vx2, vx) and np.array_equal(dphi2, dphi)


def test_dropin_batched_and_vorticities(oracle, csh):
    L, rho, kappa, depth = 1e-6, 0.0, 0.0, 15e-9
    u = oracle.adimensionalize(L, rho, kappa, depth)
    B, n = 4, 256
    Zs = [_surface(n, 30 + b, amp=0.02) for b in range(B)]
    phis = [0.001 * np.cos(b + np.real(z)) for b, z in enumerate(Zs)]
    x = np.concatenate([z.real for z in Zs])
    y = np.concatenate([z.imag for z in Zs])
    p = np.concatenate(phis)
    sm = csh.SimulationManager()
    res, vx, vy, dphi = sm.calculate_rhs256_from_vectors_batched(x, y, p, L, rho, kappa, depth, B)
    assert res == 0
    P = oracle.Props(kind=1, rho=u["rho"], depth=u["depth"])
    for b in range(B):
        q, dp = oracle.rhs(P, Zs[b], phis[b])
        _close((vx + 1j * vy)[b * n:(b + 1) * n], q, 1e-12)
        _close(dphi[b * n:(b + 1) * n], dp, 1e-12)
    res, a, zp, zpp = sm.calculate_vorticities256_from_vectors(Zs[0], phis[0] + 0j, L, rho, kappa, depth)
    assert res == 0
    _, _, d = oracle.rhs(P, Zs[0], phis[0], diagnostics=True)
    _close(a, d["mu"], 1e-12)
    _close(zp, d["Zp"], 1e-12)
    _close(zpp, d["Zpp"], 1e-12)


def test_dropin_derivative_fft256(oracle, csh):
    rng = np.random.default_rng(3)
    f = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    res, out = csh.SimulationManager().calculate_derivativeFFT256(f)
    assert res == 0
    _close(out, oracle.fft_derivative(f), 1e-12)


def test_dropin_integrate_rk4_trajectory(oracle, csh):
    n = 64
    L, rho, kappa, depth = 1e-6, 0.0, 0.0, 15e-9
    u = oracle.adimensionalize(L, rho, kappa, depth)
    Z = _surface(n, 9, amp=0.02)
    y0 = np.concatenate([Z.real, Z.imag, np.zeros(n)])
    sp = csh.CSimulationProperties(L, rho, kappa, depth, False, 1, False)
    dt_nd = 0.01
    opts = csh.CRK4Options(dt_nd * u["base_time"], 0.0, 5.5 * dt_nd * u["base_time"], True)
    res, times, states = csh.SimulationManager().integrate_rk4(y0, sp, opts)
    assert res == 0 and states.shape == (5, 3 * n) and times.shape == (5,)
    assert np.allclose(times, dt_nd * np.arange(1, 6), rtol=1e-12)
    Zo, Po, tr = oracle.rk4(oracle.Props(kind=1, depth=u["depth"]), Z, np.zeros(n), dt_nd, 5, trajectory=True)
    for s in range(5):
        _close(states[s, :n] + 1j * states[s, n:2 * n], tr[s, :n], 1e-11)
        _close(states[s, 2 * n:], tr[s, n:].real, 1e-11)
    opts.returnTrajectory = False
    res, times, states = csh.SimulationManager().integrate_rk4(y0, sp, opts)
    assert res == 0 and states.shape == (1, 3 * n) and times.size == 0
    _close(states[0, :n] + 1j * states[0, n:2 * n], Zo, 1e-11)


def test_dropin_rejects_unsupported_sizes(csh):
    x = np.zeros(100)
    res, *_ = csh.SimulationManager().calculate_rhs_from_vectors(x, x, x, 1e-6, 0.0, 0.0, 15e-9)
    assert res == -2


# ---------------------------------------------------------------- GMRES alternative
@pytest.mark.parametrize("n,kw", [(256, {}), (2048, {}), (1024, dict(kind=1, depth=1.0)), (96, {})])
def test_gmres_rhs_matches_oracle(oracle, csh, n, kw):
    # GMRES (tol 1e-12 relative residual) against the oracle's GEPP solve
    po, pc = oracle.Props(**kw), csh.Props(**kw)
    Z = _surface(n, 40 + n, amp=0.05)
    Phi = 0.01 * np.sin(2 * np.pi * np.arange(n) / n + 0.3)
    q, dp, d = oracle.rhs(po, Z, Phi, diagnostics=True)
    with csh.Ensemble([n], pc, solver=csh.GMRES, tol=1e-13, max_iter=32) as e:
        e.set_state(Z, Phi)
        qg, dpg, dg = e.rhs(diagnostics=True)
        st = e.stats()
    print(f"gmres n={n} {kw}: iterations {st['last_iterations']}", flush=True)
    assert st["unconverged"] == 0 and st["last_iterations"] > 0
    _close(dg["mu"], d["mu"], 1e-11)
    _close(qg, q, 1e-11)
    _close(dpg, dp, 1e-11)


def test_gmres_rk4_matches_oracle_and_warm_starts(oracle, csh):
    n = 512
    Z0 = _surface(n, 77)
    Zo, Po = oracle.rk4(oracle.Props(), Z0, np.zeros(n), 0.002, 10)
    with csh.Ensemble([n], solver=csh.GMRES, tol=1e-12, max_iter=24) as e:
        e.set_state(Z0, np.zeros(n))
        e.step(0.002, 10)
        Zg, Pg = e.get_state()
        st = e.stats()
    print("gmres rk4 warm-start iterations", st["last_iterations"], st["max_iterations"])
    assert st["unconverged"] == 0 and st["max_iterations"] <= 12
    assert np.abs(Zg - Zo).max() < 1e-12
    assert np.abs(Pg - Po).max() < 1e-12


@pytest.mark.parametrize("kw", [{}, dict(kind=1, depth=1.0)])
def test_gmres_trimmed_sequence_matches_lu(csh, kw):
    # After the warm-start history fills (5 steps) and one more full-length step, the step graph
    # keeps only the launches the solves of that step needed.  Trajectories split over several step() calls
    # (cold steps, trimmed batches, a re-captured graph) against the LU stepper on the same surface;
    # both FP64, agreement 1e-11 after 40 steps, no solve runs out of launches.
    n = 1024
    Z0 = _surface(n, 91, amp=0.04)
    P0 = 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    pc = csh.Props(**kw)
    with csh.Ensemble([n], pc, solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(0.002, 40)
        Zl, Pl = e.get_state()
    with csh.Ensemble([n], pc, solver=csh.GMRES, tol=1e-13, max_iter=32) as e:
        e.set_state(Z0, P0)
        for k in (3, 2, 7, 28):
            e.step(0.002, k)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"trimmed gmres {kw}: last {st['last_iterations']} max {st['max_iterations']} "
          f"unconverged {st['unconverged']} |dZ| {np.abs(Zg - Zl).max():.2e}", flush=True)
    assert st["unconverged"] == 0
    assert np.abs(Zg - Zl).max() < 1e-11
    assert np.abs(Pg - Pl).max() < 1e-11


@pytest.mark.parametrize("kw,redo", [(dict(kind=1, depth=1.0), True), ({}, False)])
def test_gmres_trimmed_chunks_are_redone_until_converged(csh, kw, redo):
    # force_trim = 2 gives the helium solves (4-6 Arnoldi steps from the warm start) one Arnoldi
    # step per solve: every solve of the first trimmed chunk runs out of launches, so the chunk is
    # restored from its checkpoint (state, last solution, warm-start history) and run again with the
    # sequence doubled until its solves converge.  No kept step may have an unconverged solve, and
    # the trajectory (split over several step() calls) matches the LU stepper to 1e-11 after 40
    # steps.  The water surface takes 1-2 Arnoldi steps per solve at tol 1e-13, so length 2 runs out
    # at most in the first chunk (2 steps), after which length 4 holds.
    n = 1024
    Z0 = _surface(n, 91, amp=0.04)
    P0 = 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    pc = csh.Props(**kw)
    with csh.Ensemble([n], pc, solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(0.002, 40)
        Zl, Pl = e.get_state()
    with csh.Ensemble([n], pc, solver=csh.GMRES, tol=1e-13, max_iter=32, force_trim=2) as e:
        e.set_state(Z0, P0)
        for k in (3, 2, 7, 28):
            e.step(0.002, k)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"forced trim 2 {kw}: redone {st['redone_steps']} unconverged {st['unconverged']} "
          f"max iterations {st['max_iterations']} |dZ| {np.abs(Zg - Zl).max():.2e}", flush=True)
    assert st["unconverged"] == 0
    assert st["redone_steps"] > 0 if redo else st["redone_steps"] <= 2
    assert st["steps"] == 40
    assert np.abs(Zg - Zl).max() < 1e-11
    assert np.abs(Pg - Pl).max() < 1e-11


@pytest.mark.parametrize("solver", ["gmres", "lu", "batched"])
def test_step_record_matches_stepping_and_reading(csh, solver):
    # 100 steps recorded every 3rd (33 frames through the 32-slot device ring: it wraps) against a
    # twin handle stepped 3 at a time and read back with sfh_get_state; the leftover step runs too
    sizes = [256, 128]
    sv = {"gmres": csh.GMRES, "lu": csh.LU, "batched": csh.BATCHED}[solver]
    Z0 = np.concatenate([_surface(n, 60 + i) for i, n in enumerate(sizes)])
    P0 = 0.005 * np.sin(np.arange(sum(sizes)) * 0.1)
    with csh.Ensemble(sizes, solver=sv, tol=1e-12) as e:
        e.set_state(Z0, P0)
        Zf, Pf, tf = e.step_record(0.002, 100, 3)
        Zend, Pend = e.get_state()
    assert Zf.shape == (33, sum(sizes)) and tf.shape == (33,)
    with csh.Ensemble(sizes, solver=sv, tol=1e-12) as e:
        e.set_state(Z0, P0)
        for f in range(33):
            e.step(0.002, 3)
            Z, P = e.get_state()
            assert np.array_equal(Z, Zf[f]) and np.array_equal(P, Pf[f]), f
        e.step(0.002, 1)
        Z, P = e.get_state()
    assert np.array_equal(Z, Zend) and np.array_equal(P, Pend)
    t = 0.0
    for f in range(33):
        for _ in range(3):
            t += 0.002
        assert tf[f] == t


def test_gmres_one_step_sequence_ends_in_velocity_kernel(csh):
    # Once every solve takes one Arnoldi step, the two-launch sequence (the product, then
    # velocity_gm1_kernel, which finishes the step, stores mu and forms mu' from the partial sums
    # of the product launch) has no gm_close_mu_kernel launch.  Kernel-timed steps run it eagerly
    # between graph-replayed ones; the trajectory against the LU stepper after 30 steps (FP64,
    # 1e-11).  Water: helium solves take 4-6 Arnoldi steps from the warm start (deep or finite
    # depth, dt = 0.002 or 1e-4), so they never reach this path.
    kw, dt, tol, bound = {}, 0.002, 1e-12, 1e-11
    n = 2048
    Z0 = _surface(n, 17)
    P0 = 0.005 * np.sin(2 * np.pi * np.arange(n) / n)
    pc = csh.Props(**kw)
    with csh.Ensemble([n], pc, solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(dt, 30)
        Zl, Pl = e.get_state()
    with csh.Ensemble([n], pc, solver=csh.GMRES, tol=tol, max_iter=32) as e:
        e.set_state(Z0, P0)
        e.step(dt, 12)
        e.kernel_timing(True)
        e.step(dt, 3)
        e.sync()
        close = e.kernel_report("gm_close_mu")
        vel = e.kernel_report("velocity")
        prod = e.kernel_report("gm_iter")
        e.kernel_timing(False)
        e.step(dt, 15)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"one-step sequence {kw}: close {close['launches']} velocity {vel['launches']} products "
          f"{prod['launches']} unconverged {st['unconverged']} |dZ| {np.abs(Zg - Zl).max():.2e}", flush=True)
    assert close["launches"] == 0 and vel["launches"] == 12 and prod["launches"] == 12
    assert st["unconverged"] == 0
    assert np.abs(Zg - Zl).max() < bound
    assert np.abs(Pg - Pl).max() < bound


def test_gmres_one_step_sequence_uneven_ensemble(csh):
    # the fused one-step path on an ensemble of uneven surfaces: each grid slice reduces its own
    # surface's partial sums, stores its own mu and history, and closes its own control block;
    # every surface against the LU stepper after 24 steps
    sizes = [2048, 1024, 512]
    Zs = [_surface(n, 40 + i) for i, n in enumerate(sizes)]
    Ps = [0.004 * np.cos(2 * np.pi * np.arange(n) / n + i) for i, n in enumerate(sizes)]
    Z0, P0 = np.concatenate(Zs), np.concatenate(Ps)
    with csh.Ensemble(sizes, solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(0.002, 24)
        Zl, Pl = e.get_state()
    with csh.Ensemble(sizes, solver=csh.GMRES, tol=1e-12, max_iter=32) as e:
        e.set_state(Z0, P0)
        e.step(0.002, 10)
        e.kernel_timing(True)
        e.step(0.002, 2)
        e.sync()
        close = e.kernel_report("gm_close_mu")
        vel = e.kernel_report("velocity")
        e.kernel_timing(False)
        e.step(0.002, 12)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"one-step ensemble: close {close['launches']} velocity {vel['launches']} unconverged "
          f"{st['unconverged']} |dZ| {np.abs(Zg - Zl).max():.2e}", flush=True)
    assert close["launches"] == 0 and vel["launches"] == 8
    assert st["unconverged"] == 0
    assert np.abs(Zg - Zl).max() < 1e-11
    assert np.abs(Pg - Pl).max() < 1e-11


def test_gmres_one_step_sequence_n4096_matches_lu(csh):
    # at N = 4096 the one-step sequence's velocity kernel has 1024 workgroups, more than the GPU holds
    # at once, so some start after workgroup 0 has stamped the solve finished: they must still apply
    # the step's correction (only a finish stamped by the product launch keeps the initial guess).
    # Water, 14 steps (the last 8 on the one-step sequence) against the LU stepper.
    n, dt = 4096, 0.002
    Z0 = _surface(n, 23)
    P0 = 0.004 * np.sin(2 * np.pi * np.arange(n) / n)
    with csh.Ensemble([n], solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(dt, 14)
        Zl, Pl = e.get_state()
    with csh.Ensemble([n], solver=csh.GMRES, tol=1e-12, max_iter=32) as e:
        e.set_state(Z0, P0)
        e.step(dt, 8)
        e.kernel_timing(True)
        e.step(dt, 2)
        e.sync()
        vel = e.kernel_report("velocity")
        close = e.kernel_report("gm_close_mu")
        e.kernel_timing(False)
        e.step(dt, 4)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"one-step sequence n=4096: velocity {vel['launches']} close {close['launches']} unconverged "
          f"{st['unconverged']} |dZ| {np.abs(Zg - Zl).max():.2e}", flush=True)
    assert vel["launches"] == 8 and close["launches"] == 0
    assert st["unconverged"] == 0
    assert np.abs(Zg - Zl).max() < 1e-11
    assert np.abs(Pg - Pl).max() < 1e-11


def test_gmres_uneven_ensemble_matches_oracle(oracle, csh):
    # several surfaces of different N in one GMRES solve sequence (one grid slice per surface,
    # each with its own control block, partial dots and history), 12 steps: the trimmed sequence
    # starts after step 6; every surface against the oracle's GEPP stepper
    sizes = [256, 96, 1024, 12, 384]
    Zs = [_surface(n, 300 + i, amp=0.05) for i, n in enumerate(sizes)]
    Ps = [0.01 * np.cos(2 * np.pi * np.arange(n) / n + i) for i, n in enumerate(sizes)]
    with csh.Ensemble(sizes, solver=csh.GMRES, tol=1e-13, max_iter=24) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(0.004, 5)
        e.step(0.004, 7)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"gmres ensemble: last {st['last_iterations']} max {st['max_iterations']} "
          f"unconverged {st['unconverged']}", flush=True)
    assert st["unconverged"] == 0
    o = 0
    for i, n in enumerate(sizes):
        Zo, Po = oracle.rk4(oracle.Props(), Zs[i], Ps[i], 0.004, 12)
        assert np.abs(Zg[o:o + n] - Zo).max() < 1e-11, n
        assert np.abs(Pg[o:o + n] - Po).max() < 1e-11, n
        o += n


# ---------------------------------------------------------------- mixed precision, multi-workgroup GMRES
def test_mixed_gmres_uneven_ensemble_matches_oracle(oracle, csh):
    # FP32 matrix, GMRES over it from an FP64 r0, then the FP64 residual of the solution formed
    # matrix-free and a refinement solve from it: the trajectory must stay within 1e-9 of the FP64
    # oracle (GEPP), over the cold steps and the trimmed sequence after them
    sizes = [256, 96, 1024, 384]
    Zs = [_surface(n, 400 + i, amp=0.05) for i, n in enumerate(sizes)]
    Ps = [0.01 * np.cos(2 * np.pi * np.arange(n) / n + i) for i, n in enumerate(sizes)]
    with csh.Ensemble(sizes, precision=csh.MIXED, solver=csh.GMRES, tol=1e-12, max_iter=24) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(0.004, 5)
        e.step(0.004, 7)
        Zg, Pg = e.get_state()
        st = e.stats()
    print(f"mixed gmres ensemble: last {st['last_iterations']} max {st['max_iterations']} "
          f"unconverged {st['unconverged']} refinements {st['refinements']}", flush=True)
    assert st["unconverged"] == 0
    o = 0
    for i, n in enumerate(sizes):
        Zo, Po = oracle.rk4(oracle.Props(), Zs[i], Ps[i], 0.004, 12)
        assert np.abs(Zg[o:o + n] - Zo).max() < 1e-9, n
        assert np.abs(Pg[o:o + n] - Po).max() < 1e-9, n
        o += n


def test_mixed_gmres_film_matches_oracle(oracle, csh):
    # the helium film with the image term (an outlier eigenvalue, several Arnoldi steps per solve)
    n = 1024
    kw = dict(kind=1, depth=1.0)
    Z0 = _surface(n, 7, amp=0.05)
    Zo, Po = oracle.rk4(oracle.Props(**kw), Z0, np.zeros(n), 0.002, 6)
    with csh.Ensemble([n], csh.Props(**kw), precision=csh.MIXED, solver=csh.GMRES, tol=1e-12, max_iter=24) as e:
        e.set_state(Z0, np.zeros(n))
        e.step(0.002, 6)
        Zg, Pg = e.get_state()
        st = e.stats()
    assert st["unconverged"] == 0
    assert np.abs(Zg - Zo).max() < 1e-9
    assert np.abs(Pg - Po).max() < 1e-9


def test_c5_film_500_steps_mixed_and_gmres_match_the_plain_lu(csh):
    # the N = 4096 film over the config's 500 steps: mixed GMRES (FP32 matrix, FP64 refinement) and FP64
    # GMRES against the plain FP64 stepper (blocked LU, itself checked against the oracle's golden states
    # over its first 20 steps)
    from cusuperhelium import workloads as W
    Z0, P0 = W.c5()
    n = len(Z0)
    out = {}
    for solver, prec in ((csh.LU, csh.FP64), (csh.GMRES, csh.FP64), (csh.GMRES, csh.MIXED)):
        print(f"c5 film 500 steps solver {solver} precision {prec}", flush=True)
        with csh.Ensemble([n], csh.Props(**W.C5_PROPS), precision=prec, solver=solver, tol=1e-12, max_iter=24) as e:
            e.set_state(Z0, P0)
            e.step(0.002, 500)
            out[solver, prec] = e.get_state()
    ZL, PL = out[csh.LU, csh.FP64]
    assert np.all(np.isfinite(ZL))
    for key in ((csh.GMRES, csh.FP64), (csh.GMRES, csh.MIXED)):
        Z, P = out[key]
        dz, dp = np.abs(Z - ZL).max(), np.abs(P - PL).max()
        print(f"  {key}: max |dz| {dz:.2e} |dphi| {dp:.2e}", flush=True)
        assert dz <= 1e-9 and dp <= 1e-9
