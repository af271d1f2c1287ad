"""GPU checks of the hand-written adjoint (water, with and without an upper fluid, and the helium film)
through the C ABI.

- right-hand side VJP: dot-product test <J^T l, d> = d/dh <l, f(y + h d)> against central
  differences of the CPU FP64 oracle (h = 1e-6, relative agreement 1e-7);
- one RK4 step VJP: the same test against the oracle's RK4 step;
- the differentiable-stepper workload: N = 1024, 16 Fourier amplitudes c_k ~ U(0, 0.01),
  y0 = sum c_k cos(k x), loss = 1/2 sum_j (y_j(T) - y*_j)^2 against a seeded target (seed + 1),
  checkpoint every 20 steps; gradient from torch.autograd through the adjoint against central
  finite differences (h = 1e-6) of the GPU stepper (all 32 perturbed surfaces advanced as one
  batched ensemble), relative error <= 1e-6.  With dt = 0.005 the unfiltered method (the
  reference's, reproduced by the oracle) is unstable on this surface: the Nyquist mode grows about
  100x every 10 steps from rounding level and overflows near step 100, so T = 200 dt is checked at
  dt = 0.001 (T = 0.2) and dt = 0.005 is checked over 60 steps;
- the same pipeline at N = 64 over 20 steps against central differences of the CPU FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


def _surface(n, seed, amp=0.03, modes=8):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, modes + 1))
    th = rng.uniform(0, 2 * np.pi, modes)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(modes))


def _inner(gZ, gP, dZ, dP):
    return float(np.real(np.sum(np.conj(gZ) * dZ)) + np.sum(gP * dP))


def _smooth_direction(rng, n, modes=8):
    # band-limited perturbation (modes 1..8 of x and phi): white noise would make the central
    # difference's truncation error grow like N^2 through the curvature terms
    x = 2 * np.pi * np.arange(n) / n
    k = np.arange(1, modes + 1)[:, None]
    cz = (rng.standard_normal(modes) + 1j * rng.standard_normal(modes)) / k[:, 0]
    cp = rng.standard_normal(modes) / k[:, 0]
    dZ = (cz[:, None] * np.exp(1j * k * x)).sum(0)
    dP = (cp[:, None] * np.cos(k * x)).sum(0)
    return dZ, dP


def _richardson(f, h=1e-5):
    # central differences at h and h/2 combined to cancel the O(h^2) term
    d1 = (f(h) - f(-h)) / (2 * h)
    d2 = (f(h / 2) - f(-h / 2)) / h
    return (4 * d2 - d1) / 3


FILMS = [dict(kind=1, depth=1.0), dict(kind=1, depth=0.4), dict(kind=1, depth=0.7, infinite_depth=True),
         dict(kind=1, depth=1.0, use_expansions=True, expansion_order=3), dict(kind=1, depth=1.0, g=0.15),
         dict(kind=1, depth=1.0, kappa=0.01)]
RHO = [dict(rho=0.3), dict(rho=0.3, rho_quirk=False)]   # water under a lighter fluid; the V1[1] quirk on, off
CASES = [([64], {}), ([256, 96, 1024], {})] + [([128], f) for f in FILMS] + [([256, 64, 512], FILMS[0])] + \
    [([128, 64], RHO[0]), ([96], RHO[1])]


@pytest.mark.parametrize("sizes,kw", CASES, ids=lambda v: "-".join(f"{k}{x}" for k, x in v.items()) if isinstance(v, dict) else str(v))
def test_rhs_vjp_dot_product(oracle, csh, sizes, kw):
    print(f"rhs vjp sizes {sizes} {kw}", flush=True)
    rng = np.random.default_rng(11)
    Zs = [_surface(n, 100 + n) for n in sizes]
    Ps = [0.02 * np.sin(2 * np.pi * np.arange(n) / n + 0.3) for n in sizes]
    lq = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in sizes]
    lp = [rng.standard_normal(n) for n in sizes]
    with csh.Ensemble(sizes, csh.Props(**kw), solver=csh.BATCHED, tol=1e-13) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        gZ, gP = e.rhs_vjp(np.concatenate(lq), np.concatenate(lp))
    o = 0
    for i, n in enumerate(sizes):
        dZ, dP = _smooth_direction(rng, n)
        fd = _richardson(lambda h: _inner(lq[i], lp[i], *oracle.rhs(oracle.Props(**kw), Zs[i] + h * dZ,
                                                                      Ps[i] + h * dP)))
        ad = _inner(gZ[o:o + n], gP[o:o + n], dZ, dP)
        print(f"  n={n}: adjoint {ad:.12e}  fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1e-3)
        o += n


@pytest.mark.parametrize("kw", [{}, FILMS[0], RHO[0]], ids=["water", "film", "rho"])
def test_step_vjp_dot_product(oracle, csh, kw):
    n, dt = 128, 0.005
    rng = np.random.default_rng(5)
    Z, P = _surface(n, 7), 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    aZ = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    aP = rng.standard_normal(n)
    with csh.Ensemble([n], csh.Props(**kw), solver=csh.BATCHED, tol=1e-13) as e:
        e.set_state(Z, P)
        gZ, gP = e.step_vjp(dt, aZ, aP)
        Zs, _ = e.get_state()
    assert np.array_equal(Zs, Z)   # the handle's state is left unchanged
    dZ, dP = _smooth_direction(rng, n)
    fd = _richardson(lambda h: _inner(aZ, aP, *oracle.rk4(oracle.Props(**kw), Z + h * dZ, P + h * dP, dt, 1)))
    ad = _inner(gZ, gP, dZ, dP)
    print(f"step vjp: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
    assert abs(ad - fd) <= 1e-7 * abs(fd)


MWG = [([1024], {}, "gmres"), ([768], {}, "lu"), ([1024], FILMS[0], "gmres"), ([512], FILMS[1], "lu"),
       ([256, 256], FILMS[0], "gmres"), ([2048], {}, "gmres"), ([512], FILMS[5], "gmres"),
       ([512], RHO[0], "gmres"), ([384, 192], RHO[0], "lu"), ([384], RHO[1], "gmres")]


@pytest.mark.parametrize("sizes,kw,solver", MWG, ids=lambda v: str(v))
def test_rhs_vjp_multi_workgroup_dot_product(oracle, csh, sizes, kw, solver):
    # the reverse pass over many workgroups per surface (LU and GMRES handles) against central
    # differences of the CPU oracle
    print(f"multi-workgroup rhs vjp sizes {sizes} {kw} {solver}", flush=True)
    rng = np.random.default_rng(13)
    Zs = [_surface(n, 200 + n + i) for i, n in enumerate(sizes)]
    Ps = [0.02 * np.sin(2 * np.pi * np.arange(n) / n + 0.3 + i) for i, n in enumerate(sizes)]
    lq = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in sizes]
    lp = [rng.standard_normal(n) for n in sizes]
    sv = csh.GMRES if solver == "gmres" else csh.LU
    with csh.Ensemble(sizes, csh.Props(**kw), solver=sv, tol=1e-13) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        gZ, gP = e.rhs_vjp(np.concatenate(lq), np.concatenate(lp))
    o = 0
    for i, n in enumerate(sizes):
        dZ, dP = _smooth_direction(rng, n)
        fd = _richardson(lambda h: _inner(lq[i], lp[i], *oracle.rhs(oracle.Props(**kw), Zs[i] + h * dZ,
                                                                      Ps[i] + h * dP)))
        ad = _inner(gZ[o:o + n], gP[o:o + n], dZ, dP)
        print(f"  n={n}: adjoint {ad:.12e}  fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1e-3)
        o += n


@pytest.mark.parametrize("kw", [{}, FILMS[0]], ids=["water", "film"])
def test_step_vjp_multi_workgroup_matches_batched(csh, kw):
    # one RK4 step's VJP through the multi-workgroup pass (GMRES handle) against the one-workgroup
    # pass (batched handle): the same cotangent to the solves' tolerance
    n, dt = 512, 0.004
    rng = np.random.default_rng(17)
    Z, P = _surface(n, 19), 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    aZ = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    aP = rng.standard_normal(n)
    out = {}
    for sv in (csh.BATCHED, csh.GMRES):
        with csh.Ensemble([n], csh.Props(**kw), solver=sv, tol=1e-13) as e:
            e.set_state(Z, P)
            out[sv] = e.step_vjp(dt, aZ, aP)
    for a, b in zip(out[csh.GMRES], out[csh.BATCHED]):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"step vjp multi-workgroup vs batched: rel {err:.2e}", flush=True)
        assert err < 1e-9


def _amplitude_problem(n, modes, seed):
    x = 2 * np.pi * np.arange(n) / n
    c = np.random.default_rng(seed).uniform(0, 0.01, modes)
    ct = np.random.default_rng(seed + 1).uniform(0, 0.01, modes)
    basis = np.array([np.cos((k + 1) * x) for k in range(modes)])
    return x, c, basis, ct @ basis


def _adjoint_gradient(csh, n, modes, seed, dt, steps, every, kw=None):
    import torch
    from cusuperhelium.adjoint import Trajectory, evolve
    x, c, basis, target = _amplitude_problem(n, modes, seed)
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    B = torch.tensor(basis)
    Z0 = torch.tensor(x, dtype=torch.complex128) + 1j * (ct @ B).to(torch.complex128)
    P0 = torch.zeros(n, dtype=torch.float64)
    with Trajectory([n], csh.Props(**(kw or {})), checkpoint_every=every) as tr:
        ZT, _ = evolve(Z0, P0, tr, dt, steps)
        loss = 0.5 * torch.sum((ZT.imag - torch.tensor(target)) ** 2)
        loss.backward()
    return x, c, basis, target, float(loss.detach()), ct.grad.numpy()


@pytest.mark.parametrize("dt,steps", [(0.005, 60), (0.001, 200)])
def test_differentiable_stepper_n1024_matches_fd(csh, dt, steps):
    n, modes, seed = 1024, 16, 20261202
    print(f"adjoint gradient n=1024, dt {dt}, {steps} steps", flush=True)
    x, c, basis, target, loss, grad = _adjoint_gradient(csh, n, modes, seed, dt, steps, 20)
    # central differences: the 32 perturbed initial surfaces advanced together
    h = 1e-6
    Z0s, P0s = [], []
    for k in range(modes):
        for sgn in (+1, -1):
            cc = c.copy()
            cc[k] += sgn * h
            Z0s.append(x + 1j * (cc @ basis))
            P0s.append(np.zeros(n))
    with csh.Ensemble([n] * (2 * modes), csh.Props(), solver=csh.BATCHED, tol=1e-14, max_iter=48) as e:
        e.set_state(np.concatenate(Z0s), np.concatenate(P0s))
        e.step(dt, steps)
        ZT, _ = e.get_state()
    yT = ZT.imag.reshape(2 * modes, n)
    L = 0.5 * np.sum((yT - target) ** 2, axis=1)
    fd = (L[0::2] - L[1::2]) / (2 * h)
    rel = np.abs(grad - fd).max() / np.abs(fd).max()
    print(f"loss {loss:.6e}; max |grad - fd| / max |fd| = {rel:.2e}", flush=True)
    assert np.isfinite(loss) and rel <= 1e-6


def test_differentiable_stepper_small_matches_oracle_fd(oracle, csh):
    n, modes, seed, dt, steps = 64, 4, 77, 0.005, 20
    x, c, basis, target, loss, grad = _adjoint_gradient(csh, n, modes, seed, dt, steps, 7)
    h = 1e-6
    fd = np.empty(modes)
    for k in range(modes):
        Ls = []
        for sgn in (+1, -1):
            cc = c.copy()
            cc[k] += sgn * h
            ZT, _ = oracle.rk4(oracle.Props(), x + 1j * (cc @ basis), np.zeros(n), dt, steps)
            Ls.append(0.5 * np.sum((ZT.imag - target) ** 2))
        fd[k] = (Ls[0] - Ls[1]) / (2 * h)
    rel = np.abs(grad - fd).max() / np.abs(fd).max()
    print(f"oracle fd: grad {grad} fd {fd} rel {rel:.2e}", flush=True)
    assert rel <= 1e-6


def test_differentiable_film_small_matches_oracle_fd(oracle, csh):
    # the same pipeline on the helium film (depth 1, image terms on) against central differences
    # of the CPU FP64 oracle
    kw = FILMS[0]
    n, modes, seed, dt, steps = 64, 4, 78, 0.005, 20
    x, c, basis, target, loss, grad = _adjoint_gradient(csh, n, modes, seed, dt, steps, 7, kw)
    h = 1e-6
    fd = np.empty(modes)
    for k in range(modes):
        Ls = []
        for sgn in (+1, -1):
            cc = c.copy()
            cc[k] += sgn * h
            ZT, _ = oracle.rk4(oracle.Props(**kw), x + 1j * (cc @ basis), np.zeros(n), dt, steps)
            Ls.append(0.5 * np.sum((ZT.imag - target) ** 2))
        fd[k] = (Ls[0] - Ls[1]) / (2 * h)
    rel = np.abs(grad - fd).max() / np.abs(fd).max()
    print(f"film oracle fd: grad {grad} fd {fd} rel {rel:.2e}", flush=True)
    assert rel <= 1e-6


DRIVE = dict(damping=0.05, strength=0.3, detuning=0.4, gamma=0.8, G=2.0, max_intensity=0.6, x0=np.pi,
             sigma=0.7, beta=0.9, tau=1.5)


@pytest.mark.parametrize("solver,sizes", [("BATCHED", [128, 64]), ("GMRES", [512]), ("LU", [256])])
def test_rhs_vjp_autonomous_drive_dot_product(oracle, csh, solver, sizes):
    # the optically driven film (damping q1.y and the intensity term in dPhi/dt) against central
    # differences of the oracle's driven right-hand side
    kw = FILMS[0]
    print(f"driven rhs vjp {solver} {sizes}", flush=True)
    rng = np.random.default_rng(17)
    Zs = [_surface(n, 300 + n) for n in sizes]
    Ps = [0.02 * np.sin(2 * np.pi * np.arange(n) / n + 0.5) for n in sizes]
    lq = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in sizes]
    lp = [rng.standard_normal(n) for n in sizes]
    with csh.Ensemble(sizes, csh.Props(**kw), solver=getattr(csh, solver), tol=1e-13) as e:
        e.set_drive(**DRIVE)
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        gZ, gP = e.rhs_vjp(np.concatenate(lq), np.concatenate(lp))
    drive = oracle.Drive(augmented=False, **DRIVE)
    o = 0
    for i, n in enumerate(sizes):
        dZ, dP = _smooth_direction(rng, n)

        def f(h):
            q, dp, _ = oracle.rhs_driven(oracle.Props(**kw), drive, Zs[i] + h * dZ, Ps[i] + h * dP, np.zeros(n))
            return _inner(lq[i], lp[i], q, dp)
        fd = _richardson(f)
        ad = _inner(gZ[o:o + n], gP[o:o + n], dZ, dP)
        print(f"  n={n}: adjoint {ad:.12e}  fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1e-3)
        o += n


@pytest.mark.parametrize("kw,drive", [(dict(kind=1, depth=1.0), "timed")])
def test_adjoint_rejects_unsupported_physics(csh, kw, drive):
    with csh.Ensemble([64], csh.Props(**kw), solver=csh.BATCHED) as e:
        if drive:
            e.set_drive(**{drive: True}, damping=0.1)
        e.set_state(_surface(64, 1), np.zeros(64))
        with pytest.raises(Exception, match="adjoint is implemented"):
            e.rhs_vjp(np.ones(64, complex), np.ones(64))


def _host_sweep(e, Z0, P0, dt, steps, every, gZ, gP):
    # the checkpointed reverse sweep driven step by step through the host (set_state, step, get_state,
    # step_vjp per step, each recomputing its stages from zero-guess solves): the device-resident sweep,
    # which keeps the stages of its warm-started re-run, must agree to the solver tolerance
    e.set_state(Z0, P0)
    ck = []
    done = 0
    while done < steps:
        ck.append(e.get_state())
        n = min(every, steps - done)
        e.step(dt, n)
        done += n
    ZT, PT = e.get_state()
    for seg in range(len(ck) - 1, -1, -1):
        n = min(every, steps - seg * every)
        e.set_state(*ck[seg])
        states = [ck[seg]]
        for _ in range(n - 1):
            e.step(dt, 1)
            states.append(e.get_state())
        for i in range(n - 1, -1, -1):
            e.set_state(*states[i])
            gZ, gP = e.step_vjp(dt, gZ, gP)
    return ZT, PT, gZ, gP


@pytest.mark.parametrize("solver,sizes,every,steps,kw", [
    ("GMRES", [1024], 4, 11, None),
    ("LU", [256], 3, 11, None),
    ("BATCHED", [128, 96], 5, 11, None),
    ("GMRES", [256], 4, 11, "film"),
    # two surfaces of one GMRES handle: per-surface Krylov blocks of the trimmed, warm-started transposed
    # solves (the calibration falls in the first swept segment, the second is trimmed)
    ("GMRES", [512, 256], 8, 16, None),
    ("GMRES", [384], 4, 9, "rho"),
    ("BATCHED", [128], 4, 9, "drive"),   # the film under the autonomous optical drive
])
def test_device_trajectory_sweep_matches_host_sweep(csh, solver, sizes, every, steps, kw):
    dt = 0.004
    props = csh.Props(**({"film": FILMS[0], "rho": RHO[0], "drive": FILMS[0]}.get(kw, {})))
    rng = np.random.default_rng(31)
    Z0 = np.concatenate([_surface(n, 40 + i) for i, n in enumerate(sizes)])
    P0 = np.concatenate([0.01 * np.cos(2 * np.pi * np.arange(n) / n) for n in sizes])
    gZ = rng.standard_normal(len(Z0)) + 1j * rng.standard_normal(len(Z0))
    gP = rng.standard_normal(len(Z0))
    print(f"device trajectory sweep {solver} sizes {sizes} every {every} {kw or 'water'}", flush=True)
    with csh.Ensemble(sizes, props, solver=getattr(csh, solver), tol=1e-13,
                      max_iter=48 if solver == "BATCHED" else 32) as e:
        if kw == "drive":
            e.set_drive(**DRIVE)
        ZT_h, PT_h, gZ_h, gP_h = _host_sweep(e, Z0, P0, dt, steps, every, gZ, gP)
        e.set_state(Z0, P0)
        rec = e.traj_forward(dt, steps, every)
        ZT, PT = e.get_state()
        gZ_d, gP_d = e.traj_backward(rec, gZ, gP)
        Zs, Ps = e.get_state()   # the record's initial state
        gZ_d2, gP_d2 = e.traj_backward(rec, gZ, gP)   # a record can be swept back again
        e.traj_release(rec)
        with pytest.raises(RuntimeError):
            e.traj_backward(rec, gZ, gP)
        with pytest.raises(RuntimeError):
            e.traj_release(rec)
    dz = np.abs(ZT - ZT_h).max()
    dg = max(np.abs(gZ_d - gZ_h).max(), np.abs(gP_d - gP_h).max()) / max(np.abs(gZ_h).max(), np.abs(gP_h).max())
    print(f"  final state diff {dz:.2e}, cotangent rel diff {dg:.2e}", flush=True)
    assert dz <= 1e-14 and dg <= 1e-10
    assert np.array_equal(Zs, Z0) and np.array_equal(Ps, P0)
    assert np.array_equal(gZ_d, gZ_d2) and np.array_equal(gP_d, gP_d2)


def test_device_trajectory_rejects_unsupported_problem_before_stepping(csh):
    Z0 = _surface(128, 5)
    with csh.Ensemble([128], csh.Props(**FILMS[0]), solver=csh.GMRES) as e:
        e.set_drive(augmented=True, damping=0.1)   # the augmented drive has no adjoint
        e.set_state(Z0, np.zeros(128))
        with pytest.raises(RuntimeError):
            e.traj_forward(0.004, 5, 2)
        Z, _ = e.get_state()
    assert np.array_equal(Z, Z0)   # nothing stepped


def test_device_trajectory_redoes_segments_whose_trimmed_solves_ran_out(csh):
    # force_trim = 2 gives every trimmed GMRES solve one Arnoldi step; the helium film's solves need
    # more from their warm starts, so each segment's re-run runs out of launches after its history
    # fills and must be run again with full sequences, and so does the reverse sweep of the second
    # segment, whose warm-started transposed solves get the forced length after the calibration step
    # of the first.  The forward pass's own redo doubles the length, so the handle is brought back to
    # the forced length (a new state and the 6 steps of its calibration) before the backward pass,
    # which needs only the record
    n, dt, steps, every = 256, 0.004, 14, 7
    props = csh.Props(**FILMS[0])
    rng = np.random.default_rng(32)
    Z0 = _surface(n, 60)
    P0 = 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    gZ = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    gP = rng.standard_normal(n)
    print("device trajectory sweep, film, forced trim 2", flush=True)
    with csh.Ensemble([n], props, solver=csh.GMRES, tol=1e-13, max_iter=32) as e:
        _, _, gZ_h, gP_h = _host_sweep(e, Z0, P0, dt, steps, every, gZ, gP)
    with csh.Ensemble([n], props, solver=csh.GMRES, tol=1e-13, max_iter=32, force_trim=2) as e:
        e.set_state(Z0, P0)
        rec = e.traj_forward(dt, steps, every)
        e.set_state(Z0, P0)
        e.step(dt, 6)
        redone_fwd = e.stats()["redone_steps"]
        gZ_d, gP_d = e.traj_backward(rec, gZ, gP)
        st = e.stats()
        e.traj_release(rec)
    dg = max(np.abs(gZ_d - gZ_h).max(), np.abs(gP_d - gP_h).max()) / max(np.abs(gZ_h).max(), np.abs(gP_h).max())
    print(f"  redone steps before the sweep {redone_fwd}, after it {st['redone_steps']}; "
          f"cotangent rel diff {dg:.2e}", flush=True)
    assert st["redone_steps"] >= redone_fwd + 3 * every   # both re-runs and one reverse sweep again
    assert dg <= 1e-10


@pytest.mark.parametrize("solver,sizes", [("BATCHED", [128, 64]), ("GMRES", [512]), ("LU", [256])])
def test_rhs_vjp_augmented_drive_dot_product(oracle, csh, solver, sizes):
    # the augmented system (the delayed intensity D a third field: dPhi/dt + D, dD/dt = beta I - D/tau)
    # against central differences of the oracle's augmented right-hand side
    kw = FILMS[0]
    print(f"augmented rhs vjp {solver} {sizes}", flush=True)
    rng = np.random.default_rng(19)
    Zs = [_surface(n, 400 + n) for n in sizes]
    Ps = [0.02 * np.sin(2 * np.pi * np.arange(n) / n + 0.7) for n in sizes]
    Ds = [0.05 * np.cos(2 * np.pi * np.arange(n) / n) for n in sizes]
    lq = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in sizes]
    lp = [rng.standard_normal(n) for n in sizes]
    ld = [rng.standard_normal(n) for n in sizes]
    with csh.Ensemble(sizes, csh.Props(**kw), solver=getattr(csh, solver), tol=1e-13) as e:
        e.set_drive(augmented=True, **DRIVE)
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.set_delayed(np.concatenate(Ds))
        gZ, gP, gD = e.rhs_vjp(np.concatenate(lq), np.concatenate(lp), np.concatenate(ld))
        with pytest.raises(Exception):
            e.rhs_vjp(np.concatenate(lq), np.concatenate(lp))   # D's cotangent is required
    drive = oracle.Drive(augmented=True, **DRIVE)
    o = 0
    for i, n in enumerate(sizes):
        dZ, dP = _smooth_direction(rng, n)
        dD = 0.3 * np.sin(2 * np.pi * np.arange(n) / n + 0.2)

        def f(h):
            q, dp, dd = oracle.rhs_driven(oracle.Props(**kw), drive, Zs[i] + h * dZ, Ps[i] + h * dP, Ds[i] + h * dD)
            return _inner(lq[i], lp[i], q, dp) + float(np.sum(ld[i] * dd))
        fd = _richardson(f)
        ad = _inner(gZ[o:o + n], gP[o:o + n], dZ, dP) + float(np.sum(gD[o:o + n] * dD))
        print(f"  n={n}: adjoint {ad:.12e}  fd {fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-7 * max(abs(fd), 1e-3)
        o += n


@pytest.mark.parametrize("solver", ["BATCHED", "GMRES"])
def test_step_vjp_augmented_drive_dot_product(oracle, csh, solver):
    kw = FILMS[0]
    n, dt = 128, 0.005
    rng = np.random.default_rng(23)
    Z, P = _surface(n, 9), 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    D = 0.05 * np.sin(2 * np.pi * np.arange(n) / n)
    aZ = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    aP, aD = rng.standard_normal(n), rng.standard_normal(n)
    with csh.Ensemble([n], csh.Props(**kw), solver=getattr(csh, solver), tol=1e-13) as e:
        e.set_drive(augmented=True, **DRIVE)
        e.set_state(Z, P)
        e.set_delayed(D)
        gZ, gP, gD = e.step_vjp(dt, aZ, aP, aD)
        Zs, _ = e.get_state()
        assert np.array_equal(Zs, Z) and np.array_equal(e.delayed(), D)   # the state is left unchanged
    drive = oracle.Drive(augmented=True, **DRIVE)
    dZ, dP = _smooth_direction(rng, n)
    dD = 0.3 * np.cos(2 * np.pi * np.arange(n) / n + 0.4)

    def f(h):
        Z1, P1, D1 = oracle.rk4_driven(oracle.Props(**kw), drive, Z + h * dZ, P + h * dP, D + h * dD, dt, 1)
        return _inner(aZ, aP, Z1, P1) + float(np.sum(aD * D1))
    fd = _richardson(f)
    ad = _inner(gZ, gP, dZ, dP) + float(np.sum(gD * dD))
    print(f"augmented step vjp {solver}: adjoint {ad:.12e} fd {fd:.12e}", flush=True)
    assert abs(ad - fd) <= 1e-7 * abs(fd)
