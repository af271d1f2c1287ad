"""GPU energies and volume flux (EnergyContainer + VolumeFlux, Energies.cuh:136-204) against the
CPU FP64 oracle's sfo_energies, through the C ABI.

sfh_get_energies returns, per surface, {kinetic, potential, surface, volume flux} of the last
right-hand side evaluated: after sfh_rhs the current state, after sfh_step the fourth RK4 stage of
the last step (the state the reference's getEnergy() reports after runStep).  Tolerance: 1e-12
relative to 1 + |energy| (the sums run in different orders).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _oracle_energies(O, props, Z, P):
    q, _, d = O.rhs(props, Z, P, diagnostics=True)
    return O.energies(props, Z, P, d["Zp"], q)


def _close(a, b, tol=1e-12):
    err = np.abs(np.asarray(a) - np.asarray(b))
    assert np.all(err <= tol * (1.0 + np.abs(b))), f"{a} vs {b}"


CASES = [
    ("lu", 256, dict()),
    ("gmres", 2048, dict()),
    ("lu", 1024, dict(kind=1, depth=1.0)),
    ("gmres", 1024, dict(kind=1, depth=0.5)),
    ("lu", 128, dict(kind=1, depth=1.0, kappa=0.01)),
    ("lu", 64, dict(rho=0.3)),
    ("lu", 64, dict(g=0.15, U=0.2)),
    ("batched", 512, dict()),
    ("batched", 256, dict(kind=1, depth=1.0)),
]


@pytest.mark.parametrize("solver,n,kw", CASES, ids=lambda v: str(v))
def test_energies_at_state_match_oracle(oracle, csh, solver, n, kw):
    Z = _surface(n, 5 + n)
    P = 0.01 * np.sin(2 * np.pi * np.arange(n) / n + 0.4)
    sv = {"lu": csh.LU, "gmres": csh.GMRES, "batched": csh.BATCHED}[solver]
    with csh.Ensemble([n], csh.Props(**kw), solver=sv, tol=1e-13) as e:
        e.set_state(Z, P)
        e.track_energies(True)
        e.rhs()
        eg = e.energies()[0]
    eo = _oracle_energies(oracle, oracle.Props(**kw), Z, P)
    print(f"energies {solver} n={n} {kw}: gpu {eg} oracle {eo}", flush=True)
    _close(eg, eo)


def test_energies_untracked_is_an_error(csh):
    with csh.Ensemble([64], csh.Props()) as e:
        e.set_state(_surface(64, 1), np.zeros(64))
        e.rhs()
        with pytest.raises(csh.SolverError) as ei:
            e.energies()
        assert ei.value.code == -3


@pytest.mark.parametrize("solver", ["gmres", "batched"])
def test_energies_after_step_are_those_of_the_fourth_stage(oracle, csh, solver):
    # An uneven ensemble stepped with tracking off and then on (the step graph is re-captured):
    # the energies reported after the last step are those of its fourth stage, y3 = y0 + dt k3.
    sizes = [256, 128]
    dt = 0.005
    Zs = [_surface(n, 30 + i) for i, n in enumerate(sizes)]
    P0 = np.zeros(sum(sizes))
    sv = csh.GMRES if solver == "gmres" else csh.BATCHED
    with csh.Ensemble(sizes, csh.Props(), solver=sv, tol=1e-13) as e:
        e.set_state(np.concatenate(Zs), P0)
        e.step(dt, 9)
        e.track_energies(True)
        e.step(dt, 3)
        eg = e.energies()
    po = oracle.Props()
    off = 0
    for s, n in enumerate(sizes):
        Z, P = oracle.rk4(po, Zs[s], np.zeros(n), dt, 11)
        q1, p1 = oracle.rhs(po, Z, P)
        q2, p2 = oracle.rhs(po, Z + 0.5 * dt * q1, P + 0.5 * dt * p1)
        q3, p3 = oracle.rhs(po, Z + 0.5 * dt * q2, P + 0.5 * dt * p2)
        Z3, P3 = Z + dt * q3, P + dt * p3
        eo = _oracle_energies(oracle, po, Z3, P3)
        print(f"surface {s} (n={n}) after 12 steps: gpu {eg[s]} oracle {eo}", flush=True)
        _close(eg[s], eo, 1e-11)
        off += n
