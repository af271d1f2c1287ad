"""Long-horizon GPU parity of the benchmark workloads against golden fixtures of the CPU FP64
oracle (tests/golden/*.npz, written by tests/golden/make_golden.py from the seeded generator in
cusuperhelium.workloads).

Errors are per instance at each stored time: max_j |z_j - z_j^oracle|, and max_j |mu_j - mu_j^oracle|
of the right-hand side at the final state (the GPU's state and the oracle's, each with its own
solver).  The workload metric's bound is 1e-9; the FP64 paths are held to tighter bounds where the
horizon allows:
  c1 (N = 256, 100 steps of dt = 0.01)            LU and GMRES         1e-12
  c2 (N = 2048, 1000 steps)                       GMRES and LU         1e-9 (rounding differences
                                                  grow with the surface's high modes, about x1e6
                                                  over the 1000 steps)
  c3 (512 surfaces, 20 steps; 8 surfaces, 500)    batched mixed, FP64  1e-9 per instance
  c5 (film N = 4096, 20 steps)                    GMRES FP64, mixed    1e-11 / 1e-9; LU 2 steps 1e-12
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


@pytest.fixture(scope="module")
def W():
    from cusuperhelium import workloads
    return workloads


def _golden(name):
    path = os.path.join(GOLDEN, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} not generated (tests/golden/make_golden.py)")
    with np.load(path) as f:
        return {k: f[k] for k in f.files}


def _run_marks(e, dt, marks):
    """States after each step count in marks (increasing), from the handle's current state."""
    out, done = [], 0
    for m in marks:
        e.step(dt, int(m) - done)
        done = int(m)
        out.append(e.get_state())
    return out


def _mu_scale(mu):
    return 1.0 + np.abs(mu).max()


@pytest.mark.parametrize("solver", ["lu", "gmres"])
def test_c1_n256_100_steps(csh, W, solver):
    g = _golden("c1_water_n256.npz")
    Z0, P0, dt, steps = W.c1()
    sv = csh.LU if solver == "lu" else csh.GMRES
    with csh.Ensemble([256], csh.Props(), solver=sv, tol=1e-12) as e:
        e.set_state(Z0, P0)
        (Z, P), = _run_marks(e, dt, g["steps"])
        _, _, d = e.rhs(diagnostics=True)
        st = e.stats()
    ez, ep = np.abs(Z - g["Z"][0]).max(), np.abs(P - g["phi"][0]).max()
    em = np.abs(d["mu"] - g["mu"][0]).max()
    print(f"c1 {solver}: |dz| {ez:.2e} |dphi| {ep:.2e} |dmu| {em:.2e} unconverged {st['unconverged']}", flush=True)
    assert st["unconverged"] == 0
    assert ez < 1e-12 and ep < 1e-12 and em < 1e-12 * _mu_scale(g["mu"][0])


@pytest.mark.parametrize("solver", ["gmres", "lu"])
def test_c2_n2048_1000_steps(csh, W, solver):
    g = _golden("c2_water_n2048.npz")
    Z0, P0 = W.c2()
    sv = csh.LU if solver == "lu" else csh.GMRES
    with csh.Ensemble([2048], csh.Props(), solver=sv, tol=1e-12, max_iter=32) as e:
        e.set_state(Z0, P0)
        states = _run_marks(e, W.DT, g["steps"])
        _, _, d = e.rhs(diagnostics=True)
        st = e.stats()
    ez = [float(np.abs(Z - g["Z"][k]).max()) for k, (Z, _) in enumerate(states)]
    ep = [float(np.abs(P - g["phi"][k]).max()) for k, (_, P) in enumerate(states)]
    em = float(np.abs(d["mu"] - g["mu"][-1]).max())
    print(f"c2 {solver}: steps {list(g['steps'])} |dz| {['%.2e' % v for v in ez]} |dphi| "
          f"{['%.2e' % v for v in ep]} |dmu| {em:.2e} unconverged {st['unconverged']} redone "
          f"{st['redone_steps']}", flush=True)
    assert st["unconverged"] == 0
    assert max(ez) < 1e-9 and max(ep) < 1e-9
    assert em < 1e-9 * _mu_scale(g["mu"][-1])


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_c3_every_instance_20_steps(csh, W, prec):
    g = _golden("c3_ensemble_512_20steps.npz")
    sizes, Zs = W.c3()
    assert list(g["sizes"]) == sizes
    p = csh.MIXED if prec == "mixed" else csh.FP64
    with csh.Ensemble(sizes, csh.Props(), precision=p, solver=csh.BATCHED, tol=1e-12, max_iter=16) as e:
        e.set_state(np.concatenate(Zs), np.zeros(sum(sizes)))
        e.step(W.DT, int(g["steps"]))
        Z, P = e.get_state()
        _, _, d = e.rhs(diagnostics=True)
        st = e.stats()
    off = np.concatenate([[0], np.cumsum(sizes)])
    ez = np.array([np.abs(Z[a:b] - g["Z"][a:b]).max() for a, b in zip(off[:-1], off[1:])])
    em = np.array([np.abs(d["mu"][a:b] - g["mu"][a:b]).max() / _mu_scale(g["mu"][a:b])
                   for a, b in zip(off[:-1], off[1:])])
    worst = int(np.argmax(ez))
    print(f"c3 {prec}: 512 instances |dz| max {ez.max():.2e} (#{worst}, N={sizes[worst]}) median "
          f"{np.median(ez):.2e}; |dmu|/(1+|mu|) max {em.max():.2e}; unconverged {st['unconverged']}", flush=True)
    assert st["unconverged"] == 0
    assert ez.max() < 1e-9 and em.max() < 1e-9


def test_c3_sample_500_steps(csh, W):
    g = _golden("c3_sample_500steps.npz")
    sizes, Zs = W.c3()
    idx = [int(i) for i in g["index"]]
    sub = [sizes[i] for i in idx]
    with csh.Ensemble(sub, csh.Props(), precision=csh.MIXED, solver=csh.BATCHED, tol=1e-12, max_iter=16) as e:
        e.set_state(np.concatenate([Zs[i] for i in idx]), np.zeros(sum(sub)))
        states = _run_marks(e, W.DT, g["steps"])
        _, _, d = e.rhs(diagnostics=True)
        st = e.stats()
    off = np.concatenate([[0], np.cumsum(sub)])
    rows = []
    for k, (Z, _) in enumerate(states):
        rows.append([float(np.abs(Z[a:b] - g["Z"][k][a:b]).max()) for a, b in zip(off[:-1], off[1:])])
    em = max(float(np.abs(d["mu"][a:b] - g["mu"][-1][a:b]).max()) for a, b in zip(off[:-1], off[1:]))
    for k, m in enumerate(g["steps"]):
        print(f"c3 sample after {m} steps: |dz| per surface (N={sub}) {['%.1e' % v for v in rows[k]]}", flush=True)
    print(f"c3 sample |dmu| {em:.2e} unconverged {st['unconverged']}", flush=True)
    assert st["unconverged"] == 0
    assert max(max(r) for r in rows) < 1e-9
    assert em < 1e-9 * _mu_scale(g["mu"][-1])


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_c5_film_n4096_20_steps(csh, W, prec):
    g = _golden("c5_film_n4096.npz")
    Z0, P0 = W.c5()
    p = csh.MIXED if prec == "mixed" else csh.FP64
    bound = 1e-11 if prec == "fp64" else 1e-9
    with csh.Ensemble([4096], csh.Props(**W.C5_PROPS), precision=p, solver=csh.GMRES, tol=1e-12,
                      max_iter=24) as e:
        e.set_state(Z0, P0)
        states = _run_marks(e, W.DT, g["steps"])
        _, _, d = e.rhs(diagnostics=True)
        st = e.stats()
    ez = [float(np.abs(Z - g["Z"][k]).max()) for k, (Z, _) in enumerate(states)]
    ep = [float(np.abs(P - g["phi"][k]).max()) for k, (_, P) in enumerate(states)]
    em = float(np.abs(d["mu"] - g["mu"][-1]).max())
    print(f"c5 {prec}: steps {list(g['steps'])} |dz| {['%.2e' % v for v in ez]} |dphi| "
          f"{['%.2e' % v for v in ep]} |dmu| {em:.2e} unconverged {st['unconverged']} redone "
          f"{st['redone_steps']}", flush=True)
    assert st["unconverged"] == 0
    assert max(ez) < bound and max(ep) < bound
    assert em < bound * _mu_scale(g["mu"][-1])


def test_c5_film_n4096_lu_2_steps(csh, W):
    g = _golden("c5_film_n4096.npz")
    assert int(g["steps"][0]) == 2
    Z0, P0 = W.c5()
    with csh.Ensemble([4096], csh.Props(**W.C5_PROPS), solver=csh.LU) as e:
        e.set_state(Z0, P0)
        e.step(W.DT, 2)
        Z, P = e.get_state()
        st = e.stats()
    ez, ep = np.abs(Z - g["Z"][0]).max(), np.abs(P - g["phi"][0]).max()
    print(f"c5 lu: 2 steps |dz| {ez:.2e} |dphi| {ep:.2e} gepp fallbacks {st['fallback_solves']}", flush=True)
    assert ez < 1e-12 and ep < 1e-12
