"""GPU parity of the batched form (one persistent launch per RK stage, one workgroup per surface)
against the CPU FP64 oracle, through the C ABI, in FP64 and in mixed precision.

Tolerances: the solve is iterative (relative residual 1e-12 against |b|), so mu and everything
after it agree with the oracle's GEPP solution to about 1e-12 of the scale; the mixed form refines
in FP64 against a matrix-free FP64 residual and must meet the same bound.  Over a few RK4 steps
the state agrees to 1e-11, and the surface error after the 20-step ensemble run is asserted at the
1e-9 bound of the batched mixed-precision target.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


def _surface(n, seed, amp=0.02, modes=8):
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    a = rng.uniform(0, amp / np.arange(1, modes + 1))
    th = rng.uniform(0, 2 * np.pi, modes)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(modes))


def _close(a, b, tol, what=""):
    scale = 1.0 + np.abs(b).max()
    err = np.abs(a - b).max()
    assert err <= tol * scale, f"{what} max err {err:.3e} > {tol:.1e} * {scale:.3e}"


SIZES = [1, 2, 3, 4, 6, 16, 96, 128, 192, 256, 384, 768, 1024, 512, 64]


def _ensemble_state(sizes, seed, amp=0.05):
    Zs = [_surface(n, seed + i, amp=amp) for i, n in enumerate(sizes)]
    Ps = [0.01 * np.sin(2 * np.pi * np.arange(n) / n + 0.1 * i) for i, n in enumerate(sizes)]
    return Zs, Ps


def _split(v, sizes):
    out, o = [], 0
    for n in sizes:
        out.append(v[o:o + n])
        o += n
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kw", [{}, dict(kind=1, depth=1.0), dict(kind=0, rho=0.3), dict(kind=0, g=0.15, U=0.2)],
                         ids=["water", "film", "rho", "gU"])
def test_batched_rhs_matches_oracle(oracle, csh, precision, kw):
    print(f"batched rhs {precision} {kw}", flush=True)
    Zs, Ps = _ensemble_state(SIZES, 300)
    prec = csh.MIXED if precision == "mixed" else csh.FP64
    with csh.Ensemble(SIZES, csh.Props(**kw), precision=prec, solver=csh.BATCHED, tol=1e-13) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        qg, dpg, dg = e.rhs(diagnostics=True)
        st = e.stats()
    assert st["unconverged"] == 0
    for i, n in enumerate(SIZES):
        q, dp, d = oracle.rhs(oracle.Props(**kw), Zs[i], Ps[i], diagnostics=True)
        tag = f"surface {i} n={n}"
        _close(_split(dg["Zp"], SIZES)[i], d["Zp"], 1e-12, tag + " Zp")
        _close(_split(dg["Zpp"], SIZES)[i], d["Zpp"], 1e-12, tag + " Zpp")
        _close(_split(dg["mu"], SIZES)[i], d["mu"], 2e-12, tag + " mu")
        _close(_split(qg, SIZES)[i], q, 2e-12, tag + " q")
        _close(_split(dpg, SIZES)[i], dp, 2e-12, tag + " dphi")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_rk4_matches_oracle(oracle, csh, precision):
    sizes = [256, 128, 1024, 512, 96, 256]
    Zs, Ps = _ensemble_state(sizes, 500, amp=0.02)
    steps, dt = 5, 0.002
    prec = csh.MIXED if precision == "mixed" else csh.FP64
    print(f"batched rk4 {precision} sizes {sizes}", flush=True)
    with csh.Ensemble(sizes, csh.Props(), precision=prec, solver=csh.BATCHED, tol=1e-12) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(dt, steps)
        Zg, Pg = e.get_state()
        st = e.stats()
    assert st["unconverged"] == 0
    for i, n in enumerate(sizes):
        Zr, Pr = oracle.rk4(oracle.Props(), Zs[i], Ps[i], dt, steps)
        _close(_split(Zg, sizes)[i], Zr, 1e-11, f"surface {i} n={n} Z")
        _close(_split(Pg, sizes)[i], Pr, 1e-11, f"surface {i} n={n} Phi")


def test_batched_matches_single_surface_stepper(csh):
    # the batched form and the plain FP64 LU stepper on the same surfaces
    sizes = [128, 256]
    Zs, Ps = _ensemble_state(sizes, 700, amp=0.03)
    with csh.Ensemble(sizes, csh.Props(), solver=csh.BATCHED, tol=1e-13) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(0.005, 10)
        Zb, Pb = e.get_state()
    for i, n in enumerate(sizes):
        with csh.Ensemble([n], csh.Props()) as e:
            e.set_state(Zs[i], Ps[i])
            e.step(0.005, 10)
            Z1, P1 = e.get_state()
        _close(_split(Zb, sizes)[i], Z1, 1e-11, f"n={n} Z")
        _close(_split(Pb, sizes)[i], P1, 1e-11, f"n={n} Phi")


def test_batched_graph_replay_matches_direct(csh):
    sizes = [256, 64, 1024]
    Zs, Ps = _ensemble_state(sizes, 900, amp=0.02)
    out = []
    for graph in (True, False):
        with csh.Ensemble(sizes, csh.Props(), precision=csh.MIXED, solver=csh.BATCHED, use_graph=graph) as e:
            e.set_state(np.concatenate(Zs), np.concatenate(Ps))
            e.step(0.002, 4)
            out.append(e.get_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_batched_ensemble_20_steps_within_1e9(oracle, csh):
    # a slice of the 512-surface workload: 8-mode draws, N from {128, 256, 512, 1024}, dt 0.002
    rng = np.random.default_rng(20261202)
    sizes = [int(v) for v in rng.choice([128, 256, 512, 1024], size=12)]
    Zs = [_surface(n, 20261202 + i) for i, n in enumerate(sizes)]
    Ps = [np.zeros(n) for n in sizes]
    steps, dt = 20, 0.002
    with csh.Ensemble(sizes, csh.Props(), precision=csh.MIXED, solver=csh.BATCHED, tol=1e-12) as e:
        e.set_state(np.concatenate(Zs), np.concatenate(Ps))
        e.step(dt, steps)
        Zg, _ = e.get_state()
        st = e.stats()
    assert st["unconverged"] == 0
    worst = 0.0
    for i, n in enumerate(sizes):
        Zr, _ = oracle.rk4(oracle.Props(), Zs[i], Ps[i], dt, steps)
        worst = max(worst, float(np.abs(_split(Zg, sizes)[i] - Zr).max()))
    print(f"mixed batched: max |z - z_oracle| over {len(sizes)} surfaces after {steps} steps = {worst:.2e}")
    assert worst <= 1e-9


def test_batched_c3_every_instance_500_steps_mixed_within_1e9_of_fp64(csh):
    # the whole 512-surface workload over its 500 steps: every instance of the mixed form against the
    # FP64 form of the same kernel (itself checked against the oracle above and in the golden
    # 500-step sample)
    from cusuperhelium import workloads as W
    sizes, Zs = W.c3()
    Z0 = np.concatenate(Zs)
    fin = {}
    print(f"c3 ensemble: {len(sizes)} surfaces, 500 steps, mixed and fp64", flush=True)
    for prec in (csh.MIXED, csh.FP64):
        with csh.Ensemble(sizes, csh.Props(), precision=prec, solver=csh.BATCHED, tol=1e-12, max_iter=16) as e:
            e.set_state(Z0, np.zeros(len(Z0)))
            e.step(0.002, 500)
            fin[prec] = e.get_state()[0]
            assert e.stats()["unconverged"] == 0
    dz = np.array([np.abs(a - b).max() for a, b in zip(_split(fin[csh.MIXED], sizes), _split(fin[csh.FP64], sizes))])
    print(f"  max |z_mixed - z_fp64| per instance: max {dz.max():.2e}, median {np.median(dz):.2e}")
    assert np.all(np.isfinite(dz)) and dz.max() <= 1e-9


def test_batched_rejects_large_surfaces(csh):
    with pytest.raises(Exception):
        csh.Ensemble([2048], csh.Props(), solver=csh.BATCHED)
    with pytest.raises(Exception):
        csh.Ensemble([256], csh.Props(), precision=csh.MIXED, solver=csh.LU)
    # the multi-workgroup GMRES keeps the vector being multiplied in LDS: at most 4096 nodes
    with pytest.raises(Exception):
        csh.Ensemble([8192], csh.Props(), solver=csh.GMRES)
