"""Pair kernels with 1, 2 and 4 rows per workgroup, and the autotuner that picks among them
(sfh_autotune, sfh_tuned, SFH_SOLVE_AUTO), on the GPU against the CPU FP64 oracle.

Tolerances as for the other FP64 paths: right-hand sides 1e-12 relative to 1 + max|reference|,
RK4 trajectories 1e-11 after 10 steps (mixed precision: 1e-9, the refinement tolerance bound).
"""
import os

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


def _close(a, b, tol):
    err = np.abs(np.asarray(a) - np.asarray(b)).max()
    assert err <= tol * (1.0 + np.abs(b).max()), f"max err {err:.3e}"
    return err


@pytest.mark.parametrize("rows", [1, 2, 4])
@pytest.mark.parametrize("form", ["lu", "gmres", "gmres_mixed"])
@pytest.mark.parametrize("kw", [dict(), dict(kind=1, depth=1.0)], ids=["water", "film"])
def test_pair_rows_match_oracle(oracle, csh, rows, form, kw):
    n, dt, steps = 192, 0.004, 10
    Z = _surface(n, 31 + rows)
    P = 0.01 * np.cos(2 * np.pi * np.arange(n) / n)
    solver = csh.LU if form == "lu" else csh.GMRES
    prec = csh.MIXED if form.endswith("mixed") else csh.FP64
    print(f"pair rows {rows} {form} {kw}", flush=True)
    q_o, dp_o = oracle.rhs(oracle.Props(**kw), Z, P)
    Zo, Po = oracle.rk4(oracle.Props(**kw), Z, P, dt, steps)
    with csh.Ensemble([n], csh.Props(**kw), solver=solver, precision=prec, pair_rows=rows, tol=1e-13) as e:
        e.set_state(Z, P)
        q, dp = e.rhs()
        e.step(dt, steps)
        Zg, Pg = e.get_state()
        st = e.stats()
    rtol = 1e-9 if prec == csh.MIXED else 1e-12
    _close(q, q_o, rtol)
    _close(dp, dp_o, rtol)
    _close(Zg, Zo, 1e-9 if prec == csh.MIXED else 1e-11)
    _close(Pg, Po, 1e-9 if prec == csh.MIXED else 1e-11)
    assert st["unconverged"] == 0


def test_pair_rows_rejects_other_values(csh):
    for rows in (3, 8, -1):
        with pytest.raises(Exception, match="pair_rows"):
            csh.Ensemble([64], csh.Props(), pair_rows=rows)


def test_autotune_caches_and_auto_solver_matches_oracle(oracle, csh, tmp_path):
    cache = tmp_path / "tuning.txt"
    os.environ["SFH_TUNE_CACHE"] = str(cache)
    try:
        n = 256
        t = csh.autotune(n, steps=4)
        print(f"autotune n={n}: {t}", flush=True)
        assert not t["from_cache"]
        assert t["solver"] in (csh.LU, csh.GMRES, csh.BATCHED) and t["pair_rows"] in (1, 2, 4)
        assert t["precision"] == csh.FP64 and 0.0 < t["ms_per_step"] < 1e3
        lines = [ln.split() for ln in cache.read_text().splitlines() if not ln.startswith("#")]
        assert any(ln[1:5] == ["0", "0", str(n), "0"] for ln in lines), lines
        t2 = csh.tuned(n)
        assert t2["from_cache"] and (t2["solver"], t2["pair_rows"]) == (t["solver"], t["pair_rows"])
        Z = _surface(n, 5)
        P = np.zeros(n)
        Zo, Po = oracle.rk4(oracle.Props(), Z, P, 0.004, 10)
        with csh.Ensemble([n], csh.Props(), solver=csh.AUTO, tol=1e-13) as e:
            e.set_state(Z, P)
            e.step(0.004, 10)
            Zg, Pg = e.get_state()
        _close(Zg, Zo, 1e-11)
        _close(Pg, Po, 1e-11)
        # an uneven ensemble of small surfaces takes the batched form, with no tuning run
        with csh.Ensemble([64, 96], csh.Props(), solver=csh.AUTO) as e:
            assert e.stats()["solver"] == csh.BATCHED
    finally:
        del os.environ["SFH_TUNE_CACHE"]
