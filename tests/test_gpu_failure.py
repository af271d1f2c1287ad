"""Failure paths forced from the host, through the C ABI: a NaN input state and a surface with two
coincident nodes (the dipole matrix of cot((z_k - z_j)/2) is then singular) are rejected (-3), a
trajectory that overflows is reported (-4) and a new state clears it, an unsupported size returns
-2.  No entry point aborts the process.  All kernels run to their normal end on non-finite values:
their loops have fixed trip counts, and the pivot search of the elimination fallback skips NaN
entries.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SOLVERS = ["lu", "gmres", "batched"]


@pytest.fixture(scope="module")
def csh():
    import cusuperhelium as c
    c.lib()
    return c


def _surface(n):
    x = 2 * np.pi * np.arange(n) / n
    return x + 0.01j * np.cos(x), 0.001 * np.sin(x)


def _sv(csh, name):
    return {"lu": csh.LU, "gmres": csh.GMRES, "batched": csh.BATCHED}[name]


@pytest.mark.parametrize("solver", SOLVERS)
def test_nan_state_is_rejected(csh, solver):
    n = 64
    Z, P = _surface(n)
    P[5] = np.nan
    with csh.Ensemble([n], csh.Props(), solver=_sv(csh, solver)) as e:
        with pytest.raises(csh.SolverError) as ei:
            e.set_state(Z, P)
        assert ei.value.code == -3 and "non-finite" in csh._lib.last_error()


@pytest.mark.parametrize("solver", SOLVERS)
def test_coincident_nodes_are_rejected(csh, solver):
    # two nodes at the same point (also one period apart) make the dipole matrix singular
    n = 64
    Z, P = _surface(n)
    with csh.Ensemble([n, n], csh.Props(), solver=_sv(csh, solver)) as e:
        # (node 0 sits at x = 0, so x = 2 pi lands on it exactly)
        for i, j, shift in ((10, 11, 0.0), (3, 0, 2 * np.pi)):
            Zbad = Z.copy()
            Zbad[i] = Zbad[j] + shift
            with pytest.raises(csh.SolverError) as ei:
                e.set_state(np.concatenate([Z, Zbad]), np.concatenate([P, P]))
            assert ei.value.code == -3
            assert f"coincident nodes {min(i, j)} and {max(i, j)} of surface 1" in csh._lib.last_error()


@pytest.mark.parametrize("solver", SOLVERS)
def test_blown_up_trajectory_reports_nonfinite_then_recovers(csh, solver):
    # the 16-mode N = 1024 surface at dt = 0.005: its Nyquist mode grows about 100x per 10 steps from
    # rounding level and the state overflows between steps 100 and 130 (as in the CPU oracle)
    from cusuperhelium import workloads as W
    x, amp, _, basis = W.c4()
    Z = x + 1j * (amp @ basis)
    P = np.zeros_like(x)
    print(f"blow-up, {solver}", flush=True)
    with csh.Ensemble([x.size], csh.Props(), solver=_sv(csh, solver)) as e:
        e.set_state(Z, P)
        code, done = None, 0
        while done < 200 and code is None:
            try:
                e.step(0.005, 20)
                done += 20
                e.get_state()
            except csh.SolverError as err:
                code = err.code
        print(f"  reported {code} after {done} steps: {csh._lib.last_error()}", flush=True)
        assert code == -4 and 60 <= done <= 200
        with pytest.raises(csh.SolverError):   # the condition holds until a new state
            e.get_state()
        e.set_state(Z, P)
        e.step(0.005, 5)
        Zg, Pg = e.get_state()
        assert np.all(np.isfinite(Zg)) and np.all(np.isfinite(Pg))


def test_dropin_rhs_rejects_nan_and_coincident_nodes(csh):
    # calculateRHS256FromVectors (Export.cu:116-189): nondimensional state, SI film properties
    lib = csh.lib()
    n = 256
    Z, P = _surface(n)
    L, depth = 1e-6, 15e-9
    x0, y0, p0 = Z.real, Z.imag, P
    vx, vy, dp = np.zeros(n), np.zeros(n), np.zeros(n)
    ptr = ctypes.POINTER(ctypes.c_double)

    def call(x, y, phi):
        x, y, phi = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, phi))
        return lib.calculateRHS256FromVectors(x.ctypes.data_as(ptr), y.ctypes.data_as(ptr), phi.ctypes.data_as(ptr),
                                              vx.ctypes.data_as(ptr), vy.ctypes.data_as(ptr),
                                              dp.ctypes.data_as(ptr), L, 145.0, 0.0, depth)

    assert call(x0, y0, p0) == 0 and np.all(np.isfinite(vx))
    y = y0.copy()
    y[3] = np.inf
    assert call(x0, y, p0) == -3
    x, y = x0.copy(), y0.copy()
    x[7], y[7] = x[8], y[8]
    assert call(x, y, p0) == -3
    assert call(x0, y0, p0) == 0 and np.all(np.isfinite(vx))   # the cached handle recovers


def test_unsupported_size_returns_code(csh):
    with pytest.raises(csh.SolverError) as ei:
        csh.Ensemble([100], csh.Props())
    assert ei.value.code == -2
