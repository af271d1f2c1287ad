"""The C++ plugin API (include/cusuperhelium.hpp) as a reference C++ caller uses it: the check program
tests/cpp/plugin_api_check (built by `make -C tests/cpp`) runs WaterBoundaryProblem +
BaseBoundaryIntegralCalculator::run, AutonomousRungeKuttaStepper::runEvolution with a
TrajectoryLogger from a host state, and runStep on a device state, for two N = 256 surfaces.

run() is compared with the CPU FP64 oracle's right-hand side (1e-12); the logged trajectory and the
device-state steps with the Python handle API stepped one step at a time (the same kernels:
identical bits), and the logged times with dt accumulated per step.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "plugin_api_check")


def test_cpp_plugin_api(oracle, tmp_path):
    import cusuperhelium as c
    assert os.path.exists(EXE), "build it with make -C tests/cpp"
    n, B, dt = 256, 2, 0.01
    x = 2 * np.pi * np.arange(n) / n
    Zs = [x + 1j * (0.04 * np.cos(x) + 0.01 * np.sin(3 * x)), x + 1j * 0.03 * np.cos(2 * x + 0.5)]
    Ps = [0.01 * np.sin(x), np.zeros(n)]
    Z0, P0 = np.concatenate(Zs), np.concatenate(Ps)
    y0 = np.concatenate([Z0, P0 + 0j])
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    y0.astype(np.complex128).tofile(fin)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    M = 2 * n * B
    out = np.fromfile(fout, dtype=np.uint8)
    rhs = out[:16 * M].view(np.complex128)
    states = out[16 * M:16 * M * 11].view(np.complex128).reshape(10, M)
    times = out[16 * M * 11:16 * M * 11 + 80].view(np.float64)
    yd = out[16 * M * 11 + 80:].view(np.complex128)
    # run() against the oracle
    q, dphi = oracle.rhs(oracle.Props(), Z0, P0, batch=B)
    assert np.abs(rhs[:n * B] - q).max() < 1e-12 * (1 + np.abs(q).max())
    assert np.abs(rhs[n * B:].real - dphi).max() < 1e-12 * (1 + np.abs(dphi).max())
    assert np.all(rhs[n * B:].imag == 0)
    # trajectory against the handle API
    with c.Ensemble([n] * B, c.Props(), solver=c.LU, tol=1e-12) as e:
        e.set_state(Z0, P0)
        t = 0.0
        for k in range(10):
            e.step(dt, 1)
            t += dt
            Z, P = e.get_state()
            assert np.array_equal(states[k, :n * B], Z), k
            assert np.array_equal(states[k, n * B:].real, P), k
            assert times[k] == t
    assert np.array_equal(yd, states[9])
