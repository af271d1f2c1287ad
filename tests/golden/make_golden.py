"""Writes the long-horizon golden fixtures of the GPU parity tests by running the CPU FP64 oracle
(oracle/libsfo.so, OpenMP) on the seeded workloads of cusuperhelium.workloads.

    python tests/golden/make_golden.py [c1] [c2] [c3] [c3long] [c5] [gl2] [mu]

  c1      N = 256 water surface y = 0.1 cos x, dt = 0.01: (Z, phi) after 100 RK4 steps
  c2      N = 2048 water surface, dt = 0.002: (Z, phi) after 100, 500 and 1000 RK4 steps
  c3      all 512 surfaces of the uneven ensemble: (Z, phi) after 20 steps
  c3long  two surfaces of each N in {128, 256, 512, 1024}: (Z, phi) after 100 and 500 steps
  c5      N = 4096 helium film (h0 = 1, image term, g_eff = 0.15): (Z, phi) after 2, 10 and 20 steps
  gl2     N = 32 helium film of depth 1 (the SI film L = 2 pi 1e-7 m, depth 1e-7 m): two-stage
          Gauss-Legendre frames [x, y, phi] of steps h = 0.05 to t = 0.2 (Newton tol 1e-12)

Each fixture is an uncompressed .npz of float64 / complex128 arrays (no pickles) beside this
script; `mu` adds to every fixture the oracle's dipole strengths mu of the right-hand side at each
stored state (BaseBoundaryIntegrator.cuh:275-306), where they are missing.  The oracle runs one
surface at a time; every surface evolves independently, so a fixture of a sub-ensemble holds
exactly the trajectories of the full one.
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "superfluid-dynamics_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

from cusuperhelium import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def _checkpoints(props, Z, P, dt, marks):
    """States after each step count in `marks` (increasing)."""
    out, done = {}, 0
    for m in marks:
        Z, P = O.rk4(props, Z, P, dt, m - done)
        done = m
        out[m] = (Z.copy(), P.copy())
    return out


def make_c1():
    Z0, P0, dt, steps = W.c1()
    Z, P = O.rk4(O.Props(), Z0, P0, dt, steps)
    np.savez(os.path.join(HERE, "c1_water_n256.npz"), steps=np.array([steps]), Z=Z[None], phi=P[None])


def make_c2():
    Z0, P0 = W.c2()
    marks = (100, 500, 1000)
    cp = _checkpoints(O.Props(), Z0, P0, W.DT, marks)
    np.savez(os.path.join(HERE, "c2_water_n2048.npz"), steps=np.array(marks),
             Z=np.array([cp[m][0] for m in marks]), phi=np.array([cp[m][1] for m in marks]))


def make_c3():
    sizes, Zs = W.c3()
    Zf, Pf = [], []
    for n, Z0 in zip(sizes, Zs):
        Z, P = O.rk4(O.Props(), Z0, np.zeros(n), W.DT, 20)
        Zf.append(Z)
        Pf.append(P)
    np.savez(os.path.join(HERE, "c3_ensemble_512_20steps.npz"), sizes=np.array(sizes), steps=np.array(20),
             Z=np.concatenate(Zf), phi=np.concatenate(Pf))


def make_c3long():
    sizes, Zs = W.c3()
    idx = W.c3_sample(2)
    marks = (100, 500)
    Zo = {m: [] for m in marks}
    Po = {m: [] for m in marks}
    for i in idx:
        cp = _checkpoints(O.Props(), Zs[i], np.zeros(sizes[i]), W.DT, marks)
        for m in marks:
            Zo[m].append(cp[m][0])
            Po[m].append(cp[m][1])
    np.savez(os.path.join(HERE, "c3_sample_500steps.npz"), index=np.array(idx),
             sizes=np.array([sizes[i] for i in idx]), steps=np.array(marks),
             Z=np.array([np.concatenate(Zo[m]) for m in marks]),
             phi=np.array([np.concatenate(Po[m]) for m in marks]))


def make_c5():
    Z0, P0 = W.c5()
    marks = (2, 10, 20)
    cp = _checkpoints(O.Props(**W.C5_PROPS), Z0, P0, W.DT, marks)
    np.savez(os.path.join(HERE, "c5_film_n4096.npz"), steps=np.array(marks),
             Z=np.array([cp[m][0] for m in marks]), phi=np.array([cp[m][1] for m in marks]))


def make_gl2():
    n = 32
    rng = np.random.default_rng(2)
    x = 2 * np.pi * np.arange(n) / n
    y = sum(rng.uniform(0, 0.03 / k) * np.cos(k * x + rng.uniform(0, 2 * np.pi)) for k in range(1, 6))
    s0 = np.concatenate([x, y, 0.01 * np.sin(x + 0.4)])
    d = O.adimensionalize(2 * np.pi * 1e-7, 150.0, 0.0, 1e-7)
    traj, times = O.gl2(O.Props(kind=O.HELIUM, depth=d["depth"], rho=d["rho"]), s0, 0.0, 0.2, 0.05, tol=1e-12)
    np.savez(os.path.join(HERE, "gl2_helium_n32.npz"), s0=s0, depth=np.array(d["depth"]), frames=traj, times=times)


FIXTURES = {"c1_water_n256.npz": None, "c2_water_n2048.npz": None, "c3_ensemble_512_20steps.npz": None,
            "c3_sample_500steps.npz": None, "c5_film_n4096.npz": W.C5_PROPS}


def add_mu():
    """mu at every stored state of every fixture that lacks it (one oracle RHS per surface)."""
    for name, kw in FIXTURES.items():
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            continue
        with np.load(path) as f:
            d = {k: f[k] for k in f.files}
        if "mu" in d:
            continue
        props = O.Props(**kw) if kw else O.Props()
        Zs = d["Z"] if d["Z"].ndim == 2 else d["Z"][None]
        Ps = d["phi"] if d["phi"].ndim == 2 else d["phi"][None]
        sizes = [int(v) for v in d["sizes"]] if "sizes" in d else [Zs.shape[1]]
        mus = []
        for Z, P in zip(Zs, Ps):
            mu, o = [], 0
            for n in sizes:
                _, _, diag = O.rhs(props, Z[o:o + n], P[o:o + n], diagnostics=True)
                mu.append(diag["mu"])
                o += n
            mus.append(np.concatenate(mu))
        d["mu"] = np.array(mus) if d["Z"].ndim == 2 else mus[0]
        np.savez(path, **d)


MAKERS = {"c1": make_c1, "c2": make_c2, "c3": make_c3, "c3long": make_c3long, "c5": make_c5, "gl2": make_gl2,
          "mu": add_mu}

if __name__ == "__main__":
    O.build()
    for name in sys.argv[1:] or [k for k in MAKERS if k != "mu"] + ["mu"]:
        t0 = time.perf_counter()
        MAKERS[name]()
        print(f"{name}: {time.perf_counter() - t0:.1f} s on {O.num_threads()} threads", flush=True)
