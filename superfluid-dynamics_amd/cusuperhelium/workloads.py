"""Seeded synthetic workloads of the stepper (pure numpy: no device, no library).

Every surface is periodic on [0, 2 pi) with equispaced nodes x_j = 2 pi j / N and Z_j = x_j + i y(x_j);
the velocity potential starts at zero.  The same generator feeds the benchmark, the GPU tests and
the scripts that write the golden fixtures, so all of them see identical bits.

  c1  N = 256, y = 0.1 cos x, water (g = 1, deep), dt = 0.01, 100 steps
  c2  N = 2048, y = sum_{k=1..8} a_k cos(k x + theta_k), a_k ~ U(0, 0.02/k), theta_k ~ U(0, 2 pi)
      (default_rng(SEED)), water, dt = 0.002, 1000 steps
  c3  512 surfaces, N_i uniform in {128, 256, 512, 1024} (default_rng(SEED)), surface i drawn as c2
      with seed SEED + i, water, dt = 0.002, 500 steps
  c4  N = 1024, y0 = sum_{k=1..16} c_k cos(k x), c_k ~ U(0, 0.01) (default_rng(SEED)), target drawn
      with SEED + 1, water, dt = 0.005
  c5  helium film N = 4096 over a substrate at depth h0 = 1 (image term), g_eff = 3 A / h0^4 = 0.15
      (A = 0.05), 8-mode surface with a_k ~ U(0, 0.05/k) (seed SEED + 5), dt = 0.002, 500 steps
"""
from __future__ import annotations

import numpy as np

SEED = 20261202
DT = 0.002


def nodes(n: int) -> np.ndarray:
    return 2 * np.pi * np.arange(n) / n


def surface(n: int, seed: int, amp: float = 0.02, modes: int = 8) -> np.ndarray:
    """Z of an 8-mode random surface: a_k ~ U(0, amp/k), theta_k ~ U(0, 2 pi)."""
    rng = np.random.default_rng(seed)
    x = nodes(n)
    a = rng.uniform(0, amp / np.arange(1, modes + 1))
    th = rng.uniform(0, 2 * np.pi, modes)
    return x + 1j * sum(a[k] * np.cos((k + 1) * x + th[k]) for k in range(modes))


def ensemble_sizes(count: int = 512, seed: int = SEED) -> list[int]:
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.choice([128, 256, 512, 1024], size=count)]


def c1():
    """(Z0, phi0, dt, steps) of the N = 256 surface y = 0.1 cos x."""
    x = nodes(256)
    return x + 1j * 0.1 * np.cos(x), np.zeros(256), 0.01, 100


def c2():
    """(Z0, phi0) of the N = 2048 water surface."""
    return surface(2048, SEED), np.zeros(2048)


def c3(count: int = 512):
    """(sizes, [Z0 per surface]) of the uneven ensemble."""
    sizes = ensemble_sizes(count)
    return sizes, [surface(n, SEED + i) for i, n in enumerate(sizes)]


def c3_sample(per_size: int = 2):
    """Indices of the first `per_size` surfaces of each N in the c3 ensemble."""
    sizes = ensemble_sizes()
    out = []
    for n in (128, 256, 512, 1024):
        out += [i for i, m in enumerate(sizes) if m == n][:per_size]
    return out


def c4(modes: int = 16, n: int = 1024):
    """(x, amplitudes, target amplitudes, cosine basis [modes, n])."""
    x = nodes(n)
    amp = np.random.default_rng(SEED).uniform(0, 0.01, modes)
    tgt = np.random.default_rng(SEED + 1).uniform(0, 0.01, modes)
    basis = np.array([np.cos((k + 1) * x) for k in range(modes)])
    return x, amp, tgt, basis


C5_PROPS = dict(kind=1, depth=1.0, g=0.15)


def c5():
    """(Z0, phi0) of the N = 4096 helium film; physics in C5_PROPS."""
    return surface(4096, SEED + 5, amp=0.05), np.zeros(4096)
