"""The diagonal-block elimination of the LU panel kernel (eliminate_tiles in csrc/lu.hip), emulated on
the CPU from the kernel's rules.

A block of T x T tiles of 4 x 4 (T = 16 in the kernel) is eliminated without interchanges: thread
(i, j) owns tile (i, j).  Block step kk has two barrier-separated phases: the tiles of block column
kk form L_ik = A_ik U_kk^{-1} and those of block row kk form U_kj = L_kk^{-1} A_kj from the published
diagonal tile and publish them; then every trailing tile applies A_ij -= L_ik U_kj and the thread of
tile (kk + 1, kk + 1) factors it and publishes it for the next step.  Checks:
- the factors equal those of the unblocked right-looking elimination (getf2 without interchanges,
  multipliers by the reciprocal pivot) bit for bit, with every multiply-add fused and rounded once
  (exact rational arithmetic, rounded to the nearest double);
- no phase both reads and writes one of the shared arrays (the barriers order every exchange);
- the multiplier check flags a block whose elimination needs an interchange, and rows from kb on
  (padded with identity rows) neither change the factors of the kb x kb corner nor raise the flag.
"""
from fractions import Fraction

import numpy as np
import pytest


def fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def getf2(A, kb):
    """Unblocked right-looking elimination without interchanges; the flag of a multiplier > 1."""
    A = [list(map(float, row)) for row in A]
    n, bad = len(A), False
    for p in range(n):
        inv = 1.0 / A[p][p]
        for r in range(p + 1, n):
            l = A[r][p] * inv
            bad |= r < kb and not abs(l) <= 1.0
            A[r][p] = l
            for c in range(p + 1, n):
                A[r][c] = fma(-l, A[p][c], A[r][c])
    return np.array(A), bad


def tiles(A, kb, T):
    """The kernel's schedule (tile rules, shared arrays, phases) on T x T tiles."""
    n = 4 * T
    tile = {}
    for ti in range(T):
        for tj in range(T):
            tile[ti, tj] = [[float(A[4 * ti + r][4 * tj + c]) if 4 * ti + r < kb else float(4 * ti + r == 4 * tj + c)
                             for c in range(4)] for r in range(4)]
    shared = {"Dk": None, "Dinv": None, "Lp": {}, "Up": {}}
    bad = [False]
    phase = {"read": set(), "write": set()}

    def rd(name):
        phase["read"].add(name)
        return shared[name]

    def wr(name, key, val):
        phase["write"].add(name)
        if key is None:
            shared[name] = val
        else:
            shared[name][key] = val

    def barrier():
        assert not (phase["read"] & phase["write"]), f"one phase reads and writes {phase['read'] & phase['write']}"
        phase["read"].clear()
        phase["write"].clear()

    def factor_diag(k):
        a = tile[k, k]
        inv = [0.0] * 4
        for p in range(4):
            inv[p] = 1.0 / a[p][p]
            for r in range(p + 1, 4):
                l = a[r][p] * inv[p]
                bad[0] |= 4 * k + r < kb and not abs(l) <= 1.0
                a[r][p] = l
                for c in range(p + 1, 4):
                    a[r][c] = fma(-l, a[p][c], a[r][c])
        wr("Dk", None, [row[:] for row in a])
        wr("Dinv", None, inv[:])

    factor_diag(0)
    barrier()
    for kk in range(T):
        for ti in range(T):
            for tj in range(T):
                a = tile[ti, tj]
                if tj == kk and ti > kk:      # L_ik = A_ik U_kk^{-1}, column by column
                    U, inv = rd("Dk"), rd("Dinv")
                    for c in range(4):
                        for r in range(4):
                            x = a[r][c]
                            for p in range(c):
                                x = fma(-a[r][p], U[p][c], x)
                            l = x * inv[c]
                            bad[0] |= 4 * ti + r < kb and not abs(l) <= 1.0
                            a[r][c] = l
                    wr("Lp", ti, [row[:] for row in a])
                elif ti == kk and tj > kk:    # U_kj = L_kk^{-1} A_kj, step by step
                    L = rd("Dk")
                    for p in range(3):
                        for r in range(p + 1, 4):
                            for c in range(4):
                                a[r][c] = fma(-L[r][p], a[p][c], a[r][c])
                    wr("Up", tj, [row[:] for row in a])
        barrier()
        for ti in range(T):
            for tj in range(T):
                if ti > kk and tj > kk:       # A_ij -= L_ik U_kj, then the next diagonal tile
                    L, U = rd("Lp")[ti], rd("Up")[tj]
                    a = tile[ti, tj]
                    for p in range(4):
                        for r in range(4):
                            for c in range(4):
                                a[r][c] = fma(-L[r][p], U[p][c], a[r][c])
                    if ti == kk + 1 and tj == kk + 1:
                        factor_diag(kk + 1)
        barrier()
    out = np.zeros((n, n))
    for (ti, tj), blk in tile.items():
        out[4 * ti:4 * ti + 4, 4 * tj:4 * tj + 4] = blk
    return out, bad[0]


def _dominant(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)) * 0.3 / n + np.eye(n) * 0.5


@pytest.mark.parametrize("T,kb", [(5, 20), (6, 19), (4, 13)])
def test_tiles_equal_unblocked_getf2_bitwise(T, kb):
    n = 4 * T
    A = _dominant(n, 1000 + T)
    got, bad = tiles(A, kb, T)
    Ap = np.array([[A[r][c] if r < kb else float(r == c) for c in range(n)] for r in range(n)])
    ref, bad_ref = getf2(Ap, kb)
    assert np.array_equal(got, ref)
    assert not bad and not bad_ref
    # the kb x kb corner is the factorisation of the kb x kb block alone
    corner, _ = getf2(A[:kb, :kb], kb)
    assert np.array_equal(got[:kb, :kb], corner)


def test_tiles_flag_a_block_that_needs_interchanges():
    T = 4
    n = 4 * T
    A = _dominant(n, 7)
    A[9, 2] = 3.0   # |l_92| > 1 once column 2 is eliminated
    got, bad = tiles(A, n, T)
    ref, bad_ref = getf2(A, n)
    assert bad and bad_ref and np.array_equal(got, ref)
