"""The unordered-pair tile schedule of the batched kernel, emulated on the CPU (csrc/sfh_sym.hpp).

For every node count the batched path takes (N = 2^k or 3 * 2^k, N <= 1024) the schedule of
sym_sweep is replayed wave by wave, phase by phase and offset by offset:
- every unordered pair {k, j} of nodes (k != j, both < N) is visited exactly once, and the pair's
  column sum is taken exactly when its mirror flag is set (so each ordered pair contributes once);
- every element of the stored tile-skewed matrix (the direct entry at offset r, the mirrored one at
  offset (64 - r) mod 64 of the transposed tile, and the diagonal at offset 0) is written exactly once;
- within a phase no two waves add to the same LDS slot (column sums), and a node's column-sum slot
  after the last offset of a fragment is the column the rotated register then holds;
- the product's offset rows cover every tile element of a block row exactly once.
The geometry is restated here from the kernel's rules (block rows per wave, the phase ranges, the
split of each range among the waves of a block row); the GPU tests check the kernel's results.
"""
import numpy as np
import pytest

TILE, WAVES = 64, 16


def _sizes():
    out = []
    for k in range(0, 11):
        for m in (1, 3):
            n = m << k
            if n <= 1024:
                out.append(n)
    return sorted(set(out))


def _schedule(n):
    """Yields (phase, wave, sub, I, J, r, mirror) for every offset a wave runs."""
    nb = (n + TILE - 1) // TILE
    wpb = WAVES // nb
    for ph in range(nb // 2 + 1):
        for w in range(WAVES):
            I, sub = w // wpb, w % wpb
            if I >= nb:
                continue
            J = (I + ph) % nb
            if ph == 0:
                lo, hi = 1, 33
            elif 2 * ph < nb:
                lo, hi = 0, TILE
            else:
                lo = 1 if I < nb // 2 else 0
                hi = lo + 32
            r0 = lo + ((hi - lo) * sub) // wpb
            r1 = lo + ((hi - lo) * (sub + 1)) // wpb
            for r in range(r0, r1):
                yield ph, w, sub, I, J, r, not (ph == 0 and r == 32), r0, r1


@pytest.mark.parametrize("n", _sizes())
def test_every_unordered_pair_once_and_every_matrix_slot_once(n):
    nb = (n + TILE - 1) // TILE
    wpb = WAVES // nb
    lanes = np.arange(TILE)
    seen_row = np.zeros((n, n), np.int32)   # [k, j]: contribution of j to row k
    stored = np.zeros((nb, nb, TILE, TILE), np.int32)   # tile (I, J), offset r, lane l
    for I in range(nb):   # the diagonal entries, offset 0 of the diagonal tiles
        stored[I, I, 0, :] += 1
    col_slots = {}   # phase -> set of (sub, node) written
    frag_end = {}
    for ph, w, sub, I, J, r, mirror, r0, r1 in _schedule(n):
        k = I * TILE + lanes
        j = J * TILE + ((lanes + r) & (TILE - 1))
        ok = (k < n) & (j < n)
        assert not np.any(k[ok] == j[ok])
        seen_row[k[ok], j[ok]] += 1
        if mirror:
            seen_row[j[ok], k[ok]] += 1
        stored[I, J, r, lanes] += 1
        if mirror:
            stored[J, I, (TILE - r) & (TILE - 1), (lanes + r) & (TILE - 1)] += 1
        frag_end[(ph, w)] = (sub, J, r1)
    # each ordered pair k != j contributes to row k once; nothing on the diagonal
    off = ~np.eye(n, dtype=bool)
    assert np.all(seen_row[off] == 1) and np.all(np.diag(seen_row) == 0)
    assert np.all(stored == 1)
    # column-sum slots of a phase: one writer per (sub-wave slot, node)
    for (ph, w), (sub, J, r1) in frag_end.items():
        c = J * TILE + ((lanes + r1) & (TILE - 1))
        key = col_slots.setdefault(ph, set())
        for node in c:
            assert (sub, int(node)) not in key, f"n={n} phase {ph}: slot ({sub}, {node}) written twice"
            key.add((sub, int(node)))
    # the slots fit the accumulator arrays (kBT doubles each)
    assert wpb * nb * TILE <= 1024


@pytest.mark.parametrize("n", _sizes())
def test_stored_product_covers_every_tile_element_once(n):
    nb = (n + TILE - 1) // TILE
    wpb = WAVES // nb
    span = nb * TILE
    hits = np.zeros((nb, span, TILE), np.int32)   # block row I, offset row i, lane l
    for w in range(WAVES):
        I, sub = w // wpb, w % wpb
        if I >= nb:
            continue
        i0, i1 = (span * sub) // wpb, (span * (sub + 1)) // wpb
        hits[I, i0:i1, :] += 1
    assert np.all(hits == 1)
    # offset row i of block row I, lane l reads column (i & ~63) + ((l + i) & 63): every column once
    # per row
    lanes = np.arange(TILE)
    cols = np.zeros((TILE, span), np.int32)
    for i in range(span):
        cols[lanes, (i & ~(TILE - 1)) + ((lanes + i) & (TILE - 1))] += 1
    assert np.all(cols == 1)


def test_wave_rotate_direction_matches_column_bookkeeping():
    # at offset r lane l holds the column sum of column (l + r) % 64; after the rotate (lane l takes
    # lane l + 1's value) lane l holds column (l + 1 + r) % 64, the column of offset r + 1
    lanes = np.arange(TILE)
    held = (lanes + 5) % TILE
    rotated = held[(lanes + 1) % TILE]
    assert np.array_equal(rotated, (lanes + 6) % TILE)
