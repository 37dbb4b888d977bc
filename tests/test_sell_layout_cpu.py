"""Host-side check of the sliced-ELL (SELL-64) address arithmetic.

Restates, in numpy, the index computations of k_sell_width / k_sell_fill /
k_apply_sell (openfpm_numerics_amd/csrc/dcpse_sell.hip) and checks on ragged
shapes (particle counts that are not multiples of 64, empty rows, one long
row) that
  * the lane -> row order (rows by decreasing length inside a slice, ties by
    index, -1 for lanes without a row) is a permutation of the slice's rows;
  * the fill writes every slot of the layout exactly once (dead lanes of the
    last slice included), so no slot the kernel can read is uninitialised;
  * every 16-byte load the apply kernel issues (including the clamped re-reads
    of group 0 and the groups past the slice width in the unrolled loop) stays
    inside its own slice.
"""
import numpy as np
import pytest


def layout(lengths):
    """(L, width, slice_ptr, sell_row) as k_sell_width computes them; L[i] is the
    length of the row served by lane i (0 for lanes without a row)."""
    n = len(lengths)
    ns = (n + 63) // 64
    key = np.full(ns * 64, -1, np.int64)
    key[:n] = lengths
    width = ((np.maximum(key.reshape(ns, 64).max(axis=1), 0) + 3) // 4) * 4
    slice_ptr = np.concatenate([[0], np.cumsum(64 * width)])
    sell_row = np.full(ns * 64, -1, np.int64)
    for s in range(ns):
        k = key[64 * s: 64 * s + 64]
        for lane in range(64):
            rank = int(np.sum((k > k[lane]) | ((k == k[lane]) & (np.arange(64) < lane))))
            sell_row[64 * s + rank] = 64 * s + lane if 64 * s + lane < n else -1
    L = np.where(sell_row >= 0, key[np.maximum(sell_row, 0)], 0)
    return L, width, slice_ptr, sell_row


def fill_targets(L, width, slice_ptr):
    hits = np.zeros(slice_ptr[-1], np.int64)
    for i in range(L.size):
        s, l = divmod(i, 64)
        for j in range(width[s]):
            hits[slice_ptr[s] + (j >> 2) * 256 + 4 * l + (j & 3)] += 1
    return hits


def groups_loaded(ng, unr, depth):
    """Slot groups whose col/W loads k_apply_sell (depth 2) or k_apply_sell3
    (depth 3) issue for a slice of ng groups, before clamping."""
    b = lambda g0: [g0 + u for u in range(unr)]
    if depth == 2:
        gs = b(0)
        for g0 in range(0, ng, 2 * unr):
            gs += b(g0 + unr) + b(g0 + 2 * unr)
        return gs
    gs = b(0) + b(unr)
    for g0 in range(0, ng, 2 * unr):
        gs += b(g0 + 2 * unr) + b(g0 + unr) + b(g0 + 3 * unr) + b(g0 + 2 * unr)
    return gs


def kernel_reads(L, width, slice_ptr, unr, depth=2):
    """(lo, hi) of every 16-byte load, per slice."""
    out = []
    for s in range(width.size):
        ng = width[s] // 4
        if ng == 0:
            continue
        base = slice_ptr[s]
        for lane in range(64):
            ln = L[64 * s + lane]
            gs = groups_loaded(ng, unr, depth)
            for g in gs:
                off = g * 256 if 4 * g < ln else 0
                lo = base + 4 * lane + off
                out.append((s, lo, lo + 3))
    return out


SHAPES = {
    "ragged_100": lambda rng: rng.integers(3, 60, 100),
    "with_empty_rows": lambda rng: np.where(rng.random(200) < 0.2, 0, rng.integers(1, 40, 200)),
    "one_long_row": lambda rng: np.concatenate([rng.integers(20, 30, 130), [200], rng.integers(20, 30, 7)]),
    "exact_128": lambda rng: rng.integers(30, 70, 128),
    "all_empty_tail": lambda rng: np.concatenate([rng.integers(5, 9, 64), np.zeros(10, np.int64)]),
}


@pytest.mark.parametrize("name", sorted(SHAPES))
@pytest.mark.parametrize("unr,depth", [(2, 2), (4, 2), (1, 3), (2, 3)])
def test_sell_fill_covers_and_reads_stay_in_slice(name, unr, depth):
    rng = np.random.default_rng(0)
    lengths = SHAPES[name](rng)
    L, width, slice_ptr, sell_row = layout(lengths)
    n = lengths.size
    for s in range(width.size):
        rows = sell_row[64 * s: 64 * s + 64]
        live = rows[rows >= 0]
        assert sorted(live.tolist()) == list(range(64 * s, min(64 * s + 64, n)))
        assert np.all(rows[len(live):] == -1), "lanes without a row come last"
        assert np.all(np.diff(L[64 * s: 64 * s + 64]) <= 0), "lengths decrease with the lane"
        assert width[s] >= L[64 * s] and width[s] % 4 == 0
    hits = fill_targets(L, width, slice_ptr)
    assert np.all(hits == 1), "fill must write every slot exactly once"
    for s, lo, hi in kernel_reads(L, width, slice_ptr, unr, depth):
        assert slice_ptr[s] <= lo and hi < slice_ptr[s + 1], (s, lo, hi)


def test_padding_lines_are_not_touched():
    """With lanes ordered by decreasing length, a 128-byte line (8 lanes x 16 B)
    of a slot group is read only if one of its lanes still has slots there: the
    lines fetched per slice are within one line per group of the row data."""
    rng = np.random.default_rng(1)
    lengths = rng.integers(37, 71, 64 * 40)
    L, width, slice_ptr, _ = layout(lengths)
    fetched = stored = 0
    for s in range(width.size):
        Ls = L[64 * s: 64 * s + 64]
        for g in range(width[s] // 4):
            for line in range(8):
                stored += 1
                fetched += bool(np.any(Ls[8 * line: 8 * line + 8] > 4 * g))
    need = lengths.sum() * 4 / 128
    assert fetched <= need * 1.06 + width.size * 2, (fetched, need, stored)
    assert stored > need * 1.08  # the layout itself is padded


@pytest.mark.parametrize("n_ops", [1, 2, 4])
def test_bundle16_addressing(n_ops):
    """16-bit bundle: every word written once, every read inside the slice's blocks."""
    rng = np.random.default_rng(3)
    lengths = rng.integers(20, 60, 300)
    L, width, slice_ptr, _ = layout(lengths)
    blk = 128 + 256 * n_ops
    total = slice_ptr[-1] // 256 * blk
    assert total * 2 == slice_ptr[-1] + 2 * n_ops * slice_ptr[-1]  # (0.5 + n_ops) * n_slots words
    hits16 = np.zeros(2 * total, np.int64)  # int16 units of the column part
    hitsw = np.zeros(total, np.int64)
    for i in range(L.size):
        s, l = divmod(i, 64)
        for j in range(width[s]):
            gb = (slice_ptr[s] // 256 + (j >> 2)) * blk
            hits16[2 * gb + 4 * l + (j & 3)] += 1
            for o in range(n_ops):
                hitsw[gb + 128 + o * 256 + 4 * l + (j & 3)] += 1
    col_words = hits16.reshape(-1, 2).sum(axis=1)
    assert np.all((col_words == 2) | (hitsw == 1))
    assert np.all(col_words + 2 * hitsw == 2)  # each word is either 2 offsets or 1 weight
    for s in range(width.size):
        lo_w, hi_w = slice_ptr[s] // 256 * blk, slice_ptr[s + 1] // 256 * blk
        for lane in range(64):
            for g in range(width[s] // 4):
                cw = lo_w + g * blk + 2 * lane  # this lane's 8 bytes of offsets
                assert lo_w <= cw and cw + 1 < hi_w
                for o in range(n_ops):
                    ww = lo_w + g * blk + 128 + o * 256 + 4 * lane
                    assert ww + 3 < hi_w
