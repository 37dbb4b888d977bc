"""Host-side check of the sliced-ELL (SELL-64) address arithmetic.

Restates, in numpy, the index computations of k_sell_width / k_sell_fill /
k_apply_sell (openfpm_numerics_amd/csrc/dcpse_sell.hip) and checks on ragged
shapes (particle counts that are not multiples of 64, empty rows, one long
row) that
  * the fill writes every slot of the layout exactly once (dead lanes of the
    last slice included), so no slot the kernel can read is uninitialised;
  * every 16-byte load the apply kernel issues (including the clamped re-reads
    of group 0 and the groups past the slice width in the unrolled loop) stays
    inside its own slice.
"""
import numpy as np
import pytest


def layout(lengths):
    n = len(lengths)
    ns = (n + 63) // 64
    L = np.zeros(ns * 64, np.int64)
    L[:n] = lengths
    width = ((L.reshape(ns, 64).max(axis=1) + 3) // 4) * 4
    slice_ptr = np.concatenate([[0], np.cumsum(64 * width)])
    return L, width, slice_ptr


def fill_targets(L, width, slice_ptr):
    hits = np.zeros(slice_ptr[-1], np.int64)
    for row in range(L.size):
        s, l = divmod(row, 64)
        for j in range(width[s]):
            hits[slice_ptr[s] + (j >> 2) * 256 + 4 * l + (j & 3)] += 1
    return hits


def kernel_reads(L, width, slice_ptr, n_rows, unr):
    """(lo, hi) of every 16-byte load, per slice."""
    out = []
    for s in range(width.size):
        ng = width[s] // 4
        if ng == 0:
            continue
        base = slice_ptr[s]
        for lane in range(64):
            row = 64 * s + lane
            ln = L[row] if row < n_rows else 0
            gs = list(range(0, unr))  # prologue
            for g0 in range(0, ng, 2 * unr):
                gs += [g0 + unr + u for u in range(unr)] + [g0 + 2 * unr + u for u in range(unr)]
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
@pytest.mark.parametrize("unr", [2, 4])
def test_sell_fill_covers_and_reads_stay_in_slice(name, unr):
    rng = np.random.default_rng(0)
    lengths = SHAPES[name](rng)
    L, width, slice_ptr = layout(lengths)
    hits = fill_targets(L, width, slice_ptr)
    assert np.all(hits == 1), "fill must write every slot exactly once"
    for s, lo, hi in kernel_reads(L, width, slice_ptr, lengths.size, unr):
        assert slice_ptr[s] <= lo and hi < slice_ptr[s + 1], (s, lo, hi)
