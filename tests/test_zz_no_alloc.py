"""The no-allocation slice-structure entry point, dcpse_sell_count_noalloc.

It must give exactly what dcpse_sell_count gives (integer work: bit-exact) on
two calls in a row, the second with a smaller input, so the static scratch it
shares across calls holds nothing that leaks from one call into the next.
Both are called through the C ABI.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _csr(n_rows, seed, dev):
    rng = np.random.default_rng(seed)
    lengths = rng.integers(0, 200, size=n_rows)
    lengths[rng.random(n_rows) < 0.05] = 0  # empty rows
    row_ptr = np.zeros(n_rows + 1, np.int64)
    np.cumsum(lengths, out=row_ptr[1:])
    return torch.as_tensor(row_ptr, device=dev)


def _count(ctx, name, row_ptr, n_rows):
    from openfpm_numerics_amd import _lib

    n_slices = (n_rows + 63) // 64
    dev = row_ptr.device
    slice_ptr = torch.full((n_slices + 1,), -7, dtype=torch.int64, device=dev)
    sell_row = torch.full((max(64 * n_slices, 64),), -7, dtype=torch.int32, device=dev)
    n_slots = np.zeros(1, np.int64)
    ctx.call(name, n_rows, _lib.ptr(row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row), n_slots.ctypes.data)
    torch.cuda.synchronize()
    return slice_ptr.cpu().numpy(), sell_row.cpu().numpy(), int(n_slots[0])


@pytest.mark.parametrize("sizes", [(200_003, 70_001), (4096, 1)])
def test_noalloc_sell_count_matches_allocating_entry_point(cuda, sizes):
    from openfpm_numerics_amd import _lib

    ctx = _lib.context(cuda.index or 0)
    for call, n_rows in enumerate(sizes):
        print(f"[noalloc] call {call} n_rows={n_rows}", flush=True)
        row_ptr = _csr(n_rows, 100 + call, cuda)
        want = _count(ctx, "dcpse_sell_count", row_ptr, n_rows)
        got = _count(ctx, "dcpse_sell_count_noalloc", row_ptr, n_rows)
        assert got[2] == want[2]
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        # the slice widths follow from the row lengths alone
        lens = np.diff(row_ptr.cpu().numpy())
        n_slices = (n_rows + 63) // 64
        padded = np.full(64 * n_slices, 0, np.int64)
        padded[:n_rows] = lens
        width = (padded.reshape(n_slices, 64).max(1) + 3) // 4 * 4
        assert got[2] == int(64 * width.sum())


def test_noalloc_sell_count_rejects_oversize(cuda):
    """Rows beyond DCPSE_NOALLOC_MAX_ROWS are refused on the host, before any
    device work."""
    from openfpm_numerics_amd import _lib

    ctx = _lib.context(cuda.index or 0)
    row_ptr = torch.zeros(2, dtype=torch.int64, device=cuda)
    slice_ptr = torch.zeros(2, dtype=torch.int64, device=cuda)
    sell_row = torch.zeros(64, dtype=torch.int32, device=cuda)
    n_slots = np.zeros(1, np.int64)
    ctx.bind_current_stream()
    st = ctx.lib.dcpse_sell_count_noalloc(ctx.h, 33554432 + 1, _lib.ptr(row_ptr), _lib.ptr(slice_ptr),
                                          _lib.ptr(sell_row), n_slots.ctypes.data)
    assert st == _lib.DCPSE_ERR_RANGE
    assert b"DCPSE_NOALLOC_MAX_ROWS" in ctx.lib.dcpse_last_error(ctx.h)
