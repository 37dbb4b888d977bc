"""C ABI checks that need no GPU: the library loads, exports every entry point
include/dcpse_hip.h declares (and the ctypes table binds exactly those), the
host-only helpers agree with the oracle, and device entry points fail cleanly
without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "dcpse_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dcpse_[a-z0-9_]+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    from openfpm_numerics_amd import _lib

    return _lib.load()


def test_header_declares_the_expected_surface():
    names = header_functions()
    for must in ("dcpse_ctx_create", "dcpse_support_count", "dcpse_support_fill", "dcpse_assemble",
                 "dcpse_apply_f32", "dcpse_apply_f64", "dcpse_transpose_count", "dcpse_apply_transpose_f32",
                 "dcpse_apply_sell_f32"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    handle = C.CDLL(lib._name)
    missing = [n for n in header_functions() if not hasattr(handle, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from openfpm_numerics_amd import _lib

    assert sorted(_lib.SIGNATURES) == header_functions()


def test_host_basis_and_rhs_match_oracle(lib, oracle):
    for beta in [(1, 0), (0, 1), (2, 0), (1, 1), (2, 2), (0, 0), (1, 0, 0), (2, 0, 0), (0, 1, 1), (3, 0, 0)]:
        for order in (1, 2, 3):
            b = np.ascontiguousarray(beta, dtype=np.uint32)
            m = lib.dcpse_basis_size(b.size, b.ctypes.data, order)
            e = np.zeros(m * b.size, np.uint32)
            assert lib.dcpse_basis(b.size, b.ctypes.data, order, e.ctypes.data, m) == m
            assert np.array_equal(e.reshape(m, b.size), oracle.basis(beta, order))
            r = np.zeros(m)
            assert lib.dcpse_rhs(b.size, b.ctypes.data, order, r.ctypes.data, m) == m
            assert np.array_equal(r, oracle.rhs(beta, order))
            assert lib.dcpse_basis(b.size, b.ctypes.data, order, e.ctypes.data, 0) == -m


def test_bad_arguments_are_rejected_without_a_device(lib):
    b = np.zeros(4, np.uint32)
    assert lib.dcpse_basis_size(4, b.ctypes.data, 2) == -1  # dim must be 2 or 3
    assert lib.dcpse_version() >= 100
    # entry points check the context before touching the device
    assert lib.dcpse_sync(None) != 0
    assert lib.dcpse_apply_f32(None, 0, None, None, 1, None, None, 0, 1, None, 0, None, 0, 1.0, 0.0) != 0
    assert lib.dcpse_assemble(None, 3, None, 0, None, 0, None, None, 0, None, None, 1, None, None, 2, 0.9, 0,
                              None, None, None, None, None, None) != 0
    assert lib.dcpse_last_error(None) == b"null context"


def test_context_creation_without_gpu_fails_cleanly(lib):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = C.c_void_p()
    assert lib.dcpse_ctx_create(C.byref(h), 0) != 0
    assert not h.value


def test_python_layer_refuses_to_run_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from openfpm_numerics_amd import _lib

    with pytest.raises(_lib.DcpseError):
        _lib.Context(0)
