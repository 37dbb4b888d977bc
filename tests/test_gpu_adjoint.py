"""Adjoint (transpose-apply) kernel: against the oracle's assembled operator
transposed on the CPU, the inner-product identity, and central differences."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(cuda, n=64, seed=4, spread=False):
    from openfpm_numerics_amd import Laplacian, ParticleSet, workloads as W
    from openfpm_numerics_amd.particles import VerletList

    w = W.jittered_2d(n, seed=seed)
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=[1, 1], device=cuda)
    if spread:
        # per-particle cut-offs: an unsymmetric support pattern
        rc = w["rcut"] * (1.0 + 0.4 * np.random.default_rng(seed).random(ps.size_local()))
        vl = VerletList.build(ps, torch.as_tensor(rc, device=cuda))
    else:
        vl = ps.getVerlet(w["rcut"])
    return w, ps, vl, Laplacian(ps, vl, 2, w["rcut"])


@pytest.mark.parametrize("spread", [False, True], ids=["one_cutoff", "per_particle_cutoff"])
def test_transpose_matches_value_nz_transposed(spread, oracle, cuda):
    """A^T g on the GPU == (sum_d value_nz(D_d))^T g from the oracle (scipy),
    for a symmetric support pattern (one cut-off: the transpose reuses the
    pattern) and an unsymmetric one (per-particle cut-offs)."""
    w, ps, vl, lap = _setup(cuda, spread=spread)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    A = None
    for d in range(2):
        ref = oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(2)), 2,
                              w["box_lo"], w["box_hi"], w["periodic"])
        Ad = oracle.value_nz(ref, rp, col)
        A = Ad if A is None else A + Ad
    rng = np.random.default_rng(0)
    g = rng.standard_normal((3, ps.size_local()))
    want = (A.T @ g.T).T
    got = lap.applyTranspose(torch.as_tensor(g, device=cuda)).cpu().numpy()
    rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert rel <= 1e-10, f"FP64 transpose apply error {rel:.3e}"
    got32 = lap.applyTranspose(torch.as_tensor(g, dtype=torch.float32, device=cuda)).cpu().numpy()
    want32 = (A.T @ g.astype(np.float32).astype(np.float64).T).T
    rel32 = np.linalg.norm(got32 - want32) / np.linalg.norm(want32)
    assert rel32 <= 1e-5, f"FP32 transpose apply rel L2 {rel32:.3e}"
    # single fields: the CSR kernel, then the sliced-ELL copy (16-bit bundle
    # on the operator's own support when the pattern is symmetric, the
    # transposed 32-bit layout otherwise)
    from openfpm_numerics_amd import dcpse as D

    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 2
    try:
        one = [lap.applyTranspose(torch.as_tensor(g[1], dtype=torch.float32, device=cuda)).cpu().numpy()
               for _ in range(2)]
    finally:
        D.SELL_AFTER = sell_after
    assert (lap.dcpse._T.get("bundle16") is None) == spread
    for o in one:
        rel1 = np.linalg.norm(o - want32[1]) / np.linalg.norm(want32[1])
        assert rel1 <= 1e-5, f"FP32 single-field transpose apply rel L2 {rel1:.3e}"


def test_inner_product_identity(cuda):
    """<A f, g> == <f, A^T g> to FP64 rounding."""
    _, ps, _, lap = _setup(cuda, n=80, seed=9)
    gen = torch.Generator(device="cpu").manual_seed(3)
    f = torch.randn(ps.size_local(), generator=gen, dtype=torch.float64).to(cuda)
    g = torch.randn(ps.size_local(), generator=gen, dtype=torch.float64).to(cuda)
    lhs = torch.dot(lap(f), g)
    rhs = torch.dot(f, lap.applyTranspose(g))
    assert abs(float(lhs - rhs)) <= 1e-12 * float(torch.linalg.vector_norm(lap(f)) * torch.linalg.vector_norm(g))


def test_autograd_against_central_differences(cuda):
    """dL/df from the transpose kernel vs central differences of the FP64 forward.

    L = sum_b ||Lap f_b - g_b||^2 is quadratic in f, so a central difference is
    exact up to rounding; it is evaluated as sum_p (r+_p - r-_p)(r+_p + r-_p)
    so the cancellation is confined to the rows that change.
    """
    from openfpm_numerics_amd import dcpse_apply

    w, ps, vl, lap = _setup(cuda, n=64, seed=12)
    n = ps.size_local()
    gen = torch.Generator(device="cpu").manual_seed(7)
    B = 4
    f = torch.randn(B, n, generator=gen, dtype=torch.float64).to(cuda)
    tgt = torch.randn(B, n, generator=gen, dtype=torch.float64).to(cuda)
    f.requires_grad_(True)
    loss = ((dcpse_apply(lap, f) - tgt) ** 2).sum()
    loss.backward()
    grad = f.grad.detach()
    rng = np.random.default_rng(5)
    h = 1e-3
    worst = 0.0
    with torch.no_grad():
        for _ in range(32):
            b, q = int(rng.integers(B)), int(rng.integers(n))
            fp = f.detach().clone()
            fm = f.detach().clone()
            fp[b, q] += h
            fm[b, q] -= h
            rp = lap(fp[b]) - tgt[b]
            rm = lap(fm[b]) - tgt[b]
            fd = float(((rp - rm) * (rp + rm)).sum()) / (2 * h)
            worst = max(worst, abs(float(grad[b, q]) - fd) / max(abs(fd), 1e-30))
    assert worst <= 1e-6, f"max relative gradient error vs central differences {worst:.3e}"


@pytest.mark.parametrize("sig", [(1, 0), (2, 0), (1, 1), "lap"])
def test_sparse_export_matches_value_nz(sig, oracle, cuda):
    """export_csr == the oracle's value_nz triplets summed (scipy CSR), and
    export_csr @ f == the apply."""
    from openfpm_numerics_amd import Dcpse, workloads as W

    w, ps, vl, lap = _setup(cuda, n=40, seed=2)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    sigs = [(2, 0), (0, 2)] if sig == "lap" else [sig]
    A = None
    for s in sigs:
        Ad = oracle.value_nz(oracle.assemble(w["pos"], rp, col, s, 2, w["box_lo"], w["box_hi"], w["periodic"]),
                             rp, col)
        A = Ad if A is None else A + Ad
    A = A.tocsr()
    A.sort_indices()
    op = lap.dcpse if sig == "lap" else Dcpse(ps, vl, sig, 2, w["rcut"])
    erp, ecol, eval_ = (t.cpu().numpy() for t in op.export_csr(torch.float64))
    assert np.array_equal(erp, A.indptr) and np.array_equal(ecol, A.indices)
    scale = np.abs(A.data).max()
    assert np.max(np.abs(eval_ - A.data)) <= 1e-10 * scale
    f = torch.randn(ps.size_local(), dtype=torch.float64, generator=torch.Generator().manual_seed(0)).to(cuda)
    T = op.to_sparse_csr()
    assert T.shape == (ps.size_local(), ps.size_local())
    assert torch.equal(T.crow_indices().cpu(), torch.as_tensor(A.indptr, dtype=torch.int64))
    got = (A @ f.cpu().numpy())
    want = op.computeDifferentialOperator(f).cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))
    e32 = op.export_csr(torch.float32)[2].cpu().numpy().astype(np.float64)
    assert np.max(np.abs(e32 - A.data)) <= 1e-6 * scale


@pytest.mark.parametrize("nb", [16, 13, 6])
def test_batched_fp32_applies_equal_single_field(nb, cuda):
    """Batches of 4 or more FP32 fields on one set run on a particle-major
    copy of the fields (16, 8 or 4 per pass; the rest field-major): every
    field's forward and transposed apply equals the single-field CSR
    kernel's bit for bit (same lanes, same chunks, same order per field), for
    the default launch shape and, with 6 fields, every 16-byte-load shape."""
    from openfpm_numerics_amd import _lib
    from openfpm_numerics_amd.autotune import APPLY_SHAPES

    w, ps, vl, lap = _setup(cuda)
    d = lap.dcpse
    n = ps.size_local()
    f = torch.as_tensor(np.random.default_rng(nb).standard_normal((nb, n)), dtype=torch.float32, device=cuda)
    ctx = _lib.context(cuda.index or 0)
    shapes = [None] + ([s for s in APPLY_SHAPES if s[2] == 4] if nb == 6 else [])
    try:
        for shape in shapes:
            if shape is not None:
                ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, shape[0], shape[1], shape[2], 256), "config")
            _check_batched(ctx, d, vl, f, n, nb, cuda)
    finally:
        ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, 0, 4, 4, 256), "config reset")


def _check_batched(ctx, d, vl, f, n, nb, cuda):
    from openfpm_numerics_amd import _lib

    fwd = d.computeDifferentialOperator(f)
    adj = d.applyTranspose(f)
    W = d.weights(torch.float32)
    rowT, colT, WT, diag = d.transposed(torch.float32)
    for b in range(nb):
        fb = f[b].contiguous()
        o = torch.empty(n, dtype=torch.float32, device=cuda)
        ctx.call("dcpse_apply_f32", n, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), 1, _lib.ptr_array([W]),
                 _lib.ptr(fb), n, 1, None, n, _lib.ptr_array([o]), n, 1.0, 0.0)
        assert torch.equal(o, fwd[b]), (nb, b)
        ot = torch.empty(n, dtype=torch.float32, device=cuda)
        ctx.call("dcpse_apply_transpose_f32", n, _lib.ptr(rowT), _lib.ptr(colT), _lib.ptr(WT), _lib.ptr(diag),
                 _lib.ptr(fb), n, 1, _lib.ptr(ot), n, 1.0, 0.0)
        assert torch.equal(ot, adj[b]), (nb, b)
