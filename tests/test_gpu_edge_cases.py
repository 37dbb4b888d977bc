"""Edge cases of the support, assembly and apply paths: empty particle sets
(the reference's loops simply run zero times, e.g. a process whose
subdomain holds no particles), an empty domain of two sets, a domain whose
support set is empty (every
row has too few neighbours: the reference's Vandermonde.hpp:109-112 error),
rows of several hundred neighbours (many LDS chunks in the assembly and a
wide sliced-ELL slice in the apply), and rows above the documented limit of
DCPSE_MAX_ROW neighbours (a range error from the support builder).

Tolerances: FP64 weights per row <= 1e-10 (or 16 cond(A) u, _dcpse_tol) and
FP64 applies <= 1e-10 of max|ref| against the oracle; FP32 applies <= 1e-5
relative L2 against the FP64 oracle on FP32-rounded input.
"""
import numpy as np
import pytest
import torch
from _dcpse_tol import check_weights

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_empty_particle_set(dim, cuda):
    from openfpm_numerics_amd import Derivative_x, Laplacian, ParticleSet

    ps = ParticleSet(np.zeros((0, dim)), [0.0] * dim, [1.0] * dim, bc=[1] * dim, device=cuda)
    vl = ps.getVerlet(0.1)
    assert vl.nnz == 0 and vl.row_ptr.tolist() == [0]
    lap = Laplacian(ps, vl, 2, 0.1)
    dx = Derivative_x(ps, vl, 2, 0.1)
    for dt in (torch.float32, torch.float64):
        f = torch.zeros(0, dtype=dt, device=cuda)
        assert lap(f).shape == (0,) and dx(f).shape == (0,)
        assert dx.dcpse.applyTranspose(f).shape == (0,)
    torch.cuda.synchronize()


def test_domain_without_support_particles(cuda):
    """Two sets, the support set empty: the support builder gives empty rows
    and the assembly stops with the reference's too-few-neighbours error."""
    from openfpm_numerics_amd import Derivative_x, ParticleSet
    from openfpm_numerics_amd.particles import VerletList

    sup = ParticleSet(np.zeros((0, 2)), [0.0, 0.0], [1.0, 1.0], bc=[0, 0], device=cuda)
    dom = ParticleSet(np.random.default_rng(1).random((5, 2)), [0.0, 0.0], [1.0, 1.0], bc=[0, 0], device=cuda)
    vl = VerletList.build(sup, 0.2, dom)
    assert vl.nnz == 0 and vl.row_ptr.tolist() == [0] * 6
    with pytest.raises(ValueError):
        Derivative_x(sup, vl, 2, 0.2, particlesDomain=dom)


def test_empty_domain_of_two_sets(cuda):
    """A non-empty support set with an empty domain (a separate set, and an
    empty subset): no rows, nothing assembled, empty results; the C ABI
    refuses a NULL domain (meaning the support set itself) whose row count
    differs from the support set's, before any device work."""
    from openfpm_numerics_amd import DcpseError, Derivative_x, ParticleSet, ParticleSubset, _lib
    from openfpm_numerics_amd.particles import VerletList

    pos = np.random.default_rng(2).random((400, 2))
    sup = ParticleSet(pos, [0.0, 0.0], [1.0, 1.0], bc=[1, 1], device=cuda)
    f = torch.as_tensor(np.sin(2 * np.pi * pos[:, 0]), device=cuda)
    dom = ParticleSet(np.zeros((0, 2)), [0.0, 0.0], [1.0, 1.0], bc=[1, 1], device=cuda)
    vl = VerletList.build(sup, 0.2, dom)
    assert vl.n_rows == 0 and vl.nnz == 0
    dx = Derivative_x(sup, vl, 2, 0.2, particlesDomain=dom)
    assert dx(f, f_domain=torch.zeros(0, dtype=f.dtype, device=cuda)).shape == (0,)
    sub = ParticleSubset(sup, np.zeros(0, dtype=np.int32))
    vls = sub.getVerlet(0.2)
    assert vls.n_rows == 0 and vls.nnz == 0
    dxs = Derivative_x(sup, vls, 2, 0.2, particlesDomain=sub)
    assert dxs(f).shape == (0,)
    ctx = _lib.context(0)
    rc = torch.tensor([0.2], dtype=torch.float64, device=cuda)
    row_ptr = torch.empty(1, dtype=torch.int64, device=cuda)
    nnz = np.zeros(1, np.int64)
    lo = np.zeros(2)
    hi = np.ones(2)
    with pytest.raises(DcpseError, match="n_dom == n_sup"):
        ctx.call("dcpse_support_count", 2, _lib.ptr(sup.pos_soa), 400, None, 0, _lib.ptr(rc), 0, lo.ctypes.data,
                 hi.ctypes.data, 3, _lib.ptr(row_ptr), nnz.ctypes.data)


def test_long_rows_match_oracle(oracle, cuda):
    """3D periodic lattice with rCut = 5.5h: about 700 neighbours per row."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, ParticleSet, dcpse as D

    n = 20
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")], 1)
    pos = pos + np.random.default_rng(4).uniform(-0.1 * h, 0.1 * h, pos.shape)
    rc = 5.5 * h
    lo, hi = [0.0] * 3, [1.0] * 3
    ps = ParticleSet(pos, lo, hi, bc=[1, 1, 1], device=cuda)
    vl = ps.getVerlet(rc)
    rp, col = oracle.support(pos, rc, lo, hi, 0b111)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp) and np.array_equal(vl.col.cpu().numpy(), col)
    lens = np.diff(rp)
    print(f"[long rows] N={len(lens)} neighbours {lens.min()}..{lens.max()}", flush=True)
    assert lens.min() > 600 and lens.max() <= 1024
    lap = Laplacian(ps, vl, 2, rc)
    dx = Derivative_x(ps, vl, 2, rc)
    refs = [oracle.assemble(pos, rp, col, b, 2, lo, hi, 0b111, dense_ev=False)
            for b in ((2, 0, 0), (0, 2, 0), (0, 0, 2))]
    ref_dx = oracle.assemble(pos, rp, col, (1, 0, 0), 2, lo, hi, 0b111, dense_ev=False)
    rel, n_cond, worst = check_weights(lap.dcpse.calcKernels.cpu().numpy(), sum(r.K for r in refs), rp, pos, pos,
                                       col, oracle.basis((2, 0, 0), 2), period=[1.0, 1.0, 1.0], label="lap")
    rel_dx, _, worst_dx = check_weights(dx.dcpse.calcKernels.cpu().numpy(), ref_dx.K, rp, pos, pos, col,
                                        oracle.basis((1, 0, 0), 2), period=[1.0, 1.0, 1.0], label="dx")
    print(f"[long rows] weights: Laplacian {rel:.2e}, Dx {rel_dx:.2e}", flush=True)
    f64 = np.sin(2 * np.pi * pos[:, 0]) * np.cos(2 * np.pi * pos[:, 2])
    f32 = f64.astype(np.float32)
    want64 = [sum(oracle.apply(r, rp, col, f64) for r in refs), oracle.apply(ref_dx, rp, col, f64)]
    want32 = [sum(oracle.apply(r, rp, col, f32.astype(np.float64)) for r in refs),
              oracle.apply(ref_dx, rp, col, f32.astype(np.float64))]
    bar64 = max(1e-10, 16 * max(worst, worst_dx) * 2.0 ** -53)
    for i in range(D.SELL_AFTER + 1):  # the CSR kernel, then the sliced-ELL copies
        for op, w64, w32 in zip((lap, dx), want64, want32):
            g64 = op(torch.as_tensor(f64, device=cuda)).cpu().numpy()
            assert np.abs(g64 - w64).max() <= bar64 * np.abs(w64).max(), i
            g32 = op(torch.as_tensor(f32, device=cuda)).cpu().numpy().astype(np.float64)
            assert np.linalg.norm(g32 - w32) / np.linalg.norm(w32) <= 1e-5, i


def test_rows_above_the_limit_raise(cuda):
    from openfpm_numerics_amd import DcpseError, ParticleSet

    n = 24
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")], 1)
    ps = ParticleSet(pos, [0.0] * 3, [1.0] * 3, bc=[1, 1, 1], device=cuda)
    with pytest.raises(DcpseError, match="DCPSE_MAX_ROW"):
        ps.getVerlet(7.0 * h)  # about 1437 neighbours per row
    vl = ps.getVerlet(2.0 * h)  # the context stays usable
    assert vl.nnz > 0
