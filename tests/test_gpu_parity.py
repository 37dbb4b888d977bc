"""Parity of the HIP DC-PSE path against the CPU oracle, on seeded inputs.

Tolerances (stated per test):
  * support (neighbour CSR): bit-exact (integer work).
  * FP64 kernel weights: per row max|K_gpu - K_ref| / max|K_ref| <= 1e-10.
  * FP64 apply: max|D_gpu - D_ref| / max|D_ref| <= 1e-10.
  * FP32 apply vs the FP64 oracle applied to the FP32-rounded input:
    relative L2 <= 1e-5 and max-abs / max|ref| <= 1e-5.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ps(w, dev):
    from openfpm_numerics_amd import ParticleSet

    bc = [(w["periodic"] >> d) & 1 for d in range(len(w["box_lo"]))]
    return ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=bc, device=dev)


def _row_rel(Kg, Kr, row_ptr):
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    num = np.zeros(n)
    den = np.zeros(n)
    np.maximum.at(num, rows, np.abs(Kg - Kr))
    np.maximum.at(den, rows, np.abs(Kr))
    return float(np.max(num / np.maximum(den, 1e-300)))


WORKLOADS = {
    "2d_lattice_64": lambda W: W.lattice_2d_periodic(64),
    "2d_jitter_96": lambda W: W.jittered_2d(96, seed=7),
    "3d_holes_20": lambda W: W.jittered_3d_holes(20, seed=11),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_support_matches_oracle(name, oracle, cuda):
    from openfpm_numerics_amd import workloads as W

    print(f"[support] {name}", flush=True)
    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = oracle.support(w["pos"], w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(vl.col.cpu().numpy(), col)


def test_support_cutoff_spread_and_two_sets_match_oracle(oracle, cuda):
    """The cell walk with per-particle reach: cut-offs over a 4x range (3D,
    x not periodic, cells of rmax/4), and a 2D domain set distinct from the
    support set with per-domain-particle cut-offs, against the oracle and a
    numpy brute-force search over all pairs."""
    from openfpm_numerics_amd import ParticleSet, workloads as W
    from openfpm_numerics_amd.particles import VerletList

    w = W.adaptive_3d(20_000, seed=9)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = oracle.support(w["pos"], w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    print(f"[support] adaptive N={len(w['pos'])} nnz={col.size}", flush=True)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(vl.col.cpu().numpy(), col)

    rng = np.random.default_rng(3)
    sup = rng.random((2500, 2))
    dom = rng.random((900, 2))
    h = 1.0 / 50
    rc = h * (1.5 + 4.5 * rng.random(len(dom)))
    for periodic in (0b11, 0b00):
        bc = [(periodic >> d) & 1 for d in range(2)]
        s_set = ParticleSet(sup, [0.0, 0.0], [1.0, 1.0], bc=bc, device=cuda)
        d_set = ParticleSet(dom, [0.0, 0.0], [1.0, 1.0], bc=bc, device=cuda)
        vl2 = VerletList.build(s_set, torch.as_tensor(rc, device=cuda), domain=d_set)
        rp2, col2 = oracle.support(sup, rc, [0.0, 0.0], [1.0, 1.0], periodic, pos_dom=dom)
        assert np.array_equal(vl2.row_ptr.cpu().numpy(), rp2), periodic
        assert np.array_equal(vl2.col.cpu().numpy(), col2), periodic
        d = dom[:, None, :] - sup[None, :, :]
        for k in range(2):
            if periodic >> k & 1:
                d[..., k] -= np.round(d[..., k])
        d2 = (d ** 2).sum(-1)
        gp, gc = vl2.row_ptr.cpu().numpy(), vl2.col.cpu().numpy()
        for p in range(len(dom)):
            assert np.array_equal(gc[gp[p]:gp[p + 1]], np.nonzero(d2[p] < rc[p] ** 2)[0]), (periodic, p)


@pytest.mark.parametrize("dim", [2, 3])
def test_support_clustered_cells_match_oracle(dim, oracle, cuda):
    """One cut-off, one set, a uniform background plus a dense cluster: the
    cluster's cells hold several times the candidates one LDS copy takes, so
    the per-cell walk restages them in chunks per particle; mixed periodic and
    open dimensions, a periodic dimension only 3 cells wide."""
    from openfpm_numerics_amd import ParticleSet

    rng = np.random.default_rng(21 + dim)
    if dim == 3:
        hi = np.array([1.0, 1.0, 0.35])
        bg = rng.random((2400, 3)) * hi
        cl = 0.3 + 0.2 * rng.random((600, 3))
        cl[:, 2] = 0.1 + 0.1 * rng.random(600)
        rcut, periodic = 0.1, 0b101
    else:
        hi = np.array([1.0, 0.16])
        bg = rng.random((600, 2)) * hi
        cl = 0.4 + 0.1 * rng.random((500, 2))
        cl[:, 1] = 0.05 + 0.05 * rng.random(500)
        rcut, periodic = 0.05, 0b10
    pos = np.concatenate([bg, cl])
    lo = np.zeros(dim)
    bc = [(periodic >> d) & 1 for d in range(dim)]
    ps = ParticleSet(pos, lo, hi, bc=bc, device=cuda)
    vl = ps.getVerlet(rcut)
    rp, col = oracle.support(pos, rcut, lo, hi, periodic)
    print(f"[support] clustered dim={dim} N={len(pos)} nnz={col.size} max row={np.diff(rp).max()}", flush=True)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(vl.col.cpu().numpy(), col)


@pytest.mark.parametrize("dim,rcut_h", [(3, 3.9), (2, 6.0)])
def test_support_large_cutoff_matches_oracle(dim, rcut_h, oracle, cuda):
    """Cut-offs whose cells hold more candidates than one LDS copy (the count
    walks them in chunks, particles of a cell in two rounds of 64) and rows
    longer than 128 (sorted by the separate long-row pass)."""
    from openfpm_numerics_amd import workloads as W

    w = W.jittered_3d_holes(24, seed=5, rcut_h=rcut_h) if dim == 3 else W.jittered_2d(64, seed=5, rcut_h=rcut_h)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = oracle.support(w["pos"], w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    print(f"[support] dim={dim} rcut={rcut_h}h N={len(w['pos'])} nnz={col.size} max row={np.diff(rp).max()}",
          flush=True)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp)
    assert np.array_equal(vl.col.cpu().numpy(), col)


def test_support_open_box_and_tiny_sets_match_oracle(oracle, cuda):
    """Open (non-periodic) boxes with particles outside them (they fall in
    the edge cells), and sets of one and two particles."""
    from openfpm_numerics_amd import ParticleSet

    rng = np.random.default_rng(13)
    for dim, n, rcut in ((3, 6000, 0.12), (2, 3000, 0.05)):
        pos = rng.random((n, dim)) * 1.1 - 0.05  # about 5% of each side outside [0, 1)
        lo, hi = np.zeros(dim), np.ones(dim)
        ps = ParticleSet(pos, lo, hi, bc=[0] * dim, device=cuda)
        vl = ps.getVerlet(rcut)
        rp, col = oracle.support(pos, rcut, lo, hi, 0)
        assert np.array_equal(vl.row_ptr.cpu().numpy(), rp), dim
        assert np.array_equal(vl.col.cpu().numpy(), col), dim
    for n in (1, 2):
        for periodic in (0b11, 0b00):
            pos = np.array([[0.1, 0.1], [0.15, 0.12]])[:n]
            bc = [(periodic >> d) & 1 for d in range(2)]
            ps = ParticleSet(pos, [0.0, 0.0], [1.0, 1.0], bc=bc, device=cuda)
            vl = ps.getVerlet(0.2)
            rp, col = oracle.support(pos, 0.2, [0.0, 0.0], [1.0, 1.0], periodic)
            assert np.array_equal(vl.row_ptr.cpu().numpy(), rp), (n, periodic)
            assert np.array_equal(vl.col.cpu().numpy(), col), (n, periodic)


def test_support_fill_after_other_count_matches_oracle(oracle, cuda):
    """The fill reuses the cell list and take masks of the count just before it
    only when the arguments match: a count of another cut-off in between, and a
    second fill of the same rows, take the full distance path and agree."""
    from openfpm_numerics_amd import _lib, workloads as W

    w = W.jittered_3d_holes(20, seed=11)
    rp, col = oracle.support(w["pos"], w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    ctx = _lib.context(0)
    pos = torch.as_tensor(np.ascontiguousarray(w["pos"].T), device=cuda)
    n = pos.shape[1]
    lo = np.ascontiguousarray(w["box_lo"], dtype=np.float64)
    hi = np.ascontiguousarray(w["box_hi"], dtype=np.float64)

    def count(rc):
        rct = torch.tensor([rc], dtype=torch.float64, device=cuda)
        row_ptr = torch.empty(n + 1, dtype=torch.int64, device=cuda)
        nnz = np.zeros(1, np.int64)
        ctx.call("dcpse_support_count", 3, _lib.ptr(pos), n, None, n, _lib.ptr(rct), 0, lo.ctypes.data,
                 hi.ctypes.data, w["periodic"], _lib.ptr(row_ptr), nnz.ctypes.data)
        return row_ptr, int(nnz[0])

    def fill(rc, row_ptr, nnz):
        rct = torch.tensor([rc], dtype=torch.float64, device=cuda)
        c = torch.full((nnz,), -1, dtype=torch.int32, device=cuda)
        ctx.call("dcpse_support_fill", 3, _lib.ptr(pos), n, None, n, _lib.ptr(rct), 0, lo.ctypes.data,
                 hi.ctypes.data, w["periodic"], _lib.ptr(row_ptr), _lib.ptr(c))
        return c.cpu().numpy()

    row_ptr, nnz = count(w["rcut"])
    assert np.array_equal(row_ptr.cpu().numpy(), rp)
    rp_b, nnz_b = count(0.8 * w["rcut"])
    assert np.array_equal(fill(w["rcut"], row_ptr, nnz), col)
    assert np.array_equal(fill(w["rcut"], row_ptr, nnz), col)
    row_ptr, nnz = count(w["rcut"])
    assert np.array_equal(fill(w["rcut"], row_ptr, nnz), col)
    rp_s, col_s = oracle.support(w["pos"], 0.8 * w["rcut"], w["box_lo"], w["box_hi"], w["periodic"])
    assert np.array_equal(rp_b.cpu().numpy(), rp_s)


SIGS = {
    2: [(1, 0), (0, 1), (2, 0), (0, 2), (1, 1)],
    3: [(1, 0, 0), (0, 0, 1), (2, 0, 0), (0, 1, 1)],
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_assembly_matches_oracle(name, oracle, cuda):
    from openfpm_numerics_amd import Dcpse, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    dim = ps.dims
    for sig in SIGS[dim]:
        print(f"[assembly] {name} beta={sig}", flush=True)
        d = Dcpse(ps, vl, sig, 2, w["rcut"])
        ref = oracle.assemble(w["pos"], rp, col, sig, 2, w["box_lo"], w["box_hi"], w["periodic"])
        assert ref.info.max() == 0
        np.testing.assert_allclose(d.localEps.cpu().numpy(), ref.eps, rtol=1e-13)
        np.testing.assert_allclose(d.localEpsInvPow.cpu().numpy(), ref.eps_inv_pow, rtol=1e-12)
        err = _row_rel(d.calcKernels.cpu().numpy(), ref.K, rp)
        assert err <= 1e-10, f"beta={sig}: row-relative kernel error {err:.3e}"


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_laplacian_is_sum_of_directions(name, oracle, cuda):
    """One fused Laplacian operator == sum over d of the reference's dim operators."""
    from openfpm_numerics_amd import Laplacian, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dim = ps.dims
    refs = [oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(dim)), 2,
                            w["box_lo"], w["box_hi"], w["periodic"]) for d in range(dim)]
    Kref = sum(r.K for r in refs)
    err = _row_rel(lap.dcpse.calcKernels.cpu().numpy(), Kref, rp)
    assert err <= 1e-10, f"Laplacian kernel error {err:.3e}"
    f, _, _, lap_exact = W.sin_field(w["pos"])
    ref_apply = sum(oracle.apply(r, rp, col, f) for r in refs)
    got = lap(torch.as_tensor(f, device=cuda)).cpu().numpy()
    rel = np.max(np.abs(got - ref_apply)) / np.max(np.abs(ref_apply))
    assert rel <= 1e-10, f"FP64 Laplacian apply error {rel:.3e}"


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_apply_fp64_fp32(name, oracle, cuda):
    from openfpm_numerics_amd import Derivative_x, Derivative_y, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    f, dx_exact, dy_exact, _ = W.sin_field(w["pos"])
    for Op, sig in ((Derivative_x, (1, 0, 0)), (Derivative_y, (0, 1, 0))):
        sig = sig[: ps.dims]
        print(f"[apply] {name} {Op.__name__}", flush=True)
        op = Op(ps, vl, 2, w["rcut"])
        ref = oracle.assemble(w["pos"], rp, col, sig, 2, w["box_lo"], w["box_hi"], w["periodic"])
        want = oracle.apply(ref, rp, col, f)
        got64 = op(torch.as_tensor(f, device=cuda)).cpu().numpy()
        assert np.max(np.abs(got64 - want)) / np.max(np.abs(want)) <= 1e-10
        f32 = f.astype(np.float32)
        want32 = oracle.apply(ref, rp, col, f32.astype(np.float64))
        got32 = op(torch.as_tensor(f32, device=cuda)).cpu().numpy().astype(np.float64)
        l2 = np.linalg.norm(got32 - want32) / np.linalg.norm(want32)
        mx = np.max(np.abs(got32 - want32)) / np.max(np.abs(want32))
        assert l2 <= 1e-5 and mx <= 1e-5, f"FP32 apply rel L2 {l2:.2e}, max {mx:.2e}"


def test_fused_apply_equals_separate(cuda):
    from openfpm_numerics_amd import Derivative_x, Laplacian, fused, workloads as W

    w = W.jittered_3d_holes(24, seed=5)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dx = Derivative_x(ps, vl, 2, w["rcut"])
    f = torch.as_tensor(W.sin_field(w["pos"])[0], device=cuda)
    for dt in (torch.float64, torch.float32):
        a, b = fused(lap, dx, f=f.to(dt))
        assert torch.equal(a, lap(f.to(dt)))
        assert torch.equal(b, dx(f.to(dt)))


def test_multi_field_apply_equals_single(cuda):
    from openfpm_numerics_amd import Laplacian, dcpse as D, workloads as W

    w = W.jittered_2d(64, seed=3)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    lap = Laplacian(ps, vl, 2, w["rcut"])
    g = torch.Generator(device="cpu").manual_seed(1)
    F = torch.randn(13, ps.size_local(), generator=g, dtype=torch.float64).to(cuda)
    batch = lap(F)
    # one field at a time through the same CSR kernel (the sliced-ELL copy,
    # which single fields run by default, sums in another order)
    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 1 << 30
    try:
        for b in range(F.shape[0]):
            assert torch.equal(batch[b], lap(F[b]))
    finally:
        D.SELL_AFTER = sell_after
    for b in range(F.shape[0]):
        single = lap(F[b])
        assert float((batch[b] - single).abs().max()) <= 1e-13 * float(single.abs().max())
    # FP32: the batch runs the CSR kernel, a single field the sliced-ELL one
    # (different summation order), so they agree to FP32 rounding only
    F32 = F.to(torch.float32)
    batch = lap(F32).double()
    for b in range(F.shape[0]):
        one = lap(F32[b]).double()
        assert float((batch[b] - one).abs().max()) <= 1e-6 * float(one.abs().max())


def test_every_apply_shape(oracle, cuda):
    """Every compiled (lanes, elements-per-lane) shape and block size agrees with the
    FP64 oracle (FP32 within 1e-5 relative L2, FP64 within 1e-12), single and
    multi-field, fused and transposed."""
    from openfpm_numerics_amd import Laplacian, _lib, dcpse as D, workloads as W
    from openfpm_numerics_amd.autotune import APPLY_SHAPES

    w = W.jittered_3d_holes(22, seed=13)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    lap = Laplacian(ps, vl, 2, w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    refs = [oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(3)), 2,
                            w["box_lo"], w["box_hi"], w["periodic"]) for d in range(3)]
    rng = np.random.default_rng(1)
    F = rng.standard_normal((5, ps.size_local()))
    want = np.stack([sum(oracle.apply(r, rp, col, F[b]) for r in refs) for b in range(5)])
    F32 = F.astype(np.float32)
    want32 = np.stack([sum(oracle.apply(r, rp, col, F32[b].astype(np.float64)) for r in refs) for b in range(5)])
    ctx = _lib.context(0)
    try:
        for g, e, vw in APPLY_SHAPES:
            for block in (64, 256, 512):
                if block % g:
                    continue
                print(f"[shape] G={g} epl={e} vec={vw} block={block}", flush=True)
                ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, g, e, vw, block), "config")
                got = lap(torch.as_tensor(F, device=cuda)).cpu().numpy()
                assert np.max(np.abs(got - want)) / np.max(np.abs(want)) <= 1e-12
                got32 = lap(torch.as_tensor(F32, device=cuda)).cpu().numpy().astype(np.float64)
                assert np.linalg.norm(got32 - want32) / np.linalg.norm(want32) <= 1e-5
                D.USE_SELL = False  # single-field FP32 through the CSR kernels of this shape
                try:
                    one = lap(torch.as_tensor(F32[4], device=cuda)).cpu().numpy().astype(np.float64)
                    tr = lap.applyTranspose(torch.as_tensor(F32[4], device=cuda)).cpu().numpy()
                finally:
                    D.USE_SELL = True
                assert np.linalg.norm(one - want32[4]) / np.linalg.norm(want32[4]) <= 1e-5
                assert np.all(np.isfinite(tr))
                gt = lap.applyTranspose(torch.as_tensor(F, device=cuda)).cpu().numpy()
                one = lap.applyTranspose(torch.as_tensor(F[2], device=cuda)).cpu().numpy()
                assert np.max(np.abs(gt[2] - one)) <= 1e-12 * np.max(np.abs(one))
        assert ctx.lib.dcpse_set_apply_config(ctx.h, 8, 8, 1, 1024) != 0
        assert ctx.lib.dcpse_set_apply_config(ctx.h, 8, 3, 1, 256) != 0
        assert ctx.lib.dcpse_set_apply_config(ctx.h, 64, 1, 4, 256) != 0
    finally:
        ctx.check(ctx.lib.dcpse_set_apply_config(ctx.h, 0, 4, 4, 256), "config reset")


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_sell_layout_apply(name, oracle, cuda):
    """The sliced-ELL FP32 apply (forward, fused, transposed; both unroll depths)
    against the FP64 oracle of the FP32-rounded input, and against the CSR kernel."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, _lib, dcpse as D, fused, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    dim = ps.dims
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dx = Derivative_x(ps, vl, 2, w["rcut"])
    refs = [oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(dim)), 2,
                            w["box_lo"], w["box_hi"], w["periodic"]) for d in range(dim)]
    ref_dx = oracle.assemble(w["pos"], rp, col, (1,) + (0,) * (dim - 1), 2, w["box_lo"], w["box_hi"],
                             w["periodic"])
    rng = np.random.default_rng(3)
    f32 = rng.standard_normal(ps.size_local()).astype(np.float32)
    f64 = f32.astype(np.float64)
    want_lap = sum(oracle.apply(r, rp, col, f64) for r in refs)
    want_dx = oracle.apply(ref_dx, rp, col, f64)
    ctx = _lib.context(0)
    ft = torch.as_tensor(f32, device=cuda)
    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 1  # the sliced-ELL copies from the first apply on
    try:
        for gd in ((2, 2), (4, 2)):
            print(f"[sell] {name} groups,depth={gd}", flush=True)
            ctx.check(ctx.lib.dcpse_set_sell_pipeline(ctx.h, *gd), "pipeline")
            a, b = fused(lap, dx, f=ft)
            for got, want in ((a, want_lap), (b, want_dx)):
                got = got.cpu().numpy().astype(np.float64)
                assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5
            D.USE_SELL = False
            try:
                csr = lap(ft).double()
            finally:
                D.USE_SELL = True
            assert float((a.double() - csr).abs().max()) <= 1e-6 * float(csr.abs().max())
            # transposed operator
            A = None
            for r in refs:
                Ad = oracle.value_nz(r, rp, col)
                A = Ad if A is None else A + Ad
            want_t = A.T @ f64
            got_t = lap.applyTranspose(ft).cpu().numpy().astype(np.float64)
            assert np.linalg.norm(got_t - want_t) / np.linalg.norm(want_t) <= 1e-5
            assert dx.dcpse._W.get("sell") is not None and lap.dcpse._T.get("bundle16") is not None
    finally:
        D.SELL_AFTER = sell_after
        ctx.check(ctx.lib.dcpse_set_sell_pipeline(ctx.h, 2, 2), "pipeline reset")


def test_sell_copy_built_from_the_nth_apply(cuda):
    """With SELL_AFTER raised to 4, FP32 single-field applies run the CSR
    kernel for 3 applies since assembly, then the 16-bit sliced-ELL copy
    (built then); a reassembly starts the count again. Both give the same
    result to FP32 rounding of the row sums."""
    from openfpm_numerics_amd import Laplacian, dcpse as D, workloads as W

    w = W.jittered_2d(96, seed=7)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    lap = Laplacian(ps, vl, 2, w["rcut"])
    f = torch.randn(ps.size_local(), generator=torch.Generator().manual_seed(2)).to(cuda)
    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 4
    try:
        outs = []
        for i in range(D.SELL_AFTER + 2):
            outs.append(lap(f))
            assert ("bundle16" in lap.dcpse._W) == (i + 1 >= D.SELL_AFTER), i
        ref = outs[0].double()
        for o in outs[1:]:
            assert float((o.double() - ref).abs().max()) <= 1e-6 * float(ref.abs().max())
        lap.dcpse.initializeStaticSize(ps, ps, check=False)
        lap(f)
        assert "bundle16" not in lap.dcpse._W
    finally:
        D.SELL_AFTER = sell_after


@pytest.mark.parametrize("name", ["2d_jitter_96", "3d_holes_20"])
def test_repeated_applies_are_bitwise_equal(name, cuda):
    """With the default SELL_AFTER every single-field apply of an operator runs
    one kernel on one layout, so repeated applies, the autograd forward and
    the transposed apply give identical bits, FP32 and FP64, after a
    reassembly too."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, dcpse_apply, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    ops = [Laplacian(ps, vl, 2, w["rcut"]), Derivative_x(ps, vl, 2, w["rcut"])]
    g = torch.Generator().manual_seed(5)
    for dt in (torch.float32, torch.float64):
        f = torch.randn(ps.size_local(), generator=g, dtype=torch.float64).to(dt).to(cuda)
        for op in ops:
            first = op(f)
            firstT = op.applyTranspose(f)
            for rep in range(20):
                assert torch.equal(op(f), first), (name, dt, rep)
                assert torch.equal(op.applyTranspose(f), firstT), (name, dt, rep)
            x = f.clone().requires_grad_(True)
            y = dcpse_apply(op, x)
            assert torch.equal(y.detach(), first)
            y.backward(f)
            assert torch.equal(x.grad, firstT)
            op.dcpse.initializeStaticSize(ps, ps)
            assert torch.equal(op(f), first), (name, dt, "after reassembly")


def test_fp64_weights_formed_on_demand(cuda):
    """The assembly writes the kernels and the FP32 weights only; the FP64
    pre-scaled weights eps_p^-|beta| K_pq are formed at the first FP64 use
    (dcpse_scale_rows_f64), bit for bit the single product the assembly's
    epilogue would write, and formed again after a reassembly."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, workloads as W

    w = W.jittered_3d_holes(20, seed=11)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rows = torch.repeat_interleave(torch.arange(vl.n_rows, device=cuda), torch.diff(vl.row_ptr))
    for op in (Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse):
        assert torch.float64 not in op._W
        W64 = op.weights(torch.float64)
        assert torch.equal(W64, op.calcKernels * op.localEpsInvPow[rows])
        assert op.weights(torch.float64) is W64
        op.initializeStaticSize(ps, ps)
        assert torch.float64 not in op._W
        assert torch.equal(op.weights(torch.float64), W64)


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_sell_fp64_apply(name, oracle, cuda):
    """FP64 single-field applies switch to the FP64 sliced-ELL copy at the
    SELL_AFTER-th apply since assembly: fused Laplacian + Dx against the oracle
    (1e-10 relative, the FP64 apply bar) and against the CSR kernel (FP64 sums
    in another order: 1e-13 of the largest value)."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, dcpse as D, fused, workloads as W

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    dim = ps.dims
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dx = Derivative_x(ps, vl, 2, w["rcut"])
    refs = [oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(dim)), 2,
                            w["box_lo"], w["box_hi"], w["periodic"]) for d in range(dim)]
    ref_dx = oracle.assemble(w["pos"], rp, col, (1,) + (0,) * (dim - 1), 2, w["box_lo"], w["box_hi"],
                             w["periodic"])
    f64 = np.random.default_rng(4).standard_normal(ps.size_local())
    want = [sum(oracle.apply(r, rp, col, f64) for r in refs), oracle.apply(ref_dx, rp, col, f64)]
    ft = torch.as_tensor(f64, device=cuda)
    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 2
    try:
        csr = fused(lap, dx, f=ft)  # first apply since assembly: the CSR kernel
        assert "sell64" not in lap.dcpse._W and "sell64" not in dx.dcpse._W
        got = fused(lap, dx, f=ft)
    finally:
        D.SELL_AFTER = sell_after
    assert lap.dcpse._W.get("sell64") is not None and dx.dcpse._W.get("sell64") is not None
    for g, c, wv in zip(got, csr, want):
        gn = g.cpu().numpy()
        assert np.linalg.norm(gn - wv) / np.linalg.norm(wv) <= 1e-10
        assert float((g - c).abs().max()) <= 1e-13 * float(c.abs().max())
    # alpha / beta epilogue: out = 0.5 * A f + 2 * out
    out = [c.clone() for c in csr]
    D.SELL_AFTER = 1
    try:
        D.apply_ops([lap.dcpse, dx.dcpse], ft, outs=out, alpha=0.5, beta=2.0)
    finally:
        D.SELL_AFTER = sell_after
    for o, g, c in zip(out, got, csr):
        assert float((o - (0.5 * g + 2.0 * c)).abs().max()) <= 1e-13 * float(c.abs().max())


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_sell_bundle_equals_sell(name, cuda):
    """The interleaved bundle layout reads the same slots in the same order:
    bitwise the same result as the separate-array SELL layout."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, workloads as W
    from openfpm_numerics_amd.dcpse import PreparedApply

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    ops = [Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse]
    f = torch.randn(ps.size_local(), generator=torch.Generator().manual_seed(5)).to(cuda)
    for k in (1, 2):
        a = [torch.empty_like(f) for _ in range(k)]
        b = [torch.empty_like(f) for _ in range(k)]
        PreparedApply(ops[:k], f, a, sell=True)()
        PreparedApply(ops[:k], f, b, sell="bundle")()
        for x, y in zip(a, b):
            assert torch.equal(x, y), name
        c = [torch.empty_like(f) for _ in range(k)]
        p16 = PreparedApply(ops[:k], f, c, sell="bundle16")
        assert p16.layout == "sell_bundle16" and p16.n_exc == 0, name  # every offset fits 16 bits
        p16()
        for x, y in zip(a, c):
            assert torch.equal(x, y), name


def test_sell_bundle16_far_neighbours(oracle, cuda):
    """Shuffled particle order: about a third of the column offsets exceed
    16 bits and go through the per-lane far-neighbour lists; the result
    matches the separate-array SELL apply and the FP64 oracle."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, workloads as W
    from openfpm_numerics_amd.dcpse import PreparedApply

    w = W.jittered_3d_holes(48, seed=13)
    perm = np.random.default_rng(0).permutation(w["pos"].shape[0])
    w["pos"] = np.ascontiguousarray(w["pos"][perm])
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    ops = [Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse]
    f = torch.randn(ps.size_local(), generator=torch.Generator().manual_seed(2)).to(cuda)
    a = [torch.empty_like(f) for _ in ops]
    c = [torch.empty_like(f) for _ in ops]
    PreparedApply(ops, f, a, sell=True)()
    p16 = PreparedApply(ops, f, c, sell="bundle16")
    print(f"[bundle16] N={ps.size_local()} nnz={vl.nnz} far neighbours {p16.n_exc}", flush=True)
    assert p16.n_exc > vl.nnz // 10
    p16()
    for x, y in zip(a, c):
        assert float((x - y).abs().max()) <= 1e-6 * float(x.abs().max())
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    refs = [oracle.assemble(w["pos"], rp, col, s, 2, w["box_lo"], w["box_hi"], w["periodic"])
            for s in ((2, 0, 0), (0, 2, 0), (0, 0, 2))]
    f64 = f.double().cpu().numpy()
    want = sum(oracle.apply(r, rp, col, f64) for r in refs)
    got = c[0].double().cpu().numpy()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5
    # the time-integration stage on the one-operator bundle, far lists included
    from openfpm_numerics_amd.ode import stage

    ya = torch.randn(f.numel(), generator=torch.Generator().manual_seed(3)).to(cuda)
    outs = {}
    for layout in ("sell", "bundle16"):
        o = torch.empty_like(f)
        stage(ops[0], f, 0.5, ya, 0.25, o, None, 0.0, None, layout=layout)
        outs[layout] = o
    assert ops[0].sell_bundle16()[5] > vl.nnz // 10
    scale = float((outs["sell"] - ya).abs().max())
    assert float((outs["sell"] - outs["bundle16"]).abs().max()) <= 1e-6 * scale
    # with this many far neighbours the automatic choice takes the 32-bit layout
    from openfpm_numerics_amd import dcpse as D

    assert not ops[0].bundle16_suits()
    sell_after = D.SELL_AFTER
    D.SELL_AFTER = 1
    try:
        lap_auto = ops[0].computeDifferentialOperator(f)
    finally:
        D.SELL_AFTER = sell_after
    assert ops[0]._W.get("sell") is not None
    assert float((lap_auto - a[0]).abs().max()) <= 1e-6 * float(a[0].abs().max())


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_assemble_together_equals_separate(name, cuda):
    """Operators assembled in one pass (shared gathers, epsilon, moments) give
    the kernels of separate assemblies: epsilon bit for bit; the kernels to
    rounding, since an operator with a smaller basis sums its moments in
    another order when assembled alone (more neighbours per moment pass)."""
    from openfpm_numerics_amd import Derivative_x, Derivative_y, Laplacian, workloads as W
    from openfpm_numerics_amd.dcpse import LOAD, assemble_together

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    sep = [Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse,
           Derivative_y(ps, vl, 2, w["rcut"]).dcpse]
    tog = [Laplacian(ps, vl, 2, w["rcut"], LOAD).dcpse, Derivative_x(ps, vl, 2, w["rcut"], LOAD).dcpse,
           Derivative_y(ps, vl, 2, w["rcut"], LOAD).dcpse]
    assemble_together(tog)
    for a, b in zip(sep, tog):
        scale = float(a.calcKernels.abs().max())
        assert float((a.calcKernels - b.calcKernels).abs().max()) <= 1e-13 * scale, name
        assert torch.equal(a.localEps, b.localEps) and torch.equal(a.localEpsInvPow, b.localEpsInvPow)
        w32a, w32b = a.weights(torch.float32), b.weights(torch.float32)
        assert float((w32a - w32b).abs().max()) <= 2.0 ** -22 * float(w32a.abs().max()), name
        assert int(b.info.abs().max()) == 0


@pytest.mark.parametrize("name", ["2d_jitter_96", "3d_holes_20"])
def test_refinement_step_agrees(name, cuda):
    """The optional FP64 refinement step (DCPSE_ASM_REFINE) and the plain
    Cholesky solve give the same kernels to 1e-12 row-relative on these
    well-conditioned moment systems, so refinement stays an option."""
    from openfpm_numerics_amd import workloads as W
    from openfpm_numerics_amd.dcpse import LOAD, Dcpse

    w = WORKLOADS[name](W)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    dim = len(w["box_lo"])
    lap = [tuple(2 if e == d else 0 for e in range(dim)) for d in range(dim)]
    rp = vl.row_ptr.cpu().numpy()
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    for sigs in (lap, [tuple(1 if e == 0 else 0 for e in range(dim))]):
        K = []
        for refine in (False, True):
            d = Dcpse(ps, vl, sigs[0], 2, w["rcut"], LOAD, signatures=sigs, coeffs=[1.0] * len(sigs), refine=refine)
            d.initializeStaticSize(ps, ps)
            K.append(d.calcKernels.cpu().numpy())
        num = np.zeros(rp.size - 1)
        den = np.zeros(rp.size - 1)
        np.maximum.at(num, rows, np.abs(K[0] - K[1]))
        np.maximum.at(den, rows, np.abs(K[1]))
        rel = float(np.max(num / np.maximum(den, 1e-300)))
        print(f"[refine] {name} {sigs}: row-relative {rel:.2e}", flush=True)
        assert rel <= 1e-12


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("refine", [False, True])
def test_nested_pairs_share_one_factorisation(dim, refine, oracle, cuda):
    """Pairs whose bases nest are factored once for both operators: Laplacian
    + Dx (Dx's basis is the Laplacian's degree-1-2 monomials and the
    constant), Dx + Dy and Dxx + Dyy (one basis). On an open box (10^3 in
    3D, 20^2 in 2D) at rCut 2.3h the corner particles have enough neighbours
    for Dx (10 monomials in 3D, 6 in 2D) but too few for the Laplacian (19,
    9), so the pair marks only the Laplacian TOO_FEW there. Every pair equals
    the oracle per row (1e-10, or 16 cond(A) u for ill-conditioned boundary
    rows) with the oracle's status flags, and the single-operator assembly to
    the same bar, with and without the refinement step."""
    from _dcpse_tol import check_weights
    from openfpm_numerics_amd import ParticleSet
    from openfpm_numerics_amd.dcpse import LOAD, Dcpse, assemble_together

    rng = np.random.default_rng(11)
    side = 10 if dim == 3 else 20
    h = 1.0 / side
    g = (np.arange(side) + 0.5) * h
    pos = np.stack(np.meshgrid(*([g] * dim), indexing="ij"), -1).reshape(-1, dim)
    pos += rng.uniform(-0.05 * h, 0.05 * h, pos.shape)
    lo, hi = [0.0] * dim, [1.0] * dim
    ps = ParticleSet(pos, lo, hi, bc=[0] * dim, device=cuda)
    rc = 2.3 * h
    vl = ps.getVerlet(rc)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    nn = np.diff(rp)
    m_dx, m_lap = (10, 19) if dim == 3 else (6, 9)
    few = (nn >= m_dx) & (nn < m_lap)
    assert few.sum() >= 2 ** dim, nn.min()

    def e(*k):
        return tuple(k[:dim])

    sig = {"lap": [e(2, 0, 0), e(0, 2, 0), e(0, 0, 2)][:dim], "dx": [e(1, 0, 0)], "dy": [e(0, 1, 0)],
           "dxx": [e(2, 0, 0)], "dyy": [e(0, 2, 0)]}

    def make(k):
        return Dcpse(ps, vl, sig[k][0], 2, rc, LOAD, signatures=sig[k], coeffs=[1.0] * len(sig[k]), refine=refine)

    for pair in (("lap", "dx"), ("dx", "dy"), ("dxx", "dyy")):
        sep = [make(k) for k in pair]
        for d in sep:
            assemble_together([d], ps, ps, check=False)
        tog = [make(k) for k in pair]
        assemble_together(tog, ps, ps, check=False)
        for k, s, t in zip(pair, sep, tog):
            refs = [oracle.assemble(pos, rp, col, b, 2, lo, hi, 0, dense_ev=False) for b in sig[k]]
            info = t.info.cpu().numpy()
            assert np.array_equal(info, refs[0].info), (pair, k)
            assert np.array_equal(info, s.info.cpu().numpy()), (pair, k)
            assert torch.equal(s.localEps, t.localEps)
            Kt = t.calcKernels.cpu().numpy()
            rel, nover, worst = check_weights(Kt, sum(r.K for r in refs), rp, pos, pos, col,
                                              oracle.basis(sig[k][0], 2), label=f"{pair}/{k}")
            sep_rel, _, _ = check_weights(Kt, s.calcKernels.cpu().numpy(), rp, pos, pos, col,
                                          oracle.basis(sig[k][0], 2), label=f"{pair}/{k} separate")
            print(f"[nested] dim={dim} refine={refine} {pair}/{k}: oracle {rel:.2e} ({nover} rows by cond, worst {worst:.1e}),"
                  f" separate {sep_rel:.2e}, TOO_FEW rows {int((info == 1).sum())}", flush=True)
        if pair[0] == "lap":
            assert (tog[0].info.cpu().numpy()[few] == 1).all() and (tog[1].info.cpu().numpy()[few] == 0).all()


def test_prepared_apply_fp64_runs_the_sliced_ell_copy(cuda):
    """PreparedApply with FP64 fields binds the FP64 sliced-ELL kernel and gives
    the bits of the operators' own FP64 apply."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, fused, workloads as W
    from openfpm_numerics_amd.dcpse import PreparedApply

    w = W.jittered_2d(96, seed=8)
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    ops = [Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse]
    f = torch.randn(ps.size_local(), generator=torch.Generator().manual_seed(3), dtype=torch.float64).to(cuda)
    outs = [torch.empty_like(f) for _ in ops]
    p = PreparedApply(ops, f, outs, sell=True)
    assert p.layout == "sell64"
    p()
    want = fused(*ops, f=f)
    for o, wv in zip(outs, want):
        assert torch.equal(o, wv)


@pytest.mark.parametrize("shuffle", [False, True])
def test_fp64_bundle16_apply(shuffle, oracle, cuda):
    """The FP64 16-bit-offset bundle (dcpse_sell_fill_bundle16_f64 /
    dcpse_apply_sell_bundle16_f64): one operator through the operator call
    and the fused pair through PreparedApply, against the FP64 sliced-ELL
    kernel (bit for bit without far neighbours; <= 1e-13 of max|ref| with a
    shuffled order, where far neighbours go through the per-lane lists) and
    against the oracle (<= 1e-10 of max|ref|)."""
    from openfpm_numerics_amd import Derivative_x, Laplacian, dcpse as D, workloads as W
    from openfpm_numerics_amd.dcpse import PreparedApply

    w = W.jittered_3d_holes(24 if not shuffle else 48, seed=14)  # 48^3: offsets beyond 16 bits
    if shuffle:
        perm = np.random.default_rng(1).permutation(w["pos"].shape[0])
        w["pos"] = np.ascontiguousarray(w["pos"][perm])
    ps = _ps(w, cuda)
    vl = ps.getVerlet(w["rcut"])
    lap = Laplacian(ps, vl, 2, w["rcut"])
    dx = Derivative_x(ps, vl, 2, w["rcut"])
    ops = [lap.dcpse, dx.dcpse]
    f = torch.as_tensor(np.sin(2 * np.pi * w["pos"][:, 0]) * np.cos(2 * np.pi * w["pos"][:, 1]), device=cuda)
    a = [torch.empty_like(f) for _ in ops]
    b = [torch.empty_like(f) for _ in ops]
    PreparedApply(ops, f, a, sell=True)()  # FP64 sliced-ELL, 32-bit columns
    p16 = PreparedApply(ops, f, b, sell="bundle16")
    assert p16.layout == "sell64_bundle16"
    print(f"[fp64 bundle16] shuffle={shuffle} N={ps.size_local()} nnz={vl.nnz} far neighbours {p16.n_exc}",
          flush=True)
    assert (p16.n_exc > 0) == shuffle
    p16()
    for x, y in zip(a, b):
        if shuffle:
            assert float((x - y).abs().max()) <= 1e-13 * float(x.abs().max())
        else:
            assert torch.equal(x, y)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    lo, hi, per = w["box_lo"], w["box_hi"], w["periodic"]
    refs = [oracle.assemble(w["pos"], rp, col, s, 2, lo, hi, per) for s in ((2, 0, 0), (0, 2, 0), (0, 0, 2))]
    ref_dx = oracle.assemble(w["pos"], rp, col, (1, 0, 0), 2, lo, hi, per)
    fn = f.cpu().numpy()
    wants = [sum(oracle.apply(r, rp, col, fn) for r in refs), oracle.apply(ref_dx, rp, col, fn)]
    for got, want in zip(b, wants):
        assert np.abs(got.cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()
    # one operator through the operator call: the FP64 bundle from the first apply on
    if not shuffle:
        assert D.SELL_AFTER == 1
        got = dx(f)
        assert dx.dcpse._W.get("bundle16_64") is not None
        assert torch.equal(got, a[1])
