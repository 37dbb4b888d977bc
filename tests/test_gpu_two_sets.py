"""Two particle sets (support != domain, src/DCPSE/Dcpse.hpp:111-130) on the GPU,
against the oracle.

Two forms of the domain:
  * subset: particles of the support set picked by index (``ParticleSubset``,
    the reference's vector_dist_subset,
    src/DCPSE/DCPSE_op/tests/DCPSE_op_subset_test.cpp:113-141); each row skips
    its own particle;
  * separate: an arbitrary point set; nothing is skipped (a domain point on a
    support particle lists it at distance 0).

Tolerances: support bit-exact (and equal to a numpy brute-force search);
FP64 weights per row max|K_gpu - K_ref| / max|K_ref| <= 1e-10, or
16 cond(A) u for a row whose normal equations are worse conditioned than
that allows (_dcpse_tol: corner particles of an open box with exactly m
neighbours); FP64 apply max|D_gpu - D_ref| / max|D_ref| <= 1e-10; FP32 apply
(CSR kernel and the sliced-ELL copy) relative L2 <= 1e-5 against the FP64
oracle applied to the FP32-rounded input, and per row within 1e-5 of
max|ref| or the FP32 rounding bound of its weights (_dcpse_tol.check_fp32).
"""
import numpy as np
import pytest
import torch
from _dcpse_tol import check_fp32, check_weights

pytestmark = pytest.mark.gpu


def _brute(sup, dom, rc, periodic, skip=None):
    d = dom[:, None, :] - sup[None, :, :]
    for k in range(sup.shape[1]):
        if periodic >> k & 1:
            d[..., k] -= np.round(d[..., k])
    d2 = (d ** 2).sum(-1)
    if skip is not None:
        d2[np.arange(dom.shape[0]), skip] = np.inf
    return [np.nonzero(d2[p] < rc * rc)[0] for p in range(dom.shape[0])]


def _case(kind, periodic, cuda):
    """Support set (jittered 64^2 lattice), domain and the oracle's CSR."""
    import oracle as O
    from openfpm_numerics_amd import ParticleSet, ParticleSubset, workloads as W
    from openfpm_numerics_amd.particles import VerletList

    w = W.jittered_2d(64, seed=31)
    pos, rc = w["pos"], w["rcut"]
    bc = [(periodic >> d) & 1 for d in range(2)]
    sup = ParticleSet(pos, [0, 0], [1, 1], bc=bc, device=cuda)
    rng = np.random.default_rng(32)
    if kind == "subset":
        ids = np.sort(rng.choice(pos.shape[0], size=pos.shape[0] * 3 // 5, replace=False)).astype(np.int32)
        dom = ParticleSubset(sup, ids)
        dpos = pos[ids]
        vl = dom.getVerlet(rc)
        rp, col = O.support(pos, rc, [0, 0], [1, 1], periodic, dom_ids=ids)
        want = _brute(pos, dpos, rc, periodic, skip=ids)
    else:
        dpos = rng.random((1500, 2))
        dpos[:100] = pos[200:300]  # on top of support particles
        dom = ParticleSet(dpos, [0, 0], [1, 1], bc=bc, device=cuda)
        vl = VerletList.build(sup, rc, domain=dom)
        rp, col = O.support(pos, rc, [0, 0], [1, 1], periodic, pos_dom=dpos)
        want = _brute(pos, dpos, rc, periodic)
    return dict(w=w, sup=sup, dom=dom, dpos=dpos, vl=vl, rp=rp, col=col, want=want, periodic=periodic)


CASES = [("subset", 0b11), ("subset", 0b00), ("separate", 0b11), ("separate", 0b00)]


@pytest.mark.parametrize("kind,periodic", CASES)
def test_two_set_support(kind, periodic, cuda):
    print(f"[two-sets support] {kind} periodic={periodic}", flush=True)
    c = _case(kind, periodic, cuda)
    gp, gc = c["vl"].row_ptr.cpu().numpy(), c["vl"].col.cpu().numpy()
    assert np.array_equal(gp, c["rp"])
    assert np.array_equal(gc, c["col"])
    for p, want in enumerate(c["want"]):
        assert np.array_equal(gc[gp[p]:gp[p + 1]], want), p


@pytest.mark.parametrize("kind,periodic", CASES)
def test_two_set_assembly_and_apply(kind, periodic, cuda):
    import oracle as O
    from openfpm_numerics_amd import Derivative_x, Derivative_y, Laplacian, dcpse as D

    print(f"[two-sets asm/apply] {kind} periodic={periodic}", flush=True)
    c = _case(kind, periodic, cuda)
    pos, dpos, rp, col, vl = c["w"]["pos"], c["dpos"], c["rp"], c["col"], c["vl"]
    rc = c["w"]["rcut"]
    ops = {"dx": (Derivative_x, [(1, 0)]), "dy": (Derivative_y, [(0, 1)]), "lap": (Laplacian, [(2, 0), (0, 2)])}
    f64 = np.sin(2 * np.pi * pos[:, 0]) * np.cos(2 * np.pi * pos[:, 1]) + pos[:, 0] ** 2
    fd64 = np.sin(2 * np.pi * dpos[:, 0]) * np.cos(2 * np.pi * dpos[:, 1]) + dpos[:, 0] ** 2
    f32, fd32 = f64.astype(np.float32), fd64.astype(np.float32)
    ft = torch.as_tensor(f64, device=cuda)
    separate = kind == "separate"
    for name, (cls, betas) in ops.items():
        op = cls(c["sup"], vl, 2, rc, particlesDomain=c["dom"])
        refs = [O.assemble(pos, rp, col, b, 2, [0, 0], [1, 1], periodic, pos_dom=dpos) for b in betas]
        ok = np.all([r.info == 0 for r in refs], axis=0)
        assert ok.all(), f"{name}: oracle assembly failed at {np.nonzero(~ok)[0][:5]}"
        assert (op.dcpse.info.cpu().numpy() == 0).all()
        # FP64 weights: the fused Laplacian's kernel is the sum of the per-direction kernels
        Kg = op.dcpse.calcKernels.cpu().numpy()
        Kr = sum(r.K for r in refs)
        period = [1.0 if periodic >> d & 1 else 0.0 for d in range(2)]
        err_k, n_cond, worst_cond = check_weights(Kg, Kr, rp, pos, dpos, col, O.basis(betas[0], 2), period,
                                                  label=name)
        print(f"[two-sets] {kind}/{name} periodic={periodic} weights row-rel max {err_k:.2e}; "
              f"{n_cond} rows judged by cond(A) (max {worst_cond:.2e})", flush=True)
        np.testing.assert_allclose(op.dcpse.localEps.cpu().numpy(), refs[0].eps, rtol=1e-13)
        # FP64 apply (reference form, field on the domain read from the domain set)
        want = sum(O.apply(r, rp, col, f64, f_dom=fd64) for r in refs)
        kw = {"f_domain": torch.as_tensor(fd64, device=cuda)} if separate else {}
        for i in range(D.SELL_AFTER + 1):  # the FP64 sliced-ELL copy from the SELL_AFTER-th apply
            got = op(ft, **kw).cpu().numpy()
            err = np.abs(got - want).max() / np.abs(want).max()
            assert err <= max(1e-10, 16 * worst_cond * 2.0 ** -53), (name, i, err)
        # FP32: the sliced-ELL copy from the SELL_AFTER-th apply (before it the CSR kernel)
        want32 = sum(O.apply(r, rp, col, f32.astype(np.float64), f_dom=fd32.astype(np.float64)) for r in refs)
        ft32 = torch.as_tensor(f32, device=cuda)
        kw32 = {"f_domain": torch.as_tensor(fd32, device=cuda)} if separate else {}
        outs = [op(ft32, **kw32).cpu().numpy().astype(np.float64) for _ in range(D.SELL_AFTER + 1)]
        W = sum(r.K * np.repeat(r.eps_inv_pow, np.diff(rp)) for r in refs)
        fs, fd = f32.astype(np.float64), fd32.astype(np.float64)
        for o in (outs[0], outs[-1]):
            check_fp32(o, want32, rp, col, W, fs, fd, label=name)
        assert np.array_equal(outs[-1], outs[0]), name
        # and the CSR kernel, as a caller who never builds the copy gets it
        sell_after = D.SELL_AFTER
        D.SELL_AFTER = 1 << 30
        try:
            op.dcpse._n32["fwd"] = 0
            csr = op(ft32, **kw32).cpu().numpy().astype(np.float64)
        finally:
            D.SELL_AFTER = sell_after
        check_fp32(csr, want32, rp, col, W, fs, fd, label=name)
