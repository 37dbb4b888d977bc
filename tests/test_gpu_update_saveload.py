"""Operators after particles move (initializeUpdate / update, src/DCPSE/Dcpse.hpp:661-687,
DCPSE_op.hpp:779-784, as dcpse_op_convection drives them,
src/DCPSE/DCPSE_op/tests/DCPSE_op_test_base_tests.cpp:1025-1156) and kernel
files (save / load, Dcpse.hpp:302-379), through the Python layer.

Tolerances: FP64 weights per row <= 1e-10 (or 16 cond(A) u, _dcpse_tol) and
FP64 applies <= 1e-10 of max|ref| against the oracle at the new positions;
FP32 applies <= 1e-5 relative L2 against the FP64 oracle on FP32-rounded
input; a loaded operator equals the saved one bit for bit.
"""
import numpy as np
import pytest
import torch
from _dcpse_tol import check_weights

pytestmark = pytest.mark.gpu


def _lattice(n, lo, hi):
    h = (hi - lo) / (n - 1)
    k = np.arange(n * n)
    return np.stack([lo + (k % n) * h, lo + (k // n) * h], 1), h


def _velocity(pos):
    x, y = pos[:, 0], pos[:, 1]
    g = np.exp(-10 * (x * x + y * y))
    inside = (np.abs(x) < 1) & (np.abs(y) < 1)
    return np.stack([-y * g * inside, x * g * inside], 1)


def test_update_after_motion_matches_oracle(oracle, cuda):
    from openfpm_numerics_amd import Derivative_xx, Derivative_yy, Laplacian, ParticleSet

    pos, h = _lattice(41, -1.0, 1.0)
    rc = 3.1 * h
    lo, hi = [-1.0, -1.0], [1.0, 1.0]
    ps = ParticleSet(pos, lo, hi, bc=[0, 0], device=cuda)
    vl = ps.getVerlet(rc)
    dxx = Derivative_xx(ps, vl, 2, rc)
    dyy = Derivative_yy(ps, vl, 2, rc)
    lap = Laplacian(ps, vl, 2, rc)
    C = np.exp(-8.0 * (pos[:, 0] ** 2 + (pos[:, 1] - 0.2) ** 2))
    Ct = torch.as_tensor(C, device=cuda)
    dt = 1e-2
    for step in range(5):
        print(f"[update] step {step}", flush=True)
        Ct = Ct + dt * 0.01 * (dxx(Ct) + dyy(Ct))
        pos = pos + dt * _velocity(pos)
        ps.set_positions(pos)
        with pytest.raises(RuntimeError):
            dxx(Ct)  # stale: the particles moved since assembly
        ps.updateVerlet(vl, rc)
        for op in (dxx, dyy, lap):
            op.update(ps)
    rp, col = oracle.support(pos, rc, lo, hi, 0)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp) and np.array_equal(vl.col.cpu().numpy(), col)
    Cn = Ct.cpu().numpy()
    refs = {b: oracle.assemble(pos, rp, col, b, 2, lo, hi, 0) for b in ((2, 0), (0, 2))}
    for name, op, betas in (("dxx", dxx, [(2, 0)]), ("dyy", dyy, [(0, 2)]), ("lap", lap, [(2, 0), (0, 2)])):
        Kr = sum(refs[b].K for b in betas)
        err_k, n_cond, worst = check_weights(op.dcpse.calcKernels.cpu().numpy(), Kr, rp, pos, pos, col,
                                             oracle.basis(betas[0], 2), label=name)
        want = sum(oracle.apply(refs[b], rp, col, Cn) for b in betas)
        err = np.abs(op(Ct).cpu().numpy() - want).max() / np.abs(want).max()
        print(f"[update] {name}: weights {err_k:.2e} ({n_cond} rows by cond(A)), apply {err:.2e}", flush=True)
        assert err <= max(1e-10, 16 * worst * 2.0 ** -53), (name, err)


def test_update_two_sets_after_motion(oracle, cuda):
    """A subset-domain operator follows its parent set: positions re-gathered,
    Verlet rows rebuilt, kernels reassembled."""
    from openfpm_numerics_amd import Derivative_x, ParticleSet, ParticleSubset, workloads as W

    w = W.jittered_2d(48, seed=12)
    pos, rc = w["pos"], w["rcut"]
    ps = ParticleSet(pos, [0, 0], [1, 1], bc=[1, 1], device=cuda)
    ids = np.arange(0, pos.shape[0], 3, dtype=np.int32)
    sub = ParticleSubset(ps, ids)
    vl = sub.getVerlet(rc)
    dx = Derivative_x(ps, vl, 2, rc, particlesDomain=sub)
    rng = np.random.default_rng(3)
    pos = np.mod(pos + rng.uniform(-0.3 * w["h"], 0.3 * w["h"], pos.shape), 1.0)
    ps.set_positions(pos)
    f = torch.as_tensor(np.sin(2 * np.pi * pos[:, 0]), device=cuda)
    with pytest.raises(RuntimeError):
        dx(f)
    ps.updateVerlet(vl, rc)
    dx.update(ps)
    rp, col = oracle.support(pos, rc, [0, 0], [1, 1], 3, dom_ids=ids)
    assert np.array_equal(vl.row_ptr.cpu().numpy(), rp) and np.array_equal(vl.col.cpu().numpy(), col)
    asm = oracle.assemble(pos, rp, col, (1, 0), 2, [0, 0], [1, 1], 3, pos_dom=pos[ids])
    fn = f.cpu().numpy()
    want = oracle.apply(asm, rp, col, fn, f_dom=fn[ids])
    assert np.abs(dx(f).cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()


def test_save_load_then_sliced_ell_applies(oracle, cuda, tmp_path):
    """save -> load into a LOAD operator on the same Verlet list, then more than
    SELL_AFTER applies in FP32 and FP64 (the sliced-ELL copies are built from
    the loaded weights on the list's own layout), while a second operator on
    the same list keeps working."""
    from openfpm_numerics_amd import LOAD, Derivative_x, Laplacian, ParticleSet, dcpse as D, workloads as W

    w = W.jittered_2d(64, seed=21)
    pos, rc = w["pos"], w["rcut"]
    ps = ParticleSet(pos, [0, 0], [1, 1], bc=[1, 1], device=cuda)
    vl = ps.getVerlet(rc)
    lap = Laplacian(ps, vl, 2, rc)
    dx = Derivative_x(ps, vl, 2, rc)
    f64 = np.sin(2 * np.pi * pos[:, 0]) * np.cos(2 * np.pi * pos[:, 1])
    f32 = f64.astype(np.float32)
    for _ in range(D.SELL_AFTER + 1):  # the list's sliced-ELL layout exists before the load
        dx(torch.as_tensor(f32, device=cuda))
    rp_before, col_before, sell_before = vl.row_ptr.clone(), vl.col.clone(), vl.sell()
    path = str(tmp_path / "lap")
    lap.save(ps, path)
    ld = Laplacian(ps, vl, 2, rc, LOAD)
    ld.load(ps, path)
    assert torch.equal(ld.dcpse.calcKernels, lap.dcpse.calcKernels)
    assert torch.equal(ld.dcpse.weights(torch.float64), lap.dcpse.weights(torch.float64))
    assert torch.equal(vl.row_ptr, rp_before) and torch.equal(vl.col, col_before) and vl.sell() is sell_before
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    refs = [oracle.assemble(pos, rp, col, b, 2, [0, 0], [1, 1], 3) for b in ((2, 0), (0, 2))]
    want64 = sum(oracle.apply(r, rp, col, f64) for r in refs)
    want32 = sum(oracle.apply(r, rp, col, f32.astype(np.float64)) for r in refs)
    t64, t32 = torch.as_tensor(f64, device=cuda), torch.as_tensor(f32, device=cuda)
    for i in range(D.SELL_AFTER + 2):
        g64 = ld(t64).cpu().numpy()
        assert np.abs(g64 - want64).max() <= 1e-10 * np.abs(want64).max(), i
        g32 = ld(t32).cpu().numpy().astype(np.float64)
        assert np.linalg.norm(g32 - want32) / np.linalg.norm(want32) <= 1e-5, i
        assert np.array_equal(ld(t32).cpu().numpy(), lap(t32).cpu().numpy()), i
    # the operator sharing the list is untouched
    ref_dx = oracle.assemble(pos, rp, col, (1, 0), 2, [0, 0], [1, 1], 3)
    want = oracle.apply(ref_dx, rp, col, f64)
    assert np.abs(dx(t64).cpu().numpy() - want).max() <= 1e-10 * np.abs(want).max()


def test_load_refuses_another_support(cuda, tmp_path):
    from openfpm_numerics_amd import LOAD, Derivative_x, ParticleSet, workloads as W

    w = W.jittered_2d(40, seed=22)
    ps = ParticleSet(w["pos"], [0, 0], [1, 1], bc=[1, 1], device=cuda)
    vl = ps.getVerlet(w["rcut"])
    vl_big = ps.getVerlet(1.2 * w["rcut"])
    path = str(tmp_path / "dx_big")
    Derivative_x(ps, vl_big, 2, 1.2 * w["rcut"]).save(ps, path)
    rp, col = vl.row_ptr.clone(), vl.col.clone()
    ld = Derivative_x(ps, vl, 2, w["rcut"], LOAD)
    with pytest.raises(ValueError):
        ld.load(ps, path)
    assert torch.equal(vl.row_ptr, rp) and torch.equal(vl.col, col)
