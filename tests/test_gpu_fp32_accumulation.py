"""FP32 accumulation of the FP32 sliced-ELL forward applies
(dcpse_set_apply_accumulation(ctx, 32)): every operation of the row sum in
FP32, as the reference's apply does for float properties
(src/DCPSE/Dcpse.hpp:560-659), on the three sliced-ELL layouts.

Tolerances: against the FP64 oracle applied to the FP32-rounded input,
relative L2 <= 1e-5, and per row |err| <= max(1e-5 max|ref|,
(n_p + 4) 2^-24 sum_q |W_pq| |f_q - f_p|) -- the textbook bound of an n-term
FP32 sum of FP32 products of rounded differences (Higham, Accuracy and
Stability, 3.1), n_p the row length; switching back to FP64 accumulation
gives the earlier results bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(name, W):
    if name == "2d_jitter_96":
        return W.jittered_2d(96, seed=7)
    if name == "3d_holes_20":
        return W.jittered_3d_holes(20, seed=11)
    w = W.jittered_3d_holes(40, seed=13)  # shuffled: a third of the offsets use the far lists
    perm = np.random.default_rng(0).permutation(w["pos"].shape[0])
    w["pos"] = np.ascontiguousarray(w["pos"][perm])
    return w


@pytest.mark.parametrize("name", ["2d_jitter_96", "3d_holes_20", "3d_holes_40_shuffled"])
def test_fp32_accumulation_against_oracle(name, oracle, cuda):
    from openfpm_numerics_amd import Derivative_x, Laplacian, ParticleSet, set_apply_accumulation, workloads as W
    from openfpm_numerics_amd.dcpse import PreparedApply

    w = _case(name, W)
    dim = w["pos"].shape[1]
    bc = [(w["periodic"] >> d) & 1 for d in range(dim)]
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=bc, device=cuda)
    vl = ps.getVerlet(w["rcut"])
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    ops = [Laplacian(ps, vl, 2, w["rcut"]).dcpse, Derivative_x(ps, vl, 2, w["rcut"]).dcpse]
    refs_lap = [oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(dim)), 2,
                                w["box_lo"], w["box_hi"], w["periodic"]) for d in range(dim)]
    ref_dx = oracle.assemble(w["pos"], rp, col, (1,) + (0,) * (dim - 1), 2, w["box_lo"], w["box_hi"], w["periodic"])
    x = w["pos"]
    f64 = np.sin(2 * np.pi * x[:, 0]) * np.cos(2 * np.pi * x[:, 1]) + 0.1 * np.random.default_rng(5).standard_normal(
        x.shape[0])
    f32 = f64.astype(np.float32)
    fr = f32.astype(np.float64)
    wants = [sum(oracle.apply(r, rp, col, fr) for r in refs_lap), oracle.apply(ref_dx, rp, col, fr)]
    Ws = [sum(r.K * np.repeat(r.eps_inv_pow, np.diff(rp)) for r in refs_lap),
          ref_dx.K * np.repeat(ref_dx.eps_inv_pow, np.diff(rp))]
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    lens = np.diff(rp).astype(np.float64)
    ft = torch.as_tensor(f32, device=cuda)
    for layout in (True, "bundle", "bundle16"):
        outs64 = [torch.empty_like(ft) for _ in ops]
        outs32 = [torch.empty_like(ft) for _ in ops]
        again = [torch.empty_like(ft) for _ in ops]
        p64 = PreparedApply(ops, ft, outs64, sell=layout)
        p32 = PreparedApply(ops, ft, outs32, sell=layout)
        pag = PreparedApply(ops, ft, again, sell=layout)
        print(f"[acc32] {name} layout {p32.layout} N={rp.size - 1} nnz={rp[-1]}", flush=True)
        p64()
        try:
            set_apply_accumulation(32)
            p32()
        finally:
            set_apply_accumulation(64)
        pag()
        torch.cuda.synchronize()
        for o, (a64, a32, ag, want, Wr) in enumerate(zip(outs64, outs32, again, wants, Ws)):
            assert torch.equal(a64, ag), (layout, o)
            got = a32.double().cpu().numpy()
            rel = np.linalg.norm(got - want) / np.linalg.norm(want)
            bound = np.bincount(rows, np.abs(Wr) * np.abs(fr[col] - fr[rows]), minlength=rp.size - 1)
            bar = np.maximum(1e-5 * np.abs(want).max(), (lens + 4) * 2.0 ** -24 * bound)
            err = np.abs(got - want)
            d64 = np.linalg.norm(a64.double().cpu().numpy() - want) / np.linalg.norm(want)
            print(f"[acc32]   op {o}: FP32 sums rel L2 {rel:.2e} (FP64 sums {d64:.2e}), "
                  f"max err / bar {float((err / bar).max()):.2f}", flush=True)
            assert rel <= 1e-5, (layout, o, rel)
            assert not (err > bar).any(), (layout, o, np.nonzero(err > bar)[0][:5])
        assert any(not torch.equal(a, b) for a, b in zip(outs32, outs64)), layout  # the FP32 sums really ran


def test_accumulation_setter_rejects_other_widths(cuda):
    from openfpm_numerics_amd import DcpseError, set_apply_accumulation

    for bits in (0, 16, 48, 128):
        with pytest.raises(DcpseError):
            set_apply_accumulation(bits)
    set_apply_accumulation(64)
