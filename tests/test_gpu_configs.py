"""The small-to-full-size parity cases of the DC-PSE path, on the GPU.

  lattice 64^2 periodic, FP64: Dx, Dy, Laplacian against the analytic
      derivatives of sin 2 pi x sin 2 pi y, and second-order convergence
      (error ratio 64^2 -> 128^2 close to 4).
  jittered 512^2 (seed 2): FP64 weights of Dx, Dy and the Laplacian against the
      oracle (row-relative <= 1e-10); FP32 apply against the FP64 oracle applied
      to the FP32-rounded input (rel. L2 <= 1e-5); FP64 apply vs analytic.
  jittered 1024^2 (seed 4), 16 FP32 fields: L = sum_b ||Lap f_b - g_b||^2 with
      f_b ~ N(0,1), g_b = Lap(sin field) + 0.01 N(0,1); dL/df from the
      transpose kernel through torch.autograd against central differences
      (step 1e-3, 64 random coordinates; FP64 gradient within 1e-6 relative),
      FP32 gradient against the FP64 one (rel. L2 <= 1e-5), and the FP64
      transpose against the oracle's assembled operator transposed on the CPU.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(w, cuda):
    from openfpm_numerics_amd import ParticleSet

    bc = [(w["periodic"] >> d) & 1 for d in range(len(w["box_lo"]))]
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=bc, device=cuda)
    return ps, ps.getVerlet(w["rcut"])


def _row_rel(Kg, Kr, row_ptr):
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    num = np.zeros(n)
    den = np.zeros(n)
    np.maximum.at(num, rows, np.abs(Kg - Kr))
    np.maximum.at(den, rows, np.abs(Kr))
    return float(np.max(num / np.maximum(den, 1e-300)))


def _analytic_errors(n, cuda):
    from openfpm_numerics_amd import Derivative_x, Derivative_y, Laplacian, workloads as W

    w = W.lattice_2d_periodic(n)
    ps, vl = _setup(w, cuda)
    f, dx, dy, lap = W.sin_field(w["pos"])
    ft = torch.as_tensor(f, device=cuda)
    out = {}
    for name, op, ex in (("dx", Derivative_x, dx), ("dy", Derivative_y, dy), ("lap", Laplacian, lap)):
        got = op(ps, vl, 2, w["rcut"])(ft).cpu().numpy()
        out[name] = float(np.max(np.abs(got - ex)) / np.max(np.abs(ex)))
    return out


def test_lattice_64_analytic_and_second_order(cuda):
    e64 = _analytic_errors(64, cuda)
    e128 = _analytic_errors(128, cuda)
    print(f"[lattice] 64^2 {e64}  128^2 {e128}", flush=True)
    for k in e64:
        assert e64[k] <= 0.02, (k, e64[k])
        ratio = e64[k] / e128[k]
        assert 3.5 <= ratio <= 4.5, (k, ratio)


def test_jittered_512_fp64_weights_and_fp32_error(oracle, cuda):
    from openfpm_numerics_amd import Derivative_x, Derivative_y, Laplacian, workloads as W

    w = W.jittered_2d(512, seed=2)
    ps, vl = _setup(w, cuda)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    print(f"[512] N={ps.size_local()} nnz={vl.nnz}", flush=True)
    f, dx, dy, lap = W.sin_field(w["pos"])
    f32 = f.astype(np.float32)
    ft32 = torch.as_tensor(f32, device=cuda)
    asm = lambda sig: oracle.assemble(w["pos"], rp, col, sig, 2, w["box_lo"], w["box_hi"], w["periodic"])  # noqa
    for name, op_cls, sigs, ex in (("dx", Derivative_x, [(1, 0)], dx), ("dy", Derivative_y, [(0, 1)], dy),
                                   ("lap", Laplacian, [(2, 0), (0, 2)], lap)):
        op = op_cls(ps, vl, 2, w["rcut"])
        refs = [asm(s) for s in sigs]
        if len(refs) == 1:
            err_k = _row_rel(op.dcpse.calcKernels.cpu().numpy(), refs[0].K, rp)
            assert err_k <= 1e-10, (name, err_k)
        want64 = sum(oracle.apply(r, rp, col, f) for r in refs)
        got64 = op(torch.as_tensor(f, device=cuda)).cpu().numpy()
        e64 = float(np.max(np.abs(got64 - want64)) / np.max(np.abs(want64)))
        assert e64 <= 1e-10, (name, e64)
        want32 = sum(oracle.apply(r, rp, col, f32.astype(np.float64)) for r in refs)
        got32 = op(ft32).cpu().numpy().astype(np.float64)
        rel = float(np.linalg.norm(got32 - want32) / np.linalg.norm(want32))
        mx = float(np.max(np.abs(got32 - want32)) / np.max(np.abs(want32)))
        ana = float(np.max(np.abs(got64 - ex)) / np.max(np.abs(ex)))
        print(f"[512] {name}: FP64 vs oracle {e64:.2e}; FP32 rel L2 {rel:.2e} max {mx:.2e}; analytic {ana:.2e}",
              flush=True)
        assert rel <= 1e-5 and mx <= 1e-5, (name, rel, mx)
        assert ana <= 0.01, (name, ana)


def test_jittered_1024_batched_adjoint(oracle, cuda):
    from openfpm_numerics_amd import Laplacian, dcpse_apply, workloads as W

    w = W.jittered_2d(1024, seed=4)
    ps, vl = _setup(w, cuda)
    n = ps.size_local()
    lap = Laplacian(ps, vl, 2, w["rcut"])
    B = 16
    rng = np.random.default_rng(4)
    fb = rng.standard_normal((B, n))
    s, _, _, _ = W.sin_field(w["pos"])
    lap_s = lap(torch.as_tensor(s, device=cuda)).cpu().numpy()
    gb = lap_s[None, :] + 0.01 * rng.standard_normal((B, n))
    print(f"[1024] N={n} nnz={vl.nnz} B={B}", flush=True)
    grads = {}
    for dt in (torch.float64, torch.float32):
        f = torch.as_tensor(fb, dtype=dt, device=cuda).requires_grad_(True)
        g = torch.as_tensor(gb, dtype=dt, device=cuda)
        loss = ((dcpse_apply(lap, f) - g) ** 2).sum()
        loss.backward()
        grads[dt] = f.grad.detach().double()
    g64, g32 = grads[torch.float64], grads[torch.float32]
    rel32 = float(torch.linalg.vector_norm(g32 - g64) / torch.linalg.vector_norm(g64))
    # central differences in FP64, exact for the quadratic loss up to rounding
    f = torch.as_tensor(fb, device=cuda)
    g = torch.as_tensor(gb, device=cuda)
    h = 1e-3
    worst = 0.0
    with torch.no_grad():
        for _ in range(64):
            b, q = int(rng.integers(B)), int(rng.integers(n))
            fp = f[b].clone()
            fm = f[b].clone()
            fp[q] += h
            fm[q] -= h
            rp_ = lap(fp) - g[b]
            rm_ = lap(fm) - g[b]
            fd = float(((rp_ - rm_) * (rp_ + rm_)).sum()) / (2 * h)
            worst = max(worst, abs(float(g64[b, q]) - fd) / max(abs(fd), 1e-30))
    # the transpose against the oracle's assembled operator, transposed on the CPU
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    A = None
    for sig in ((2, 0), (0, 2)):
        Ad = oracle.value_nz(oracle.assemble(w["pos"], rp, col, sig, 2, w["box_lo"], w["box_hi"], w["periodic"]),
                             rp, col)
        A = Ad if A is None else A + Ad
    r0 = rng.standard_normal(n)
    want_t = A.T @ r0
    got_t = lap.applyTranspose(torch.as_tensor(r0, device=cuda)).cpu().numpy()
    et = float(np.max(np.abs(got_t - want_t)) / np.max(np.abs(want_t)))
    print(f"[1024] FP64 grad vs central differences max rel {worst:.2e}; FP32 grad vs FP64 rel L2 {rel32:.2e}; "
          f"FP64 transpose vs oracle {et:.2e}", flush=True)
    assert worst <= 1e-6
    assert rel32 <= 1e-5
    assert et <= 1e-10
