"""On-device RK4 of df/dt = alpha * D f (fused stage kernels).

Pinned by mathematics rather than by a restated loop: for a linear operator one
classic RK4 step is exactly the degree-4 Taylor polynomial
    f1 = (I + x A + x^2 A^2/2 + x^3 A^3/6 + x^4 A^4/24) f0,   x = alpha dt,
with A the oracle's assembled operator (value_nz semantics, CPU); a wrong
stage coefficient or weight changes a term of that polynomial. Then FP32
against FP64 on the adaptive 1M-particle set (the reference pattern:
src/OdeIntegrators/tests/OdeIntegratores_base_tests.cpp:478-577).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _small(cuda, n=40, seed=7):
    from openfpm_numerics_amd import Laplacian, ParticleSet, workloads as W

    w = W.jittered_2d(n, seed=seed)
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=[1, 1], device=cuda)
    vl = ps.getVerlet(w["rcut"])
    return w, ps, vl, Laplacian(ps, vl, 2, w["rcut"])


def _oracle_matrix(oracle, w, vl, dim=2):
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    A = None
    for d in range(dim):
        ref = oracle.assemble(w["pos"], rp, col, tuple(2 if k == d else 0 for k in range(dim)), 2,
                              w["box_lo"], w["box_hi"], w["periodic"])
        Ad = oracle.value_nz(ref, rp, col)
        A = Ad if A is None else A + Ad
    return A


def test_rk4_step_is_the_taylor_polynomial(oracle, cuda):
    from openfpm_numerics_amd import rk4

    w, ps, vl, lap = _small(cuda)
    A = _oracle_matrix(oracle, w, vl)
    f0 = np.sin(2 * np.pi * w["pos"][:, 0]) * np.cos(4 * np.pi * w["pos"][:, 1]) + 0.3
    alpha, dt = 0.01, 2e-4
    x = alpha * dt
    want = f0.copy()
    term = f0.copy()
    for k in range(1, 5):
        term = x * (A @ term) / k
        want = want + term
    got = rk4(lap, torch.as_tensor(f0, device=cuda).clone(), dt, 1, alpha=alpha).cpu().numpy()
    # the step changes f by x*|A f| ~ 1e-3 of |f|: compare the increments
    rel = np.max(np.abs((got - f0) - (want - f0))) / np.max(np.abs(want - f0))
    assert rel <= 1e-9, f"RK4 step vs Taylor polynomial: {rel:.3e}"
    # two steps compose
    want2 = want.copy()
    term = want.copy()
    for k in range(1, 5):
        term = x * (A @ term) / k
        want2 = want2 + term
    got2 = rk4(lap, torch.as_tensor(f0, device=cuda).clone(), dt, 2, alpha=alpha).cpu().numpy()
    rel2 = np.max(np.abs((got2 - f0) - (want2 - f0))) / np.max(np.abs(want2 - f0))
    assert rel2 <= 1e-9, f"two RK4 steps: {rel2:.3e}"


def test_stage_epilogue_and_aliasing(cuda):
    """out_a = y_a + ca k, out_b = y_b + cb k, in place on y_b, all layouts."""
    from openfpm_numerics_amd.ode import stage

    _, ps, _, lap = _small(cuda, n=48, seed=3)
    gen = torch.Generator(device="cpu").manual_seed(1)
    n = ps.size_local()
    for dtype, layouts, tol in ((torch.float64, ("csr",), 1e-13), (torch.float32, ("bundle16", "sell", "csr"), 2e-6)):
        u = torch.randn(n, generator=gen, dtype=torch.float64).to(cuda, dtype)
        ya = torch.randn(n, generator=gen, dtype=torch.float64).to(cuda, dtype)
        yb = torch.randn(n, generator=gen, dtype=torch.float64).to(cuda, dtype)
        k = 0.7 * lap(u.double())
        want_a = ya.double() + 0.25 * k
        want_b = yb.double() - 1.5 * k
        scale = float(k.abs().max())
        for layout in layouts:
            out_a = torch.empty_like(u)
            acc = yb.clone()
            stage(lap, u, 0.7, ya, 0.25, out_a, acc, -1.5, acc, layout=layout)
            assert float((out_a.double() - want_a).abs().max()) <= tol * scale
            assert float((acc.double() - want_b).abs().max()) <= tol * scale
            only_b = yb.clone()
            stage(lap, u, 0.7, None, 0.0, None, only_b, -1.5, only_b, layout=layout)
            assert torch.equal(only_b, acc)


def test_rk4_fp32_layouts_agree(cuda):
    from openfpm_numerics_amd import rk4

    w, ps, vl, lap = _small(cuda, n=64, seed=11)
    f0 = torch.as_tensor(np.sin(2 * np.pi * w["pos"][:, 0]) * np.sin(2 * np.pi * w["pos"][:, 1]), device=cuda)
    ref = rk4(lap, f0.clone(), 1e-4, 20, alpha=0.01)
    a = rk4(lap, f0.float().clone(), 1e-4, 20, alpha=0.01, layout="sell")
    b = rk4(lap, f0.float().clone(), 1e-4, 20, alpha=0.01, layout="csr")
    c = rk4(lap, f0.float().clone(), 1e-4, 20, alpha=0.01, layout="bundle16")
    for got in (a, b, c):
        rel = float(torch.linalg.vector_norm(got.double() - ref) / torch.linalg.vector_norm(ref))
        assert rel <= 1e-6, rel


def test_adaptive_1m_rk4_fp32_vs_fp64(cuda):
    """1M particles, h(x) = h0 (1 + 3x), per-particle rCut = 2.4 h(x_p); FP64
    Laplacian; 100 RK4 steps of df/dt = 0.01 Lap f with FP32 state on the
    device against the same integration in FP64 (rel. L2 <= 1e-5), and the
    FP64 run against the analytic decay of f = sin 2 pi y sin 2 pi z away from
    the x faces (rel. max <= 1e-3).

    x is not periodic: particles whose support crosses an x face have
    one-sided supports, and their DC-PSE Laplacian rows have eigenvalues of
    positive real part (explicit integration grows there, whatever dt). As in
    a Dirichlet problem those particles are held fixed: their rows of the
    operator are zeroed before the integration."""
    from openfpm_numerics_amd import Laplacian, ParticleSet, rk4, workloads as W

    w = W.adaptive_3d(1_000_000, seed=5)
    n = w["pos"].shape[0]
    print(f"[ode] adaptive N={n}", flush=True)
    ps = ParticleSet(w["pos"], w["box_lo"], w["box_hi"], bc=[0, 1, 1], device=cuda)
    vl = ps.getVerlet(w["rcut"])
    lens = vl.row_lengths()
    print(f"[ode] nnz={vl.nnz} neighbours {int(lens.min())}..{int(lens.max())}", flush=True)
    lap = Laplacian(ps, vl, 2, w["rcut"])
    x, rc = w["pos"][:, 0], w["rcut"]
    held = torch.as_tensor((x - rc < 0.0) | (x + rc > 1.0), device=cuda)
    nz_held = torch.repeat_interleave(held, lens)
    for dt_ in (torch.float64, torch.float32):
        lap.dcpse.weights(dt_)[nz_held] = 0  # before the sliced-ELL copy is built
    # stable step: Gershgorin bound of the difference-form operator, 2 max_p sum_q |W_pq|
    W64 = lap.dcpse.weights(torch.float64)
    rows = torch.repeat_interleave(torch.arange(n, device=cuda), lens)
    gersh = 2.0 * float(torch.zeros(n, dtype=torch.float64, device=cuda).index_add_(0, rows, W64.abs()).max())
    alpha = 0.01
    dt = min(1e-4, 0.5 * 2.78 / (alpha * gersh))
    steps = 100
    y, z = w["pos"][:, 1], w["pos"][:, 2]
    f0 = np.sin(2 * np.pi * y) * np.sin(2 * np.pi * z)
    f64 = rk4(lap, torch.as_tensor(f0, device=cuda).clone(), dt, steps, alpha=alpha)
    f32 = torch.as_tensor(f0, dtype=torch.float32, device=cuda)
    rk4(lap, f32.clone(), dt, 1, alpha=alpha)  # builds the sliced-ELL copy
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    rk4(lap, f32, dt, steps, alpha=alpha)
    ev[1].record()
    torch.cuda.synchronize()
    print(f"[ode] FP32 RK4: {steps} steps in {ev[0].elapsed_time(ev[1]):.2f} ms "
          f"({ev[0].elapsed_time(ev[1]) / (4 * steps) * 1e3:.1f} us per fused stage)", flush=True)
    rel = float(torch.linalg.vector_norm(f32.double() - f64) / torch.linalg.vector_norm(f64))
    exact = f0 * np.exp(-alpha * 8 * np.pi ** 2 * dt * steps)
    far = (x - 3 * rc > 0.0) & (x + 3 * rc < 1.0)
    ana = float(np.max(np.abs(f64.cpu().numpy() - exact)[far]) / np.max(np.abs(exact)))
    print(f"[ode] held {int(held.sum())} rows; dt={dt:.3e} FP32 vs FP64 rel L2 {rel:.3e}; "
          f"FP64 vs analytic (3 rCut from the x faces) {ana:.3e}", flush=True)
    assert torch.isfinite(f32).all() and torch.isfinite(f64).all()
    assert rel <= 1e-5
    assert ana <= 1e-3
