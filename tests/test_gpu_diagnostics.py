"""Operator diagnostics (dcpse_operator_diagnostics) against numpy on the same
kernels: the moment residual (checkMomenta, src/DCPSE/Dcpse.hpp:381-457), the
smallest neighbour distance and <h> (DCPSE_VERBOSE, :786-795,
src/DCPSE/Vandermonde.hpp:63-81) and the 1-norm condition estimate of each
particle's system A = (E V)^T (E V).

Tolerances: residual within 1e-12 of numpy's from the same K (and <= 1e-9 for
an assembled operator); min spacing and the neighbour statistics to 1e-14;
the condition estimate is a lower bound: cond1 <= (1 + 1e-6) cond_1(A) for
every row and >= cond_1(A) / 10 for every row, exact (within 1e-6) for at
least half of them (the estimator's usual behaviour: within a factor 3,
mostly exact).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _numpy_rows(pos, pos_dom, rp, col, K, eps, basis, b, period=None):
    n_dom = rp.size - 1
    res = np.zeros(n_dom)
    cond = np.zeros(n_dom)
    msp = np.zeros(n_dom)
    for p in range(n_dom):
        q = col[rp[p]:rp[p + 1]]
        x = pos_dom[p] - pos[q]
        if period is not None:
            for k, L in enumerate(period):
                if L:
                    x[:, k] -= L * np.round(x[:, k] / L)
        msp[p] = np.sqrt((x ** 2).sum(1)).min()
        z = x / eps[p]
        V = np.prod(z[:, None, :] ** basis[None, :, :], axis=2)
        res[p] = np.abs(V.T @ K[rp[p]:rp[p + 1]] - b).max() / np.abs(b).max()
        A = (np.exp(-(z ** 2).sum(1))[:, None] * V).T @ V
        cond[p] = np.linalg.cond(A, 1) if q.size >= basis.shape[0] else np.inf
    return res, cond, msp


@pytest.mark.parametrize("case", ["2d_periodic", "2d_open_perturbed", "3d_periodic"])
def test_diagnostics_match_numpy(case, oracle, cuda):
    from openfpm_numerics_amd import Derivative_x, Laplacian, ParticleSet, workloads as W

    if case == "2d_periodic":
        w = W.jittered_2d(48, seed=5)
        pos, lo, hi, per, rc = w["pos"], [0, 0], [1, 1], 3, w["rcut"]
    elif case == "3d_periodic":
        w = W.jittered_3d_holes(14, seed=5)
        pos, lo, hi, per, rc = w["pos"], [0, 0, 0], [1, 1, 1], 7, w["rcut"]
    else:
        n = 40
        h = 1.0 / (n - 1)
        rng = np.random.default_rng(6666666)
        g = np.arange(n) * h
        X, Y = np.meshgrid(g, g, indexing="ij")
        pos = np.stack([X.ravel(), Y.ravel()], 1) + rng.normal(0.0, h / 2, (n * n, 2))
        lo, hi, per, rc = [0, 0], [2 * np.pi, 2 * np.pi], 0, 4 * h
    dim = pos.shape[1]
    ps = ParticleSet(pos, lo, hi, bc=[(per >> d) & 1 for d in range(dim)], device=cuda)
    vl = ps.getVerlet(rc)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    period = [float(hi[d] - lo[d]) if per >> d & 1 else 0.0 for d in range(dim)]
    for op in (Derivative_x(ps, vl, 2, rc), Laplacian(ps, vl, 2, rc)):
        d = op.dcpse
        summary, rows = d.diagnostics()
        K = d.calcKernels.cpu().numpy()
        eps = d.localEps.cpu().numpy()
        basis = oracle.basis(d.signatures[0], 2)
        b = sum(c * oracle.rhs(s, 2) for s, c in zip(d.signatures, d.coeffs))
        res, cond, msp = _numpy_rows(pos, pos, rp, col, K, eps, basis, b, period)
        g_res = rows["residual"].cpu().numpy()
        g_cond = rows["cond1"].cpu().numpy()
        g_msp = rows["min_spacing"].cpu().numpy()
        print(f"[diag] {case} {type(op).__name__}: {summary}", flush=True)
        assert np.abs(g_res - res).max() <= 1e-12
        assert g_res.max() <= 1e-9
        np.testing.assert_allclose(g_msp, msp, rtol=1e-14)
        ratio = g_cond / cond
        print(f"[diag] cond1 estimate / numpy: min {ratio.min():.3f} max {ratio.max():.6f} "
              f"exact share {np.mean(np.abs(ratio - 1) < 1e-6):.3f}; max cond {cond.max():.3e}", flush=True)
        assert ratio.max() <= 1 + 1e-6 and ratio.min() >= 0.1
        assert np.median(ratio) >= 1 - 1e-6
        lens = np.diff(rp)
        assert summary["neighbours_min"] == lens.min() and summary["neighbours_max"] == lens.max()
        assert summary["neighbours_mean"] == pytest.approx(lens.mean(), rel=1e-14)
        assert summary["h_mean"] == pytest.approx(0.9 * eps.mean(), rel=1e-12)
        assert summary["min_spacing_mean"] == pytest.approx(msp.mean(), rel=1e-12)
        assert summary["min_spacing_min"] == pytest.approx(msp.min(), rel=1e-14)
        assert summary["residual_max"] == pytest.approx(g_res.max(), rel=1e-14)
        assert summary["cond1_max"] == pytest.approx(g_cond.max(), rel=1e-14)
        assert summary["not_factored"] == 0


def test_diagnostics_flag_rows_without_a_kernel(cuda):
    """A particle with fewer neighbours than monomials (assembled with
    check=False) has an infinite condition estimate and is counted."""
    from openfpm_numerics_amd import Dcpse, ParticleSet

    n = 40
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(6666666)
    g = np.arange(n) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel()], 1) + rng.normal(0.0, h * h / 4, (n * n, 2))
    ps = ParticleSet(pos, [0, 0], [2 * np.pi, 2 * np.pi], bc=[0, 0], device=cuda)
    rc = 3.0 * h  # a corner particle keeps 8 neighbours: too few for the 9 monomials of d2/dxdy
    vl = ps.getVerlet(rc)
    d = Dcpse(ps, vl, (1, 1), 2, rc, opt=1)  # LOAD: no assembly in the constructor
    d.initializeStaticSize(ps, ps, check=False)
    bad = int((d.info != 0).sum())
    assert bad >= 1
    summary, rows = d.diagnostics()
    assert summary["not_factored"] == bad
    assert torch.isinf(rows["cond1"][d.info != 0]).all()
