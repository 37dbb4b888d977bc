"""Higher convergence orders (src/DCPSE/MonomialBasis.hpp:115-156 takes any
order): bases beyond the compiled kernels run the generic assembly kernel
(one wavefront per particle, up to 64 monomials: |beta| + order <= 6 in 3D,
<= 10 in 2D), against the oracle and by polynomial reproduction.

Tolerances: FP64 weights per row <= 1e-10 or 16 cond(A) u (_dcpse_tol; high
orders make A ill-conditioned); FP64 apply <= max(1e-10, 16 max cond(A) u)
of max|ref|; derivatives of polynomials of degree < |beta| + order exact to
1e-6 relative at interior particles (the conditioning bounds the rounding).
"""
import numpy as np
import pytest
import torch
from _dcpse_tol import check_weights

pytestmark = pytest.mark.gpu


def _cloud(dim, n, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(*([g] * dim), indexing="ij")], 1)
    return pos + rng.uniform(-0.15 * h, 0.15 * h, pos.shape), h


CASES = [  # dim, signature, order, lattice edge, rcut / h, polynomial, its derivative
    (3, (2, 0, 0), 4, 14, 3.3, lambda x: x[:, 0] ** 4 * x[:, 1], lambda x: 12 * x[:, 0] ** 2 * x[:, 1]),
    (3, (1, 0, 0), 4, 14, 3.0, lambda x: x[:, 0] ** 4 + x[:, 1] * x[:, 2], lambda x: 4 * x[:, 0] ** 3),
    (2, (2, 0), 6, 40, 4.6, lambda x: x[:, 0] ** 7 * x[:, 1], lambda x: 42 * x[:, 0] ** 5 * x[:, 1]),
    (2, (1, 0), 7, 40, 4.6, lambda x: x[:, 0] ** 5 * x[:, 1] ** 2, lambda x: 5 * x[:, 0] ** 4 * x[:, 1] ** 2),
    (2, (1, 1), 4, 40, 3.6, lambda x: x[:, 0] ** 3 * x[:, 1] ** 2, lambda x: 6 * x[:, 0] ** 2 * x[:, 1]),
]


@pytest.mark.parametrize("case", range(len(CASES)), ids=[f"{c[0]}d_{''.join(map(str, c[1]))}_o{c[2]}"
                                                         for c in CASES])
def test_high_order_against_oracle(case, oracle, cuda):
    from openfpm_numerics_amd import LOAD, Dcpse, ParticleSet, basis

    dim, sig, order, n, rch, poly, dpoly = CASES[case]
    pos, h = _cloud(dim, n, 30 + case)
    rc = rch * h
    lo, hi = [0.0] * dim, [1.0] * dim
    ps = ParticleSet(pos, lo, hi, bc=[0] * dim, device=cuda)
    vl = ps.getVerlet(rc)
    rp, col = vl.row_ptr.cpu().numpy(), vl.col.cpu().numpy()
    m = basis(sig, order).shape[0]
    lens = np.diff(rp)
    print(f"[high order] {dim}D {sig} order {order}: {m} monomials, neighbours {lens.min()}..{lens.max()}",
          flush=True)
    # corner particles of the open box may have fewer than m neighbours: both
    # sides mark them (Vandermonde.hpp:109-112) and write no kernel
    d = Dcpse(ps, vl, sig, order, rc, LOAD)
    d.initializeStaticSize(ps, ps, check=False)
    ref = oracle.assemble(pos, rp, col, sig, order, lo, hi, 0, dense_ev=False)
    info = d.info.cpu().numpy()
    assert np.array_equal(info != 0, ref.info != 0) and np.array_equal(info[info != 0], ref.info[ref.info != 0])
    assert (info == 0).mean() > 0.5
    op = d
    rel, n_cond, worst = check_weights(d.calcKernels.cpu().numpy(), ref.K, rp, pos, pos, col,
                                       oracle.basis(sig, order), label=str(CASES[case][:3]))
    f = np.sin(2 * np.pi * pos[:, 0]) * np.cos(np.pi * pos[:, dim - 1])
    want = oracle.apply(ref, rp, col, f)
    got = op(torch.as_tensor(f, device=cuda)).cpu().numpy()
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"[high order] weights {rel:.2e} ({n_cond} rows by cond(A), max {worst:.2e}), apply {err:.2e}", flush=True)
    assert err <= max(1e-10, 16 * worst * 2.0 ** -53)
    # polynomial reproduction at particles whose support lies inside the box
    inner = np.all((pos > rc) & (pos < 1 - rc), axis=1) & (info == 0)
    d = op(torch.as_tensor(poly(pos), device=cuda)).cpu().numpy()
    exact = dpoly(pos)
    perr = np.abs(d - exact)[inner].max() / np.abs(exact[inner]).max()
    print(f"[high order] polynomial reproduction {perr:.2e} over {inner.sum()} interior particles", flush=True)
    assert perr <= 1e-6
