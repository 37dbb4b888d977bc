"""Pins the CPU oracle (oracle/dcpse_oracle.cpp) before it is trusted as the checker.

Every check is independent of the oracle's own code path:
  * known answers from the reference's unit tests
    (src/DCPSE/tests/MonomialBasis_unit_tests.cpp, DcpseRhs_unit_tests.cpp,
    Vandermonde_unit_tests.cpp, Dcpse_unit_tests.cpp);
  * closed forms (lattice eps, basis sizes, b = (-1)^|beta| beta! e_beta);
  * invariants that fix the kernel uniquely: the moment conditions
    V^T K = b and K in range(E^2 V), re-derived here with an SVD
    minimum-norm solve (numpy) instead of the oracle's pivoted QR on A = B^T B;
  * brute-force neighbour search.
A plausible slip (dropped E^2, sign of x_pq, wrong eps factor, transposed V,
wrong RHS sign) fails at least one of them.
"""
import math

import numpy as np
import pytest


# ---- MonomialBasis_unit_tests.cpp --------------------------------------------------
def test_monomial_evaluate_known_answers(oracle):
    # MonomialBasis_unit_tests.cpp:55-71
    assert oracle.monomial_eval([1, 1, 1], [5, 3, 2]) == pytest.approx(30.0, rel=1e-15)
    assert oracle.monomial_eval([5, 3, 2], [2, 2, 2]) == pytest.approx(2.0 ** 5 * 2 ** 3 * 2 ** 2, rel=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_monomial_derivative_known_answers(oracle, axis):
    # MonomialBasis_unit_tests.cpp:73-113: d^3/dx^3 x^6 -> 6*5*4 x^3; over-differentiation -> scalar 0
    beta = [0, 0, 0]
    beta[axis] = 3
    e, s = oracle.monomial_derivative([6, 6, 6], beta)
    want = [6, 6, 6]
    want[axis] = 3
    assert list(e) == want and s == 120
    beta[axis] = 7
    e, s = oracle.monomial_derivative([6, 6, 6], beta)
    want[axis] = 0
    assert list(e) == want and s == 0


def test_basis_membership_known_answers(oracle):
    # MonomialBasis_unit_tests.cpp:115-168
    b = {tuple(x) for x in oracle.basis([1, 0], 2)}
    assert {(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (0, 2)} <= b
    b = [tuple(x) for x in oracle.basis([1, 1], 2)]
    assert sorted(b) == sorted([(1, 0), (0, 1), (1, 1), (2, 0), (0, 2), (2, 1), (1, 2), (3, 0), (0, 3)])


@pytest.mark.parametrize("dim,beta,order", [(2, (1, 0), 2), (2, (2, 0), 2), (2, (1, 1), 3), (3, (1, 0, 0), 2),
                                            (3, (2, 0, 0), 2), (3, (0, 0, 0), 2), (2, (2, 2), 2), (3, (3, 0, 0), 1)])
def test_basis_closed_form(oracle, dim, beta, order):
    """Size = #{alpha : amin <= |alpha| < |beta|+r} (stars and bars), dim-0-fastest order."""
    s = sum(beta)
    L = s + order
    amin = 1 if (s % 2 == 0 and s > 0) else 0
    count = sum(math.comb(k + dim - 1, dim - 1) for k in range(amin, L))
    B = oracle.basis(beta, order)
    assert B.shape == (count, dim)
    assert all(amin <= e.sum() < L for e in B)
    # enumeration order of the reference grid iterator: linear index with dimension 0 fastest
    lin = [sum(int(e[d]) * L ** d for d in range(dim)) for e in B]
    assert lin == sorted(lin)


# ---- DcpseRhs_unit_tests.cpp -----------------------------------------------------------
@pytest.mark.parametrize("beta,sign,count", [((1, 0), -1, 6), ((1, 1), 1, 9), ((2, 2), 1, 20)])
def test_rhs_known_answers(oracle, beta, sign, count):
    """b_i = sign * (d^beta m_i)(0) with sign -1 for odd |beta| (DcpseRhs_unit_tests.cpp:16-109),
    i.e. (-1)^|beta| beta! at alpha == beta and 0 elsewhere."""
    b = oracle.rhs(beta, 2)
    B = oracle.basis(beta, 2)
    assert b.size == count
    fact = np.prod([math.factorial(x) for x in beta])
    for i, e in enumerate(B):
        want = sign * fact if tuple(e) == tuple(beta) else 0.0
        assert b[i] == want


# ---- Vandermonde_unit_tests.cpp ---------------------------------------------------------
def test_vandermonde_rows_known_answers(oracle):
    B = oracle.basis([1, 0], 2)
    # x = (2,2)/eps=2 -> all ones (Vandermonde_unit_tests.cpp:19-34)
    assert np.all(oracle.vandermonde_row(B, np.array([2.0, 2.0]) / 2.0) == 1.0)
    # x = (1,0): 1 where the monomial has no y (:36-53); x = (0,1): 1 where no x (:55-72)
    r = oracle.vandermonde_row(B, np.array([1.0, 0.0]))
    assert np.all(r == (B[:, 1] == 0))
    r = oracle.vandermonde_row(B, np.array([0.0, 1.0]))
    assert np.all(r == (B[:, 0] == 0))


# ---- eps on lattices (Vandermonde.hpp:63-81 with c = 0.9) -------------------------------
@pytest.mark.parametrize("dim,rc", [(2, 3.1), (2, 3.9), (3, 2.4), (3, 3.1)])
def test_lattice_eps_closed_form(oracle, dim, rc):
    n = 24 if dim == 2 else 14
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(*([g] * dim), indexing="ij")], 1)
    rp, col = oracle.support(pos, rc * h, [0] * dim, [1] * dim, (1 << dim) - 1)
    # lattice vectors strictly inside rc (self excluded), counted and l1-summed directly
    r = int(rc) + 1
    vecs = [v for v in np.ndindex(*([2 * r + 1] * dim))]
    vecs = np.array(vecs) - r
    d2 = (vecs ** 2).sum(1)
    inside = (d2 > 0) & (d2 < rc ** 2)
    nn = int(inside.sum())
    l1 = float(np.abs(vecs[inside]).sum()) / nn
    assert np.all(np.diff(rp) == nn)
    beta = (1,) + (0,) * (dim - 1)
    a = oracle.assemble(pos, rp, col, beta, 2, [0] * dim, [1] * dim, (1 << dim) - 1)
    np.testing.assert_allclose(a.eps, l1 / 0.9 * h, rtol=1e-12)
    assert {28: 2.857142857142857, 56: 2.738095238095238}.get(nn, l1 / 0.9) == pytest.approx(l1 / 0.9)


# ---- kernels: moment invariant + independent SVD min-norm recomputation ------------------
def _jitter_cloud(dim, n, seed):
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    g = (np.arange(n) + 0.5) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(*([g] * dim), indexing="ij")], 1)
    pos = np.mod(pos + rng.uniform(-0.2 * h, 0.2 * h, pos.shape), 1.0)
    return pos, h


def _min_image(d):
    return d - np.round(d)


@pytest.mark.parametrize("dim,beta", [(2, (1, 0)), (2, (0, 2)), (2, (1, 1)), (3, (0, 0, 1)), (3, (2, 0, 0))])
def test_kernel_moments_and_svd_min_norm(oracle, dim, beta):
    n = 20 if dim == 2 else 9
    pos, h = _jitter_cloud(dim, n, 7)
    rc = (3.1 if dim == 2 else 2.4) * h
    per = (1 << dim) - 1
    rp, col = oracle.support(pos, rc, [0] * dim, [1] * dim, per)
    a = oracle.assemble(pos, rp, col, beta, 2, [0] * dim, [1] * dim, per)
    B = oracle.basis(beta, 2)
    b = oracle.rhs(beta, 2)
    worst_m = worst_k = 0.0
    for p in range(0, pos.shape[0], 7):
        q = col[rp[p]:rp[p + 1]]
        x = _min_image(pos[p] - pos[q])
        eps = np.abs(x).sum(1).mean() / 0.9
        assert a.eps[p] == pytest.approx(eps, rel=1e-13)
        z = x / eps
        V = np.prod(z[:, None, :] ** B[None, :, :], axis=2)  # (n, m)
        E = np.exp(-(z ** 2).sum(1) / 2.0)
        K = a.K[rp[p]:rp[p + 1]]
        # (i) moment conditions V^T K = b
        worst_m = max(worst_m, np.abs(V.T @ K - b).max() / np.abs(b).max())
        # (ii) independent solve: K = E u, u = minimum-norm solution of (E V)^T u = b
        u = np.linalg.pinv((E[:, None] * V).T) @ b
        Kref = E * u
        worst_k = max(worst_k, np.abs(K - Kref).max() / np.abs(Kref).max())
    assert worst_m <= 1e-10, worst_m
    assert worst_k <= 1e-9, worst_k


# ---- analytic derivatives with the reference's own tolerances ---------------------------------
def test_dcpse_2d_dx_sin_reference_tolerance(oracle):
    """Dcpse_unit_tests.cpp:36-133: 40x40 grid, spacing 1/39 on [0,1]^2 (box [0,2pi]^2,
    non-periodic), Dx of sin(x), rCut = 3.1 h, max error < 2 (2h)^2."""
    n = 40
    h = 1.0 / (n - 1)
    g = np.arange(n) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(g, g, indexing="ij")], 1)
    box = [0.0, 0.0], [2 * np.pi, 2 * np.pi]
    rp, col = oracle.support(pos, 3.1 * h, *box, 0)
    a = oracle.assemble(pos, rp, col, (1, 0), 2, *box, 0)
    d = oracle.apply(a, rp, col, np.sin(pos[:, 0]))
    assert np.abs(d - np.cos(pos[:, 0])).max() < 2 * (2 * h) ** 2


def test_dcpse_3d_dz_sin_reference_tolerance(oracle):
    """Dcpse_unit_tests.cpp:243-337: 40^3 grid, spacing 1/39, Dz of sin(z), error < (3h)^2."""
    n = 40
    h = 1.0 / (n - 1)
    g = np.arange(n) * h
    pos = np.stack([a.ravel() for a in np.meshgrid(g, g, g, indexing="ij")], 1)
    box = [0.0] * 3, [2 * np.pi] * 3
    rp, col = oracle.support(pos, 3.1 * h, *box, 0)
    a = oracle.assemble(pos, rp, col, (0, 0, 1), 2, *box, 0, dense_ev=False)
    d = oracle.apply(a, rp, col, np.sin(pos[:, 2]))
    assert np.abs(d - np.cos(pos[:, 2])).max() < (3 * h) ** 2


def test_dcpse_op_2dx_plus_dy_and_lap_reference_tolerances(oracle):
    """DCPSE_op_test_base_tests.cpp:30-101 (80x80 on [0,2pi]^2, rCut 3.9h: 2Dx+Dy of
    sin x + sin y, max error < 0.03) and :394-467 (163x163, Lap, max error < 0.3)."""
    for n, rcf, what, tol in ((80, 3.9, "grad", 0.03), (163, 3.9, "lap", 0.3)):
        h = 2 * np.pi / (n - 1)
        g = np.arange(n) * h
        pos = np.stack([a.ravel() for a in np.meshgrid(g, g, indexing="ij")], 1)
        box = [0.0, 0.0], [2 * np.pi, 2 * np.pi]
        rp, col = oracle.support(pos, rcf * h, *box, 0)
        f = np.sin(pos[:, 0]) + np.sin(pos[:, 1])
        if what == "grad":
            dx = oracle.apply(oracle.assemble(pos, rp, col, (1, 0), 2, *box, 0, dense_ev=False), rp, col, f)
            dy = oracle.apply(oracle.assemble(pos, rp, col, (0, 1), 2, *box, 0, dense_ev=False), rp, col, f)
            err = np.abs(2 * dx + dy - (2 * np.cos(pos[:, 0]) + np.cos(pos[:, 1]))).max()
        else:
            lap = sum(oracle.apply(oracle.assemble(pos, rp, col, b, 2, *box, 0, dense_ev=False), rp, col, f)
                      for b in ((2, 0), (0, 2)))
            err = np.abs(lap + f).max()
        assert err < tol, (what, err)


def test_sign_forms_agree(oracle):
    """For odd |beta|, sum_q K_pq = 0 (constant monomial, b_0 = 0), so the reference's
    (f_q + f_p) form equals the difference form used on the GPU."""
    pos, h = _jitter_cloud(2, 24, 3)
    rp, col = oracle.support(pos, 3.1 * h, [0, 0], [1, 1], 3)
    a = oracle.assemble(pos, rp, col, (1, 0), 2, [0, 0], [1, 1], 3)
    f = np.sin(2 * np.pi * pos[:, 0]) + pos[:, 1]
    ref = oracle.apply(a, rp, col, f)
    rows = np.repeat(np.arange(pos.shape[0]), np.diff(rp))
    diff = np.bincount(rows, a.K * (f[col] - f[rows]), minlength=pos.shape[0]) * a.eps_inv_pow
    assert np.abs(ref - diff).max() <= 1e-12 * np.abs(ref).max()


def test_dense_and_diagonal_ev_agree(oracle):
    pos, h = _jitter_cloud(3, 8, 1)
    rp, col = oracle.support(pos, 2.4 * h, [0] * 3, [1] * 3, 7)
    a = oracle.assemble(pos, rp, col, (2, 0, 0), 2, [0] * 3, [1] * 3, 7, dense_ev=True)
    b = oracle.assemble(pos, rp, col, (2, 0, 0), 2, [0] * 3, [1] * 3, 7, dense_ev=False)
    assert np.abs(a.K - b.K).max() <= 1e-10 * np.abs(a.K).max()


def test_polynomial_reproduction(oracle):
    """Exact for polynomials spanned by the basis (degree < |beta| + r):
    Dx(x^2 + xy + y^2) = 2x + y (degree 2), Lap(x^3 + x y^2 + 3 y^2) = 8x + 6 (degree 3)."""
    rng = np.random.default_rng(9)
    pos = rng.random((900, 2)) * 0.5 + 0.25  # interior cloud, non-periodic
    rp, col = oracle.support(pos, 0.12, [0, 0], [1, 1], 0)
    keep = np.diff(rp) >= 12
    dx = oracle.assemble(pos, rp, col, (1, 0), 2, [0, 0], [1, 1], 0)
    x, y = pos[:, 0], pos[:, 1]
    d = oracle.apply(dx, rp, col, x ** 2 + x * y + y ** 2)
    assert np.abs(d - (2 * x + y))[keep & (dx.info == 0)].max() < 1e-9
    lap = [oracle.assemble(pos, rp, col, b, 2, [0, 0], [1, 1], 0) for b in ((2, 0), (0, 2))]
    L = sum(oracle.apply(a, rp, col, x ** 3 + x * y ** 2 + 3 * y ** 2) for a in lap)
    ok = keep & (lap[0].info == 0) & (lap[1].info == 0)
    assert np.abs(L - (8 * x + 6))[ok].max() < 1e-7
    # and not exact one degree higher (the check has teeth)
    d4 = oracle.apply(dx, rp, col, x ** 2 * y)
    assert np.abs(d4 - 2 * x * y)[keep].max() > 1e-6


# ---- support vs brute force ---------------------------------------------------------------
@pytest.mark.parametrize("dim,periodic", [(2, 3), (2, 0), (3, 7), (3, 5)])
def test_support_brute_force(oracle, dim, periodic):
    rng = np.random.default_rng(dim + periodic)
    pos = rng.random((700 if dim == 2 else 900, dim))
    rc = 0.09 if dim == 2 else 0.17
    rp, col = oracle.support(pos, rc, [0] * dim, [1] * dim, periodic)
    d = pos[:, None, :] - pos[None, :, :]
    for k in range(dim):
        if periodic >> k & 1:
            d[..., k] -= np.round(d[..., k])
    d2 = (d ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    for p in range(pos.shape[0]):
        want = np.nonzero(d2[p] < rc * rc)[0]
        assert np.array_equal(col[rp[p]:rp[p + 1]], want)


@pytest.mark.parametrize("dim", [2, 3])
def test_support_open_box_outside_particles_brute_force(oracle, dim):
    """Open box with particles up to 0.05 outside it on every side: they
    fall in the edge cells, and the rows still equal brute force."""
    rng = np.random.default_rng(40 + dim)
    pos = rng.random((700 if dim == 2 else 900, dim)) * 1.1 - 0.05
    rc = 0.09 if dim == 2 else 0.17
    rp, col = oracle.support(pos, rc, [0] * dim, [1] * dim, 0)
    d2 = ((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    for p in range(pos.shape[0]):
        assert np.array_equal(col[rp[p]:rp[p + 1]], np.nonzero(d2[p] < rc * rc)[0])


def test_support_per_particle_rcut(oracle):
    rng = np.random.default_rng(4)
    pos = rng.random((600, 2))
    rc = 0.04 + 0.06 * pos[:, 0]
    rp, col = oracle.support(pos, rc, [0, 0], [1, 1], 3)
    d = pos[:, None, :] - pos[None, :, :]
    d -= np.round(d)
    d2 = (d ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    for p in range(pos.shape[0]):
        assert np.array_equal(col[rp[p]:rp[p + 1]], np.nonzero(d2[p] < rc[p] ** 2)[0])


def test_too_few_neighbours_reported(oracle):
    """Vandermonde.hpp:109-112 raises when n < m; the oracle reports info = 1."""
    pos = np.array([[0.1, 0.1], [0.12, 0.1], [0.9, 0.9]])
    rp, col = oracle.support(pos, 0.05, [0, 0], [1, 1], 0)
    a = oracle.assemble(pos, rp, col, (1, 0), 2, [0, 0], [1, 1], 0)
    assert list(a.info) == [1, 1, 1]


# ---- two particle sets: support != domain (Dcpse.hpp:111-130) ---------------------------------
def _brute_rows(pos_sup, pos_dom, rc, periodic, skip=None):
    dim = pos_sup.shape[1]
    d = pos_dom[:, None, :] - pos_sup[None, :, :]
    for k in range(dim):
        if periodic >> k & 1:
            d[..., k] -= np.round(d[..., k])
    d2 = (d ** 2).sum(-1)
    if skip is not None:
        d2[np.arange(pos_dom.shape[0]), skip] = np.inf
    r2 = np.broadcast_to(np.asarray(rc, dtype=np.float64) ** 2, (pos_dom.shape[0],))
    return [np.nonzero(d2[p] < r2[p])[0] for p in range(pos_dom.shape[0])]


@pytest.mark.parametrize("dim,periodic,per_particle", [(2, 3, False), (2, 0, True), (3, 7, True), (3, 2, False)])
def test_support_two_sets_brute_force(oracle, dim, periodic, per_particle):
    """A domain set distinct from the support set: nothing is skipped, so a
    domain point on top of a support particle lists it (distance 0), and
    cells come from the domain positions (some domain points outside the
    open box)."""
    rng = np.random.default_rng(60 + dim + periodic)
    ns, nd = (700, 500) if dim == 2 else (900, 600)
    sup = rng.random((ns, dim))
    dom = rng.random((nd, dim)) * 1.1 - 0.05
    if periodic:
        dom = np.mod(dom, 1.0)
    dom[:40] = sup[100:140]  # coincident points
    base = 0.09 if dim == 2 else 0.17
    rc = base * (0.6 + 0.8 * rng.random(nd)) if per_particle else base
    rp, col = oracle.support(sup, rc, [0] * dim, [1] * dim, periodic, pos_dom=dom)
    want = _brute_rows(sup, dom, rc, periodic)
    assert rp[-1] == sum(len(w) for w in want)
    for p in range(nd):
        assert np.array_equal(col[rp[p]:rp[p + 1]], want[p])
    assert all(100 + p in col[rp[p]:rp[p + 1]] for p in range(40))


@pytest.mark.parametrize("dim,periodic", [(2, 3), (3, 0)])
def test_support_subset_brute_force(oracle, dim, periodic):
    """A domain that is a subset of the support set (vector_dist_subset,
    DCPSE_op_subset_test.cpp:113-141): each row skips its own particle and
    nothing else."""
    rng = np.random.default_rng(80 + dim)
    n = 800 if dim == 2 else 900
    pos = rng.random((n, dim))
    ids = np.sort(rng.choice(n, size=n // 2, replace=False)).astype(np.int32)
    rc = 0.09 if dim == 2 else 0.17
    rp, col = oracle.support(pos, rc, [0] * dim, [1] * dim, periodic, dom_ids=ids)
    want = _brute_rows(pos, pos[ids], rc, periodic, skip=ids)
    for p in range(ids.size):
        assert np.array_equal(col[rp[p]:rp[p + 1]], want[p])
    # the subset's rows are exactly those rows of the one-set list
    rpf, colf = oracle.support(pos, rc, [0] * dim, [1] * dim, periodic)
    for p, i in enumerate(ids):
        assert np.array_equal(col[rp[p]:rp[p + 1]], colf[rpf[i]:rpf[i + 1]])


@pytest.mark.parametrize("mode,beta", [("separate", (1, 0)), ("separate", (0, 2)), ("subset", (1, 1)),
                                       ("subset", (2, 0))])
def test_kernel_moments_two_sets(oracle, mode, beta):
    """Two-set assembly: x_pq = x_p(domain) - x_q(support). The moment
    conditions V^T K = b hold and K equals the SVD minimum-norm solution of
    the weighted system, as for one set."""
    pos, h = _jitter_cloud(2, 24, 17)
    rng = np.random.default_rng(5)
    rc = 3.1 * h
    if mode == "separate":
        dom = np.mod(pos[rng.choice(pos.shape[0], 150, replace=False)] + rng.uniform(-0.5 * h, 0.5 * h, (150, 2)),
                     1.0)
        rp, col = oracle.support(pos, rc, [0, 0], [1, 1], 3, pos_dom=dom)
    else:
        ids = np.sort(rng.choice(pos.shape[0], 150, replace=False)).astype(np.int32)
        dom = pos[ids]
        rp, col = oracle.support(pos, rc, [0, 0], [1, 1], 3, dom_ids=ids)
    a = oracle.assemble(pos, rp, col, beta, 2, [0, 0], [1, 1], 3, pos_dom=dom)
    assert (a.info == 0).all()
    B = oracle.basis(beta, 2)
    b = oracle.rhs(beta, 2)
    worst_m = worst_k = 0.0
    for p in range(dom.shape[0]):
        q = col[rp[p]:rp[p + 1]]
        x = _min_image(dom[p] - pos[q])
        eps = np.abs(x).sum(1).mean() / 0.9
        assert a.eps[p] == pytest.approx(eps, rel=1e-13)
        z = x / eps
        V = np.prod(z[:, None, :] ** B[None, :, :], axis=2)
        E = np.exp(-(z ** 2).sum(1) / 2.0)
        K = a.K[rp[p]:rp[p + 1]]
        worst_m = max(worst_m, np.abs(V.T @ K - b).max() / np.abs(b).max())
        u = np.linalg.pinv((E[:, None] * V).T) @ b
        worst_k = max(worst_k, np.abs(K - E * u).max() / np.abs(E * u).max())
    assert worst_m <= 1e-10, worst_m
    assert worst_k <= 1e-9, worst_k


def test_apply_two_sets_reference_form(oracle):
    """Two-set apply: Df_p = eps_p^-|b| sum_q (f_sup[q] + s f_dom[p]) K_pq
    (Dcpse.hpp:467-499 with the field read from the support set at q and
    from the domain set at p); exact for polynomials of the basis."""
    pos, h = _jitter_cloud(2, 24, 19)
    rng = np.random.default_rng(6)
    dom = 0.25 + 0.5 * rng.random((120, 2))
    rp, col = oracle.support(pos, 3.1 * h, [0, 0], [1, 1], 0, pos_dom=dom)
    fx = lambda P: P[:, 0] ** 2 + 3 * P[:, 0] * P[:, 1] - P[:, 1] ** 2  # noqa: E731
    for beta, want in (((1, 0), 2 * dom[:, 0] + 3 * dom[:, 1]), ((0, 1), 3 * dom[:, 0] - 2 * dom[:, 1])):
        a = oracle.assemble(pos, rp, col, beta, 2, [0, 0], [1, 1], 0, pos_dom=dom)
        got = oracle.apply(a, rp, col, fx(pos), f_dom=fx(dom))
        assert np.abs(got - want).max() < 1e-9
    a = oracle.assemble(pos, rp, col, (2, 0), 2, [0, 0], [1, 1], 0, pos_dom=dom)
    got = oracle.apply(a, rp, col, fx(pos), f_dom=fx(dom))
    assert np.abs(got - 2.0).max() < 1e-7
