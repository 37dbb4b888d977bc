"""Row-wise FP64 weight comparison shared by the GPU parity tests.

The GPU solves each particle's normal equations A a = b by Cholesky, the
reference (and the oracle) by a column-pivoted QR of the same A
(src/DCPSE/Dcpse.hpp:755-766). Both are backward stable, so each solution
carries a forward error of order cond(A) * u (u = 2^-53), and two of them can
differ by that much whatever either does: a row whose A has cond(A) = 4e7 (a
corner particle with exactly m neighbours) is only determined to ~5e-9 by
FP64 arithmetic on A. The bar per row is therefore 1e-10, or 16 * cond(A) * u
where that is larger; cond(A) is computed with numpy for the rows above 1e-10
only.
"""
import numpy as np

U = 2.0 ** -53


def row_rel(Kg, Kr, row_ptr):
    """max_j |Kg - Kr| / max_j |Kr| for every row."""
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    num = np.zeros(n)
    den = np.zeros(n)
    np.maximum.at(num, rows, np.abs(Kg - Kr))
    np.maximum.at(den, rows, np.abs(Kr))
    return num / np.maximum(den, 1e-300)


def cond_rows(pos_sup, pos_dom, row_ptr, col, basis, rows, period=None, h_over_eps=0.9):
    """cond(A) of the normal equations of the given rows (numpy)."""
    out = np.zeros(len(rows))
    for i, p in enumerate(rows):
        q = col[row_ptr[p]:row_ptr[p + 1]]
        x = pos_dom[p] - pos_sup[q]
        if period is not None:
            per = np.asarray(period, dtype=np.float64)
            m = per > 0
            x[:, m] -= per[m] * np.round(x[:, m] / per[m])
        eps = np.abs(x).sum(1).mean() / h_over_eps
        z = x / eps
        V = np.prod(z[:, None, :] ** basis[None, :, :], axis=2)
        EV = np.exp(-(z ** 2).sum(1) / 2.0)[:, None] * V
        out[i] = np.linalg.cond(EV.T @ EV)
    return out


def check_weights(Kg, Kr, row_ptr, pos_sup, pos_dom, col, basis, period=None, label=""):
    """Asserts the per-row bar; returns (max row error, rows judged by cond(A), their max cond)."""
    rel = row_rel(Kg, Kr, row_ptr)
    over = np.nonzero(rel > 1e-10)[0]
    worst_cond = 0.0
    if over.size:
        cond = cond_rows(pos_sup, pos_dom, row_ptr, col, basis, over, period)
        worst_cond = float(cond.max())
        bad = rel[over] > 16.0 * cond * U
        assert not bad.any(), (label, [(int(p), float(rel[p]), float(c)) for p, c in zip(over[bad], cond[bad])][:5])
    return float(rel.max()), int(over.size), worst_cond


def fp32_row_bound(row_ptr, col, W, f_sup, f_dom):
    """sum_q |W_pq| |f_q - f_p| per row: an FP32 apply rounds each weight to
    FP32 (relative 2^-24), so its error in row p can reach 2^-24 times this,
    which exceeds 1e-5 of the field's scale only in rows with large weights
    of both signs (ill-conditioned supports)."""
    n = row_ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(row_ptr))
    return np.bincount(rows, np.abs(W) * np.abs(f_sup[col] - f_dom[rows]), minlength=n)


def check_fp32(got, want, row_ptr, col, W, f_sup, f_dom, label=""):
    """FP32 apply against the FP64 oracle on FP32-rounded input: relative L2
    <= 1e-5, and per row |err| <= max(1e-5 max|ref|, 4 * 2^-24 * bound)."""
    assert np.linalg.norm(got - want) / np.linalg.norm(want) <= 1e-5, label
    bar = np.maximum(1e-5 * np.abs(want).max(), 4.0 * 2.0 ** -24 * fp32_row_bound(row_ptr, col, W, f_sup, f_dom))
    bad = np.abs(got - want) > bar
    assert not bad.any(), (label, np.nonzero(bad)[0][:5])
