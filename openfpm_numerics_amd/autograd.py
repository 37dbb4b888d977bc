"""torch.autograd binding of the DC-PSE apply and its hand-written adjoint.

Forward: the CSR gather kernel (dcpse_apply_f32/f64). Backward: the transposed
operator applied by the same gather kernel over a transposed CSR built once per
assembly (dcpse_transpose_*), so the gradient needs no float atomics and is
deterministic. The transposed operator is the transpose of the sparse matrix
the reference exports row by row through value_nz
(src/DCPSE/DCPSE_op/DCPSE_op.hpp:69-83) into its implicit solvers
(src/DCPSE/DCPSE_op/DCPSE_Solver.hpp:978-1036).
"""
from __future__ import annotations

import torch

from .dcpse import Dcpse


class _DcpseApply(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f: torch.Tensor, op: Dcpse):
        ctx.op = op
        return op.computeDifferentialOperator(f.detach())

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        return ctx.op.applyTranspose(grad_out.contiguous()), None


def dcpse_apply(op, f: torch.Tensor) -> torch.Tensor:
    """Differentiable D f; ``op`` is a Dcpse or an operator object from operators.py."""
    d = op.dcpse if hasattr(op, "dcpse") else op
    return _DcpseApply.apply(f, d)
