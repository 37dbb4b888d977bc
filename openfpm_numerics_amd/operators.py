"""DC-PSE differential operators with the reference's operator-class API.

Mirrors src/DCPSE/DCPSE_op/DCPSE_op.hpp: ``Derivative_x`` (:649-786),
``Derivative_y`` (:799-...), ``Derivative_z``, the second-order
``Derivative_xx/yy/zz/xy/yz/xz``, ``Laplacian`` (:1350-1487) and the generic
``Derivative_G`` (any signature), each constructed as
``Op(parts, verletList, ord, rCut, opt=RADIUS)`` (or with a separate support
and domain set) and applied as ``Op(field)``; ``update``, ``save``, ``load``,
``deallocate``, ``checkMomenta`` and ``DrawKernel`` forward to the underlying
``Dcpse`` exactly as the reference classes forward to their ``Dcpse_type``.

The Laplacian differs in structure, not in result: the reference holds ``dim``
separate Dcpse objects with beta = 2 e_d and sums their outputs per particle
(VECT_DCPSE_V_SUM, DCPSE_op.hpp:389-401). All ``dim`` terms share one basis,
one support and one eps, so here they are assembled as one operator whose
kernel is the sum of the ``dim`` kernels (one Cholesky, summed right-hand
side) and applied in one weight stream.
"""
from __future__ import annotations

import torch

from .dcpse import RADIUS, Dcpse, apply_ops
from .particles import ParticleSet, VerletList


class _DcpseOperator:
    signatures: list = []
    coeffs = None

    def __init__(self, parts: ParticleSet, verletList: VerletList, ord: int, rCut, opt: int = RADIUS, *,
                 particlesDomain: ParticleSet | None = None, signature=None, **kw):
        dim = parts.dims
        sigs = self._signatures(dim) if signature is None else [tuple(signature)]
        self.dcpse = Dcpse(parts, verletList, sigs[0], ord, rCut, opt, particlesDomain=particlesDomain,
                           signatures=sigs, coeffs=self.coeffs if signature is None else None, **kw)

    def _signatures(self, dim):
        out = []
        for s in self.signatures:
            if len(s) > dim and any(s[dim:]):
                raise ValueError(f"{type(self).__name__} needs {len(s)} dimensions")
            out.append(tuple(s[:dim]) + (0,) * max(0, dim - len(s)))
        return out

    def __call__(self, f: torch.Tensor, f_domain: torch.Tensor | None = None) -> torch.Tensor:
        return self.dcpse.computeDifferentialOperator(f, f_domain=f_domain)

    def apply(self, f, out=None, alpha=1.0, beta=0.0):
        return self.dcpse.computeDifferentialOperator(f, out=out, alpha=alpha, beta=beta)

    def applyTranspose(self, g, out=None, alpha=1.0, beta=0.0):
        return self.dcpse.applyTranspose(g, out=out, alpha=alpha, beta=beta)

    def export_csr(self, dtype=torch.float64):
        return self.dcpse.export_csr(dtype)

    def to_sparse_csr(self, dtype=torch.float64):
        return self.dcpse.to_sparse_csr(dtype)

    # forwarders (DCPSE_op.hpp:733-784)
    def update(self, parts: ParticleSet):
        self.dcpse.initializeUpdate(parts)

    def save(self, parts: ParticleSet, file: str):
        self.dcpse.save(file)

    def load(self, parts: ParticleSet, file: str):
        self.dcpse.load(file)

    def deallocate(self, parts: ParticleSet | None = None):
        self.dcpse = None

    def checkMomenta(self, parts: ParticleSet | None = None):
        return self.dcpse.checkMomenta(parts)

    def DrawKernel(self, prop: torch.Tensor, parts: ParticleSet, k: int):
        return self.dcpse.DrawKernel(prop, k)

    def DrawKernelNN(self, prop: torch.Tensor, parts: ParticleSet, k: int):
        return self.dcpse.DrawKernelNN(prop, k)


def _op(name, sigs, doc):
    return type(name, (_DcpseOperator,), {"signatures": sigs, "__doc__": doc})


Derivative_x = _op("Derivative_x", [(1, 0, 0)], "d/dx (beta = e_0).")
Derivative_y = _op("Derivative_y", [(0, 1, 0)], "d/dy (beta = e_1).")
Derivative_z = _op("Derivative_z", [(0, 0, 1)], "d/dz (beta = e_2).")
Derivative_xx = _op("Derivative_xx", [(2, 0, 0)], "d2/dx2.")
Derivative_yy = _op("Derivative_yy", [(0, 2, 0)], "d2/dy2.")
Derivative_zz = _op("Derivative_zz", [(0, 0, 2)], "d2/dz2.")
Derivative_xy = _op("Derivative_xy", [(1, 1, 0)], "d2/dxdy.")
Derivative_yz = _op("Derivative_yz", [(0, 1, 1)], "d2/dydz.")
Derivative_xz = _op("Derivative_xz", [(1, 0, 1)], "d2/dxdz.")


class Laplacian(_DcpseOperator):
    """sum_d d2/dx_d2, assembled as one operator (see module docstring)."""

    def _signatures(self, dim):
        return [tuple(2 if k == d else 0 for k in range(dim)) for d in range(dim)]


class Derivative_G(_DcpseOperator):
    """Generic derivative for a signature given at construction (DCPSE_op.hpp Derivative_G_T)."""

    def __init__(self, parts, verletList, ord, rCut, signature, opt=RADIUS, **kw):
        super().__init__(parts, verletList, ord, rCut, opt, signature=signature, **kw)


class Gradient:
    """(d/dx_0, ..., d/dx_{dim-1}) built from dim first-derivative operators."""

    def __init__(self, parts, verletList, ord, rCut, opt=RADIUS, **kw):
        dim = parts.dims
        self.ops = [Derivative_G(parts, verletList, ord, rCut, tuple(1 if k == d else 0 for k in range(dim)),
                                 opt, **kw) for d in range(dim)]

    def __call__(self, f):
        return torch.stack(apply_ops([o.dcpse for o in self.ops], f), dim=-2)


def fused(*ops, f: torch.Tensor):
    """Applies operators that share one Verlet list in a single HBM pass."""
    return apply_ops([o.dcpse if isinstance(o, _DcpseOperator) else o for o in ops], f)
