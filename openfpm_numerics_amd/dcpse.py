"""``Dcpse``: one DC-PSE operator, assembled and applied on the GPU.

Mirrors the interface of the reference class
``Dcpse<dim, VerletList, vector_type, vector_type2>`` (src/DCPSE/Dcpse.hpp:43-826):
constructor arguments (particles, verletList, differentialSignature,
convergenceOrder, rCut, opt) and the two-set form (:111-130), the bulk apply
``computeDifferentialOperator`` (:467-499), the row accessors ``getNumNN``,
``getCoeffNN``, ``getIndexNN``, ``getSign``, ``getEpsilonInvPrefactor``
(:507-550), ``initializeUpdate`` (:661-687), ``save``/``load`` (:302-379),
``checkMomenta`` (:381-457) and ``DrawKernel`` (:151-193).

Assembly runs the FP64 HIP kernel (dcpse_assemble) and keeps in HBM the
reference's four arrays (localEps, localEpsInvPow, kerOffsets, calcKernels) plus
the eps^-|beta| pre-scaled FP32 weights the apply kernels stream; the FP64 ones
are formed from calcKernels at the first FP64 use.
A ``Dcpse`` can carry a linear combination of signatures that share one basis
(``signatures``/``coeffs``); the Laplacian uses that to assemble
sum_d D^{2 e_d} as one operator.
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np
import torch

from . import _lib
from .particles import ParticleSet, ParticleSubset, VerletList, sell_layout, sell_weights

# support_options (src/DCPSE/SupportBuilder.hpp:13-19)
RADIUS = 0
LOAD = 1
ADAPTIVE = 2
AT_LEAST_N_PARTICLES = 3

_FILE_MAGIC = b"DCPSEHIP"

#: Single-field applies run on the sliced-ELL layouts: FP32 on the 16-bit-offset
#: bundle (one operator on one particle set) or 32-bit columns, FP64 on 32-bit
#: columns with FP64 weights (fused Laplacian + Dx at 128^3: 437 us against the
#: CSR kernel's 512 us); set False to use the CSR kernels
USE_SELL = True
#: ... from an operator's SELL_AFTER-th such apply after its assembly on; the
#: ones before run the CSR kernel. The default, 1, builds the copy at the first
#: apply, so every single-field apply of an operator runs the same kernel and
#: repeated applies (and the autograd forward) agree bit for bit. Building the
#: copy costs about 2.4 ms for 1.9M particles and 105M neighbours, what some 50
#: applies save: a caller that reassembles every few applies (moving
#: particles) can raise SELL_AFTER to skip it, and its results then change at
#: rounding level (summation order) from that apply on.
SELL_AFTER = 1
#: The 16-bit-offset bundle serves neighbours whose index offset does not fit
#: 16 bits from a per-lane exception list; above this fraction of such
#: neighbours (e.g. Morton-ordered particles: 6.7% at 128^3) the 32-bit
#: sliced-ELL layout is faster and the automatic choice takes it.
BUNDLE16_MAX_EXC = 0.01
_FILE_VERSION = 1


def _batch2d(f: torch.Tensor) -> torch.Tensor:
    """A field (N,) or batch (..., N) as (B, N), also for N = 0."""
    return f.reshape(math.prod(f.shape[:-1]), f.shape[-1])


def basis(beta, order):
    """Monomial exponents (m, dim) of the operator's basis, reference order."""
    lib = _lib.load()
    b = np.ascontiguousarray(beta, dtype=np.uint32)
    m = lib.dcpse_basis_size(b.size, b.ctypes.data, order)
    if m < 0:
        raise ValueError("dimension must be 2 or 3")
    out = np.zeros(max(m, 1) * b.size, np.uint32)
    lib.dcpse_basis(b.size, b.ctypes.data, order, out.ctypes.data, m)
    return out[: m * b.size].reshape(m, b.size)


def set_apply_accumulation(bits: int, device: int = 0):
    """Row sums of the FP32 sliced-ELL forward applies (the kernels of every
    FP32 single-field apply and of apply_ops' sliced-ELL path): 64 (default)
    accumulates the FP32 products in FP64, 32 does every operation in FP32.
    Transposed applies, RK4 stages and the CSR kernels always sum in FP64."""
    ctx = _lib.context(device)
    ctx.check(ctx.lib.dcpse_set_apply_accumulation(ctx.h, int(bits)), "dcpse_set_apply_accumulation")


def rhs(beta, order):
    lib = _lib.load()
    b = np.ascontiguousarray(beta, dtype=np.uint32)
    m = lib.dcpse_basis_size(b.size, b.ctypes.data, order)
    out = np.zeros(max(m, 1))
    lib.dcpse_rhs(b.size, b.ctypes.data, order, out.ctypes.data, m)
    return out[:m]


class Dcpse:
    #: c in eps = <h>/c (src/DCPSE/Dcpse.hpp:80)
    HOverEpsilon = 0.9

    def __init__(self, particles: ParticleSet, verletList: VerletList, differentialSignature,
                 convergenceOrder: int, rCut, opt: int = RADIUS, *, particlesDomain: ParticleSet | None = None,
                 signatures=None, coeffs=None, h_over_eps: float | None = None, refine: bool = False):
        self.particlesSupport = particles
        self.particlesDomain = particles if particlesDomain is None else particlesDomain
        self.same_set = particlesDomain is None
        self.verletList = verletList
        sig = tuple(int(x) for x in differentialSignature)
        self.signatures = [sig] if signatures is None else [tuple(int(x) for x in s) for s in signatures]
        self.coeffs = [1.0] * len(self.signatures) if coeffs is None else [float(c) for c in coeffs]
        if len(self.coeffs) != len(self.signatures):
            raise ValueError("one coefficient per signature")
        self.differentialSignature = self.signatures[0]
        self.differentialOrder = sum(self.differentialSignature)
        if any(sum(s) != self.differentialOrder for s in self.signatures):
            raise ValueError("combined signatures must share |beta|")
        self.convergenceOrder = int(convergenceOrder)
        self.rCut = rCut
        self.opt = opt
        if h_over_eps is not None:
            self.HOverEpsilon = float(h_over_eps)
        self.refine = refine
        self.monomialBasis = basis(self.differentialSignature, self.convergenceOrder)
        self.localEps = None
        self.localEpsInvPow = None
        self.calcKernels = None
        self.info = None
        self._W = {}
        self._T = {}
        self._n32 = {"fwd": 0, "adj": 0, "fwd64": 0}  # single-field applies since assembly
        self.update_ctr = particles.map_ctr
        if opt != LOAD:
            self.initializeStaticSize(self.particlesSupport, self.particlesDomain)

    # ---- assembly --------------------------------------------------------------
    @property
    def dims(self) -> int:
        return self.particlesSupport.dims

    @property
    def kerOffsets(self) -> torch.Tensor:
        return self.verletList.row_ptr

    def initializeStaticSize(self, particlesSupport: ParticleSet, particlesDomain: ParticleSet,
                             check: bool = True):
        assemble_together([self], particlesSupport, particlesDomain, check=check)

    def _outputs(self, n_dom: int, dev):
        nnz = self.verletList.nnz
        return (torch.empty(n_dom, dtype=torch.float64, device=dev), torch.empty(n_dom, dtype=torch.float64, device=dev),
                torch.empty(max(nnz, 1), dtype=torch.float64, device=dev),
                torch.empty(max(nnz, 1), dtype=torch.float32, device=dev),
                None,  # FP64 pre-scaled weights: derived at the first FP64 use (weights)
                torch.empty(n_dom, dtype=torch.int32, device=dev))

    def _adopt(self, out, map_ctr):
        eps, eip, K64, W32, W64, info = out
        nnz = self.verletList.nnz
        self.localEps, self.localEpsInvPow = eps, eip
        self.calcKernels = K64[:nnz]
        self._W = {torch.float32: W32[:nnz]}
        self._T = {}
        self._n32 = {"fwd": 0, "adj": 0, "fwd64": 0}
        self.info = info
        self.update_ctr = map_ctr

    def check_info(self):
        """Raises like the reference on the first particle whose assembly failed."""
        info, vl = self.info, self.verletList
        bad = info != 0
        if bool(bad.any()):
            first = int(torch.nonzero(bad)[0])
            code = int(info[first])
            msg = _lib.DCPSE_INFO.get(code, str(code))
            if code == 1:
                raise ValueError(f"Not enough neighbour points passed for Vandermonde matrix "
                                 f"construction! (particle {first}: {vl.getNNPart(first)} neighbours, "
                                 f"{len(self.monomialBasis)} monomials)")
            raise ArithmeticError(f"DC-PSE assembly failed at particle {first}: {msg}")

    def initializeUpdate(self, particlesSupport: ParticleSet | None = None,
                         particlesDomain: ParticleSet | None = None):
        sup = particlesSupport or self.particlesSupport
        dom = particlesDomain or (sup if self.same_set else self.particlesDomain)
        self.initializeStaticSize(sup, dom)

    # ---- apply -------------------------------------------------------------------
    def weights(self, dtype) -> torch.Tensor:
        """Pre-scaled weights W = eps_p^-|beta| K_pq in FP32 (written by the
        assembly) or FP64 (formed from calcKernels and localEpsInvPow at the
        first FP64 use, the product the assembly would have written)."""
        if dtype == torch.float64 and dtype not in self._W and torch.float32 in self._W:
            vl = self.verletList
            K = self.calcKernels
            W = torch.empty(max(vl.nnz, 1), dtype=torch.float64, device=K.device)
            _lib.context(K.device.index or 0).call("dcpse_scale_rows_f64", vl.n_rows, _lib.ptr(vl.row_ptr),
                                                   _lib.ptr(K), _lib.ptr(self.localEpsInvPow), _lib.ptr(W))
            self._W[dtype] = W[:vl.nnz]
        if dtype not in self._W:
            raise TypeError(f"no weights for {dtype}; use float32 or float64")
        return self._W[dtype]

    def sell_weights(self) -> torch.Tensor:
        """FP32 weights in the sliced-ELL layout of this operator's Verlet list."""
        if self._W.get("sell") is None:
            self._W["sell"] = sell_weights(self.verletList.row_ptr, self._W[torch.float32], self.verletList.sell(),
                                           self.verletList.n_rows)
        return self._W["sell"]

    def sell_weights64(self) -> torch.Tensor:
        """FP64 weights in the sliced-ELL layout of this operator's Verlet list."""
        if self._W.get("sell64") is None:
            vl = self.verletList
            slice_ptr, sell_row, n_slots = vl.sell_slices()
            sw = torch.empty(n_slots, dtype=torch.float64, device=vl.row_ptr.device)
            _lib.context(vl.row_ptr.device.index or 0).call(
                "dcpse_sell_fill_weights_f64", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(self.weights(torch.float64)),
                _lib.ptr(slice_ptr), _lib.ptr(sell_row), _lib.ptr(sw))
            self._W["sell64"] = sw
        return self._W["sell64"]

    def sell_bundle16(self, dtype=torch.float32):
        """This operator alone as a 16-bit-offset bundle (see sell_bundle16)
        with FP32 or FP64 weights, built once per assembly."""
        key = "bundle16" if dtype == torch.float32 else "bundle16_64"
        if self._W.get(key) is None:
            self._W[key] = sell_bundle16([self], dtype)
        return self._W[key]

    def bundle16_suits(self) -> bool:
        """Whether the 16-bit bundle is the faster FP32 layout of this operator:
        one particle set, and at most BUNDLE16_MAX_EXC of the neighbours in the
        exception lists (counted once per support)."""
        if not self.same_set:
            return False
        vl = self.verletList
        if getattr(vl, "_n_exc16", None) is None:
            sell_row = vl.sell_slices()[1]
            exc_ptr = torch.empty(sell_row.numel() + 1, dtype=torch.int64, device=vl.row_ptr.device)
            n_exc = np.zeros(1, np.int64)
            _lib.context(vl.row_ptr.device.index or 0).call(
                "dcpse_sell_bundle16_count", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(sell_row),
                _lib.ptr(exc_ptr), n_exc.ctypes.data)
            vl._n_exc16 = int(n_exc[0])
        return vl._n_exc16 <= BUNDLE16_MAX_EXC * max(vl.nnz, 1)

    def computeDifferentialOperator(self, f: torch.Tensor, out: torch.Tensor | None = None,
                                    f_domain: torch.Tensor | None = None, alpha=1.0, beta=0.0):
        """D f for a field (N,) or a batch of fields (B, N) over the support set."""
        return apply_ops([self], f, [out] if out is not None else None, f_domain, alpha, beta)[0]

    def __call__(self, f):
        return self.computeDifferentialOperator(f)

    # ---- row accessors (value_nz contract, DCPSE_op.hpp:69-83) ------------------------
    def _host_rows(self):
        if getattr(self, "_hk", None) is None or self._hk[0] is not self.calcKernels:
            self._hk = (self.calcKernels, self.calcKernels.cpu().numpy(), self.localEpsInvPow.cpu().numpy())
        return self._hk

    def getNumNN(self, p: int) -> int:
        return self.verletList.getNNPart(p)

    def getCoeffNN(self, p: int, j: int) -> float:
        _, K, _ = self._host_rows()
        rp, _ = self.verletList._mirror()
        return float(K[rp[p] + j])

    def getIndexNN(self, p: int, j: int) -> int:
        return self.verletList.get(p, j)

    def getSign(self) -> float:
        return -1.0 if (self.differentialOrder % 2 == 0 and self.differentialOrder != 0) else 1.0

    def getEpsilonInvPrefactor(self, p: int) -> float:
        return float(self._host_rows()[2][p])

    # ---- sparse export (value_nz semantics, DCPSE_op.hpp:69-83) ----------------------
    def export_csr(self, dtype=torch.float64):
        """The operator as a CSR matrix (row_ptr int64, col int32, val) on the device:
        W_pq at every neighbour q and getSign() * sum_q W_pq on the diagonal,
        columns ascending - the matrix the reference's implicit solvers assemble
        from value_nz (DCPSE_Solver.hpp:978-1036)."""
        if not self.same_set:
            raise ValueError("export needs an operator whose support and domain sets are the same")
        vl = self.verletList
        n, dev = vl.n_rows, vl.row_ptr.device
        rp = torch.empty(n + 1, dtype=torch.int64, device=dev)
        col = torch.empty(vl.nnz + n, dtype=torch.int32, device=dev)
        val = torch.empty(vl.nnz + n, dtype=dtype, device=dev)
        name = "dcpse_export_csr_f64" if dtype == torch.float64 else "dcpse_export_csr_f32"
        _lib.context(dev.index or 0).call(name, n, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(self.weights(dtype)),
                                          float(self.getSign()), _lib.ptr(rp), _lib.ptr(col), _lib.ptr(val))
        return rp, col, val

    def to_sparse_csr(self, dtype=torch.float64) -> torch.Tensor:
        """export_csr() as a torch sparse CSR tensor (N x N) on the device."""
        rp, col, val = self.export_csr(dtype)
        n = self.verletList.n_rows
        return torch.sparse_csr_tensor(rp, col.to(torch.int64), val, size=(n, n))

    # ---- adjoint -----------------------------------------------------------------
    def transposed(self, dtype):
        """(rowT_ptr, colT, WT, diag) of the transpose of the difference-form operator."""
        if dtype in self._T:
            return self._T[dtype]
        if not self.same_set:
            raise NotImplementedError("the transposed operator needs the domain set to be the support set")
        vl = self.verletList
        dev = vl.row_ptr.device
        n = vl.n_rows
        ctx = _lib.context(dev.index or 0)
        rowT = torch.empty(n + 1, dtype=torch.int64, device=dev)
        ctx.call("dcpse_transpose_count", n, vl.n_cols, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(rowT))
        colT = torch.empty(max(vl.nnz, 1), dtype=torch.int32, device=dev)
        WT = torch.empty(max(vl.nnz, 1), dtype=dtype, device=dev)
        diag = torch.empty(n, dtype=dtype, device=dev)
        name = "dcpse_transpose_fill_f32" if dtype == torch.float32 else "dcpse_transpose_fill_f64"
        ctx.call(name, n, vl.n_cols, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(self.weights(dtype)),
                 _lib.ptr(rowT), _lib.ptr(colT), _lib.ptr(WT), _lib.ptr(diag))
        self._T[dtype] = (rowT, colT[: vl.nnz], WT[: vl.nnz], diag)
        return self._T[dtype]

    def transposed_bundle16(self):
        """The FP32 transposed weights as a 16-bit bundle on this operator's own
        support (sell_bundle16's tuple), when the transposed CSR has the same
        row_ptr and col (a symmetric support: one set, one cut-off); else None."""
        if "bundle16" not in self._T:
            rowT, colT, WT, _ = self.transposed(torch.float32)
            vl = self.verletList
            same = (self.same_set and torch.equal(rowT, vl.row_ptr) and colT.numel() == vl.col.numel()
                    and torch.equal(colT, vl.col))
            self._T["bundle16"] = bundle16(vl, [WT]) if same else None
        return self._T["bundle16"]

    def transposed_sell(self):
        """SELL layout of the FP32 transposed operator: (slice_ptr, sell_rowT, sell_colT, sell_WT)."""
        if self._T.get("sell") is None:
            rowT, colT, WT, _ = self.transposed(torch.float32)
            n = self.verletList.n_cols
            layout = sell_layout(rowT, colT, n)
            self._T["sell"] = (*layout, sell_weights(rowT, WT, layout, n))
        return self._T["sell"]

    def _count_apply32(self, kind: str) -> bool:
        """Counts a single-field apply (kind "fwd" or "adj" in FP32, "fwd64" in
        FP64); True once the sliced-ELL copy is due (SELL_AFTER)."""
        self._n32[kind] += 1
        return self._n32[kind] >= SELL_AFTER

    def applyTranspose(self, g: torch.Tensor, out: torch.Tensor | None = None, alpha=1.0, beta=0.0):
        """A^T g for a field (N,) or a batch (B, N)."""
        if g.dtype not in (torch.float32, torch.float64):
            raise TypeError("fields must be float32 or float64")
        g2 = _batch2d(g).contiguous()
        nb, n = g2.shape
        if n != self.verletList.n_cols:
            raise ValueError("field length differs from the support set")
        if n == 0 or self.verletList.n_rows == 0:  # an empty set: A^T g = 0
            o = torch.zeros_like(g2) if out is None else out.reshape(nb, n).zero_()
            return o.reshape(g.shape)
        rowT, colT, WT, diag = self.transposed(g.dtype)
        o = torch.empty_like(g2) if out is None else out.reshape(nb, n)
        ctx = _lib.context(g.device.index or 0)
        if g.dtype == torch.float32 and nb == 1 and USE_SELL and self._count_apply32("adj"):
            b16 = self.transposed_bundle16() if self.bundle16_suits() else None
            if b16 is not None:
                bundle, row16, exc_ptr, exc_col, exc_W, ne = b16
                vl = self.verletList
                ctx.call("dcpse_apply_transpose_bundle16_f32", n, _lib.ptr(vl.row_ptr), _lib.ptr(vl.sell_slices()[0]),
                         _lib.ptr(row16), _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne,
                         _lib.ptr(diag), _lib.ptr(g2), _lib.ptr(o), alpha, beta)
                return o.reshape(g.shape)
            slice_ptr, sell_rowT, sell_colT, sell_WT = self.transposed_sell()
            ctx.call("dcpse_apply_transpose_sell_f32", n, _lib.ptr(rowT), _lib.ptr(slice_ptr), _lib.ptr(sell_rowT),
                     _lib.ptr(sell_colT), _lib.ptr(sell_WT), _lib.ptr(diag), _lib.ptr(g2), _lib.ptr(o), alpha, beta)
            return o.reshape(g.shape)
        name = "dcpse_apply_transpose_f32" if g.dtype == torch.float32 else "dcpse_apply_transpose_f64"
        ctx.call(name, n, _lib.ptr(rowT), _lib.ptr(colT), _lib.ptr(WT), _lib.ptr(diag), _lib.ptr(g2), n, nb,
                 _lib.ptr(o), n, alpha, beta)
        return o.reshape(g.shape)

    # ---- persistence (Dcpse.hpp:302-379) -------------------------------------------
    def save(self, file: str, rank: int = 0):
        """Writes localEps, localEpsInvPow, calcKernels, kerOffsets and the support."""
        path = f"{file}_{rank}"
        eps = self.localEps.cpu().numpy()
        eip = self.localEpsInvPow.cpu().numpy()
        K = self.calcKernels.cpu().numpy()
        rp = self.verletList.row_ptr.cpu().numpy()
        col = self.verletList.col.cpu().numpy()
        sig = np.array(self.signatures, dtype=np.uint32)
        with open(path, "wb") as fh:
            fh.write(_FILE_MAGIC)
            fh.write(struct.pack("<IIIId", _FILE_VERSION, self.dims, len(self.signatures),
                                 self.convergenceOrder, float(self.HOverEpsilon)))
            fh.write(struct.pack("<qq", eps.size, K.size))
            fh.write(sig.astype("<u4").tobytes())
            fh.write(np.array(self.coeffs, dtype="<f8").tobytes())
            for arr, dt in ((eps, "<f8"), (eip, "<f8"), (K, "<f8"), (rp, "<i8"), (col, "<i4")):
                fh.write(np.ascontiguousarray(arr).astype(dt).tobytes())

    def load(self, file: str, rank: int = 0):
        path = f"{file}_{rank}"
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        with open(path, "rb") as fh:
            if fh.read(8) != _FILE_MAGIC:
                raise ValueError(f"{path} is not a DC-PSE kernel file")
            ver, dim, nsig, order, hoe = struct.unpack("<IIIId", fh.read(24))
            if ver != _FILE_VERSION:
                raise ValueError(f"unsupported kernel file version {ver}")
            n, nnz = struct.unpack("<qq", fh.read(16))
            sig = np.frombuffer(fh.read(4 * nsig * dim), dtype="<u4").reshape(nsig, dim)
            coeffs = np.frombuffer(fh.read(8 * nsig), dtype="<f8")
            eps = np.frombuffer(fh.read(8 * n), dtype="<f8")
            eip = np.frombuffer(fh.read(8 * n), dtype="<f8")
            K = np.frombuffer(fh.read(8 * nnz), dtype="<f8")
            rp = np.frombuffer(fh.read(8 * (n + 1)), dtype="<i8")
            col = np.frombuffer(fh.read(4 * nnz), dtype="<i4")
        if [tuple(int(v) for v in s) for s in sig] != self.signatures:
            raise ValueError("kernel file holds a different differential signature")
        dev = self.particlesSupport.device
        vl = self.verletList
        # the kernels are only meaningful on the support they were assembled on
        # (the reference loads them against the Verlet list it is given,
        # Dcpse.hpp:336-379); a file for another support is refused, and the
        # Verlet list, which other operators may share, is never touched
        if vl.n_rows != n or vl.nnz != nnz:
            raise ValueError(f"kernel file support ({n} rows, {nnz} neighbours) differs from this operator's "
                             f"Verlet list ({vl.n_rows} rows, {vl.nnz} neighbours)")
        hrp, hcol = vl._mirror()
        if not np.array_equal(hrp, rp) or not np.array_equal(hcol, col):
            raise ValueError("kernel file support differs from this operator's Verlet list (neighbour lists)")
        self.convergenceOrder, self.HOverEpsilon, self.coeffs = int(order), float(hoe), list(coeffs)
        self.localEps = torch.as_tensor(eps.copy(), device=dev)
        self.localEpsInvPow = torch.as_tensor(eip.copy(), device=dev)
        self.calcKernels = torch.as_tensor(K.copy(), device=dev)
        rows = torch.repeat_interleave(torch.arange(n, device=dev), vl.row_lengths())
        W64 = self.calcKernels * self.localEpsInvPow[rows]
        self._W = {torch.float64: W64, torch.float32: W64.to(torch.float32)}
        self._T = {}
        self._n32 = {"fwd": 0, "adj": 0, "fwd64": 0}
        self.info = torch.zeros(n, dtype=torch.int32, device=dev)
        self.update_ctr = self.particlesSupport.map_ctr

    # ---- diagnostics -------------------------------------------------------------
    def checkMomenta(self, particles: ParticleSet | None = None):
        """Per-monomial min/max over particles of sum_q K_pq m_alpha(x_pq/eps).

        For a valid operator every row reproduces b_alpha (DcpseRhs.hpp:27-60).
        """
        sup = particles or self.particlesSupport
        dom = self.particlesDomain if not self.same_set else sup
        vl = self.verletList
        dev = sup.device
        rows = torch.repeat_interleave(torch.arange(vl.n_rows, device=dev), vl.row_lengths())
        cols = vl.col.long()
        d = dom.pos_soa[:, rows] - sup.pos_soa[:, cols]
        for k in range(sup.dims):
            if sup.periodic_mask >> k & 1:
                L = float(sup.box_hi[k] - sup.box_lo[k])
                d[k] -= L * torch.round(d[k] / L)
        z = d / self.localEps[rows]
        out = []
        for e in self.monomialBasis:
            mval = torch.ones_like(z[0])
            for k in range(sup.dims):
                mval = mval * z[k] ** int(e[k])
            mom = torch.zeros(vl.n_rows, dtype=torch.float64, device=dev).index_add_(0, rows, mval * self.calcKernels)
            out.append((float(mom.min()), float(mom.max())))
        return out

    def diagnostics(self):
        """Per-particle diagnostics of the assembled operator (dcpse_operator_diagnostics):
        the reference's DCPSE_VERBOSE statistics (Dcpse.hpp:786-795), checkMomenta's
        moment residual (:381-457) and a 1-norm condition estimate of each solved
        system. Returns (summary dict, {"residual", "cond1", "min_spacing"} device
        tensors of the domain particle count)."""
        sup = self.particlesSupport
        dom = self.particlesDomain
        vl = self.verletList
        dev = sup.device
        n_dom = vl.n_rows
        res = torch.empty(max(n_dom, 1), dtype=torch.float64, device=dev)
        cond = torch.empty_like(res)
        msp = torch.empty_like(res)
        summary = np.zeros(10)
        betas = np.ascontiguousarray(np.array(self.signatures, dtype=np.uint32).ravel())
        coeffs = np.ascontiguousarray(np.array(self.coeffs, dtype=np.float64))
        lo = np.ascontiguousarray(sup.box_lo)
        hi = np.ascontiguousarray(sup.box_hi)
        _lib.context(dev.index or 0).call(
            "dcpse_operator_diagnostics", sup.dims, _lib.ptr(sup.pos_soa), sup.size_local(),
            None if self.same_set else _lib.ptr(dom.pos_soa), n_dom, lo.ctypes.data, hi.ctypes.data,
            sup.periodic_mask, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col) if vl.nnz else None, len(self.signatures),
            betas.ctypes.data, coeffs.ctypes.data, self.convergenceOrder, float(self.HOverEpsilon),
            _lib.ptr(self.localEps), _lib.ptr(self.calcKernels) if vl.nnz else _lib.ptr(self.localEps),
            _lib.ptr(res), _lib.ptr(cond), _lib.ptr(msp), summary.ctypes.data)
        keys = ("neighbours_min", "neighbours_mean", "neighbours_max", "h_mean", "min_spacing_mean",
                "min_spacing_min", "min_spacing_max", "residual_max", "cond1_max", "not_factored")
        info = dict(zip(keys, (float(v) for v in summary)))
        return info, {"residual": res[:n_dom], "cond1": cond[:n_dom], "min_spacing": msp[:n_dom]}

    def DrawKernel(self, prop: torch.Tensor, p: int):
        """prop[q] += K_pq for every neighbour q of p."""
        vl = self.verletList
        rp, col = vl._mirror()
        s, e = int(rp[p]), int(rp[p + 1])
        prop.index_add_(0, vl.col[s:e].long(), self.calcKernels[s:e].to(prop.dtype))
        return prop

    def DrawKernelNN(self, prop: torch.Tensor, p: int):
        vl = self.verletList
        rp, _ = vl._mirror()
        prop[vl.col[int(rp[p]):int(rp[p + 1])].long()] = 1.0
        return prop


def assemble_together(ops, particlesSupport: ParticleSet | None = None, particlesDomain: ParticleSet | None = None,
                      check: bool = True):
    """(Re)assembles several operators on one Verlet list in one library call
    (dcpse_assemble_multi): operators whose bases pair up (the Laplacian or a
    second derivative with a first derivative, two operators of one basis)
    share the neighbour gathers, epsilon and the moments in one kernel pass."""
    if not ops:
        return
    d0 = ops[0]
    vl = d0.verletList
    for d in ops[1:]:
        if d.verletList is not vl or d.same_set != d0.same_set or d.convergenceOrder != d0.convergenceOrder \
                or d.HOverEpsilon != d0.HOverEpsilon or d.refine != d0.refine:
            raise ValueError("operators assembled together must share the Verlet list, order, h/eps and flags")
    sup = particlesSupport or d0.particlesSupport
    dom = particlesDomain or (sup if d0.same_set else d0.particlesDomain)
    dev = sup.device
    n_dom = dom.size_local()
    if vl.n_rows != n_dom:
        raise ValueError(f"Verlet list has {vl.n_rows} rows, domain has {n_dom} particles")
    outs = [d._outputs(n_dom, dev) for d in ops]
    if n_dom == 0:  # an empty domain: no rows to assemble
        for d, o in zip(ops, outs):
            d._adopt(o, sup.map_ctr)
        return
    n_sig = np.array([len(d.signatures) for d in ops], dtype=np.int32)
    betas = np.ascontiguousarray(np.concatenate([np.array(d.signatures, dtype=np.uint32).ravel() for d in ops]))
    coeffs = np.ascontiguousarray(np.concatenate([np.array(d.coeffs, dtype=np.float64) for d in ops]))
    lo = np.ascontiguousarray(sup.box_lo)
    hi = np.ascontiguousarray(sup.box_hi)
    arrays = [_lib.ptr_array([o[k] for o in outs]) for k in range(6)]
    ctx = _lib.context(dev.index or 0)
    ctx.call("dcpse_assemble_multi", sup.dims, _lib.ptr(sup.pos_soa), sup.size_local(),
             None if d0.same_set else _lib.ptr(dom.pos_soa), n_dom,
             lo.ctypes.data, hi.ctypes.data, sup.periodic_mask,
             _lib.ptr(vl.row_ptr), _lib.ptr(vl.col) if vl.nnz else None,
             len(ops), n_sig.ctypes.data, betas.ctypes.data, coeffs.ctypes.data,
             d0.convergenceOrder, float(d0.HOverEpsilon), _lib.DCPSE_ASM_REFINE if d0.refine else 0,
             arrays[0], arrays[1], arrays[2], arrays[3], arrays[4], arrays[5])
    for d, o in zip(ops, outs):
        d._adopt(o, sup.map_ctr)
    if check:
        for d in ops:
            d.check_info()


def apply_ops(ops, f: torch.Tensor, outs=None, f_domain: torch.Tensor | None = None, alpha=1.0, beta=0.0):
    """Applies several operators that share one Verlet list in a single pass.

    ``f``: (N,) or (B, N) float32/float64 over the support set. Returns one
    output per operator, shaped like ``f`` but over the domain set.
    """
    if not ops:
        return []
    vl = ops[0].verletList
    for op in ops[1:]:
        if op.verletList is not vl:
            raise ValueError("fused operators must share one Verlet list")
    if f.dtype not in (torch.float32, torch.float64):
        raise TypeError("fields must be float32 or float64")
    for op in ops:
        if op.update_ctr != op.particlesSupport.map_ctr:
            raise RuntimeError("particles moved since assembly: call update()/initializeUpdate() first")
    f2 = _batch2d(f).contiguous()
    nb, n = f2.shape
    if n != vl.n_cols:
        raise ValueError(f"field has {n} values, support set has {vl.n_cols} particles")
    fd = None
    if f_domain is not None:
        fd = f_domain.reshape(nb, -1).contiguous().to(f.dtype)
    elif isinstance(ops[0].particlesDomain, ParticleSubset):
        # a subset domain reads the field at its own particles of the support set
        fd = ops[0].particlesDomain.gather(f2).contiguous()
    elif not ops[0].same_set:
        raise ValueError("two-set operators need f_domain (the field on the domain particles)")
    n_rows = vl.n_rows
    if outs is None:
        outs = [torch.empty((nb, n_rows), dtype=f.dtype, device=f.device) for _ in ops]
    else:
        outs = [o.reshape(nb, n_rows) for o in outs]
    if n_rows == 0:  # an empty domain: nothing to compute
        return [o.reshape(*f.shape[:-1], n_rows) for o in outs]
    ctx = _lib.context(f.device.index or 0)
    use_sell = False
    if f.dtype == torch.float32 and nb == 1 and USE_SELL:
        use_sell = all([op._count_apply32("fwd") for op in ops])
    if use_sell and len(ops) == 1 and ops[0].bundle16_suits():
        # one operator: its 16-bit-offset bundle, built once per assembly
        slice_ptr = vl.sell_slices()[0]
        bundle, row16, exc_ptr, exc_col, exc_W, ne = ops[0].sell_bundle16()
        ctx.call("dcpse_apply_sell_bundle16_f32", n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(row16),
                 1, _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne, _lib.ptr(f2),
                 _lib.ptr(fd), _lib.ptr_array(outs), alpha, beta)
        return [o.reshape(*f.shape[:-1], n_rows) for o in outs]
    if use_sell:
        slice_ptr, sell_row, sell_col = vl.sell()
        SW = [op.sell_weights() for op in ops]
        ctx.call("dcpse_apply_sell_f32", n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                 _lib.ptr(sell_col), len(ops), _lib.ptr_array(SW), _lib.ptr(f2), _lib.ptr(fd), _lib.ptr_array(outs), alpha, beta)
        return [o.reshape(*f.shape[:-1], n_rows) for o in outs]
    if f.dtype == torch.float64 and nb == 1 and USE_SELL and all([op._count_apply32("fwd64") for op in ops]):
        if len(ops) == 1 and ops[0].bundle16_suits():
            # one operator: its FP64 16-bit-offset bundle (18 bytes per neighbour)
            slice_ptr = vl.sell_slices()[0]
            bundle, row16, exc_ptr, exc_col, exc_W, ne = ops[0].sell_bundle16(torch.float64)
            ctx.call("dcpse_apply_sell_bundle16_f64", n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr),
                     _lib.ptr(row16), 1, _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne,
                     _lib.ptr(f2), _lib.ptr(fd), _lib.ptr_array(outs), alpha, beta)
            return [o.reshape(*f.shape[:-1], n_rows) for o in outs]
        # FP64: the 32-bit-column layout with FP64 weights in its slot order
        slice_ptr, sell_row, sell_col = vl.sell()
        SW = [op.sell_weights64() for op in ops]
        ctx.call("dcpse_apply_sell_f64", n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                 _lib.ptr(sell_col), len(ops), _lib.ptr_array(SW), _lib.ptr(f2), _lib.ptr(fd), _lib.ptr_array(outs),
                 alpha, beta)
        return [o.reshape(*f.shape[:-1], n_rows) for o in outs]
    W = [op.weights(f.dtype) for op in ops]
    name = "dcpse_apply_f32" if f.dtype == torch.float32 else "dcpse_apply_f64"
    if len(ops) > 1 and nb > 1:
        res = [apply_ops([op], f, [o], f_domain, alpha, beta)[0] for op, o in zip(ops, outs)]
        return [r.reshape(*f.shape[:-1], n_rows) for r in res]
    ctx.call(name, n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), len(ops), _lib.ptr_array(W),
             _lib.ptr(f2), n, nb, _lib.ptr(fd), n_rows if fd is not None else n, _lib.ptr_array(outs), n_rows,
             alpha, beta)
    return [o.reshape(*f.shape[:-1], n_rows) for o in outs]


def sell_bundle(ops) -> torch.Tensor:
    """The interleaved sliced-ELL copy of several FP32 operators sharing one
    support (columns and every operator's weights per slot group)."""
    vl = ops[0].verletList
    slice_ptr, sell_row, n_slots = vl.sell_slices()
    Ws = [op.weights(torch.float32) for op in ops]
    bundle = torch.empty((1 + len(ops)) * n_slots, dtype=torch.int32, device=vl.row_ptr.device)
    _lib.context(vl.row_ptr.device.index or 0).call(
        "dcpse_sell_fill_bundle_f32", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), len(ops), _lib.ptr_array(Ws),
        _lib.ptr(slice_ptr), _lib.ptr(sell_row), _lib.ptr(bundle))
    return bundle


def sell_bundle16(ops, dtype=torch.float32):
    """The interleaved sliced-ELL copy with 16-bit column offsets (operators on
    one particle set) with FP32 or FP64 weights: (bundle, sell_row16, exc_ptr,
    exc_col, exc_W, n_exc), the far neighbours (offset beyond +-32767 modulo
    N) listed per lane."""
    if not all(op.same_set for op in ops):
        raise ValueError("16-bit column offsets need operators whose support and domain sets are the same")
    return bundle16(ops[0].verletList, [op.weights(dtype) for op in ops])


def bundle16(vl: VerletList, Ws):
    """sell_bundle16 of weight arrays Ws (all FP32 or all FP64) in the CSR
    order of vl."""
    slice_ptr, sell_row, n_slots = vl.sell_slices()
    dev = vl.row_ptr.device
    ctx = _lib.context(dev.index or 0)
    n_lanes = sell_row.numel()
    exc_ptr = torch.empty(n_lanes + 1, dtype=torch.int64, device=dev)
    n_exc = np.zeros(1, np.int64)
    ctx.call("dcpse_sell_bundle16_count", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(sell_row),
             _lib.ptr(exc_ptr), n_exc.ctypes.data)
    ne = int(n_exc[0])
    dt = Ws[0].dtype
    if dt not in (torch.float32, torch.float64) or any(W.dtype != dt for W in Ws):
        raise TypeError("bundle weights must be all float32 or all float64")
    words = 1 if dt == torch.float32 else 2  # 32-bit words per weight
    bundle = torch.empty(max(n_slots // 2 + words * len(Ws) * n_slots, 8), dtype=torch.int32, device=dev)
    sell_row16 = torch.empty_like(sell_row)
    exc_col = torch.empty(max(ne, 1), dtype=torch.int32, device=dev)
    exc_W = torch.empty(max(len(Ws) * ne, 1), dtype=dt, device=dev)
    ctx.call("dcpse_sell_fill_bundle16_f32" if dt == torch.float32 else "dcpse_sell_fill_bundle16_f64", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), len(Ws),
             _lib.ptr_array(Ws), _lib.ptr(slice_ptr), _lib.ptr(sell_row), _lib.ptr(exc_ptr), ne, _lib.ptr(bundle),
             _lib.ptr(sell_row16), _lib.ptr(exc_col), _lib.ptr(exc_W))
    return bundle, sell_row16, exc_ptr, exc_col, exc_W, ne


class PreparedApply:
    """A fused apply with every pointer bound once: calling it costs one C call.

    For timing loops and graph capture; the tensors passed in must stay alive
    and keep their storage. ``f`` is (N,) or (B, N); ``outs`` one tensor per op.
    """

    def __init__(self, ops, f: torch.Tensor, outs, alpha=1.0, beta=0.0, sell=True):
        vl = ops[0].verletList
        if any(op.verletList is not vl for op in ops):
            raise ValueError("fused operators must share one Verlet list")
        if not f.is_contiguous() or any(not o.is_contiguous() for o in outs):
            raise ValueError("contiguous tensors only")
        f2 = _batch2d(f)
        nb, n = f2.shape
        self.ctx = _lib.context(f.device.index or 0)
        self.O = _lib.ptr_array(outs)
        if sell == "bundle16" and nb == 1:
            slice_ptr = vl.sell_slices()[0]
            b16 = sell_bundle16(ops, f.dtype)
            self._keep = (ops, f, outs, b16, slice_ptr)
            bundle, sell_row16, exc_ptr, exc_col, exc_W, ne = b16
            self.n_exc = ne
            self.fn = (self.ctx.lib.dcpse_apply_sell_bundle16_f32 if f.dtype == torch.float32
                       else self.ctx.lib.dcpse_apply_sell_bundle16_f64)
            self.args = (self.ctx.h, vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row16),
                         len(ops), _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne,
                         _lib.ptr(f2), None, self.O, alpha, beta)
        elif sell == "bundle" and f.dtype == torch.float32 and nb == 1:
            slice_ptr, sell_row, _ = vl.sell_slices()
            bundle = sell_bundle(ops)
            self._keep = (ops, f, outs, bundle, slice_ptr, sell_row)
            self.fn = self.ctx.lib.dcpse_apply_sell_bundle_f32
            self.args = (self.ctx.h, vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                         len(ops), _lib.ptr(bundle), _lib.ptr(f2), None, self.O, alpha, beta)
        elif sell and f.dtype == torch.float32 and nb == 1:
            slice_ptr, sell_row, sell_col = vl.sell()
            self._keep = (ops, f, outs, [op.sell_weights() for op in ops], slice_ptr, sell_row, sell_col)
            self.W = _lib.ptr_array(self._keep[3])
            self.fn = self.ctx.lib.dcpse_apply_sell_f32
            self.args = (self.ctx.h, vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                         _lib.ptr(sell_col), len(ops), self.W, _lib.ptr(f2), None, self.O, alpha, beta)
        elif sell and f.dtype == torch.float64 and nb == 1:
            slice_ptr, sell_row, sell_col = vl.sell()
            self._keep = (ops, f, outs, [op.sell_weights64() for op in ops], slice_ptr, sell_row, sell_col)
            self.W = _lib.ptr_array(self._keep[3])
            self.fn = self.ctx.lib.dcpse_apply_sell_f64
            self.args = (self.ctx.h, vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                         _lib.ptr(sell_col), len(ops), self.W, _lib.ptr(f2), None, self.O, alpha, beta)
        else:
            self._keep = (ops, f, outs, [op.weights(f.dtype) for op in ops])
            self.fn = self.ctx.lib.dcpse_apply_f32 if f.dtype == torch.float32 else self.ctx.lib.dcpse_apply_f64
            self.W = _lib.ptr_array(self._keep[3])
            self.args = (self.ctx.h, vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), len(ops), self.W,
                         _lib.ptr(f2), n, nb, None, n, self.O, vl.n_rows, alpha, beta)
        lib = self.ctx.lib
        self.layout = next((name for fn, name in ((lib.dcpse_apply_sell_f32, "sell"),
                                                  (lib.dcpse_apply_sell_bundle_f32, "sell_bundle"),
                                                  (lib.dcpse_apply_sell_bundle16_f32, "sell_bundle16"),
                                                  (lib.dcpse_apply_sell_bundle16_f64, "sell64_bundle16"),
                                                  (lib.dcpse_apply_sell_f64, "sell64"))
                            if self.fn is fn), "csr")
        self.sell = self.layout != "csr"
        self.ctx.bind_current_stream()

    def __call__(self):
        st = self.fn(*self.args)
        if st:
            self.ctx.check(st, "dcpse_apply")
