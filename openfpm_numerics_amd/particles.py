"""Particle sets and Verlet (support) lists on the device.

``ParticleSet`` stands in for openfpm's ``vector_dist`` as far as the DC-PSE
path needs it (positions, box, boundary conditions, named properties), and
``VerletList`` for the ``VL_NON_SYMMETRIC | VL_SKIP_REF_PART`` Verlet list the
reference builds with ``vector_dist::getVerlet`` (e.g.
src/DCPSE/tests/Dcpse_unit_tests.cpp:97): row p lists every q != p with
|x_p - x_q| < rCut, here in ascending particle index. Both live in HBM; the
neighbour search runs on the GPU (dcpse_support_count / dcpse_support_fill).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

PERIODIC = 1
NON_PERIODIC = 0


class ParticleSet:
    """Positions (structure of arrays, float64) plus named properties."""

    def __init__(self, pos, box_lo, box_hi, bc=None, device=None):
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        p = torch.as_tensor(pos, dtype=torch.float64)
        if p.dim() != 2 or p.shape[1] not in (2, 3):
            raise ValueError("pos must be (N, 2) or (N, 3)")
        self.dims = int(p.shape[1])
        self.pos_soa = p.t().contiguous().to(dev)
        self.box_lo = np.asarray(box_lo, dtype=np.float64).reshape(self.dims).copy()
        self.box_hi = np.asarray(box_hi, dtype=np.float64).reshape(self.dims).copy()
        bc = [NON_PERIODIC] * self.dims if bc is None else list(bc)
        self.periodic_mask = sum(1 << d for d in range(self.dims) if bc[d] == PERIODIC)
        self.props: dict[str, torch.Tensor] = {}
        self.map_ctr = 0  # bumped whenever positions change (SE_CLASS1-style staleness check)

    @property
    def device(self):
        return self.pos_soa.device

    def size_local(self) -> int:
        return int(self.pos_soa.shape[1])

    def getPos(self, p: int) -> np.ndarray:
        return self.pos_soa[:, p].cpu().numpy()

    def positions(self) -> torch.Tensor:
        """(N, dim) view of the positions."""
        return self.pos_soa.t()

    def set_positions(self, pos):
        p = torch.as_tensor(pos, dtype=torch.float64, device=self.device)
        self.pos_soa.copy_(p.t())
        self.map_ctr += 1

    def getProp(self, name: str) -> torch.Tensor:
        return self.props[name]

    def setProp(self, name: str, value) -> torch.Tensor:
        t = torch.as_tensor(value, device=self.device)
        self.props[name] = t
        return t

    def reorder(self, order: str = "morton") -> np.ndarray:
        """Sorts the particles, the counterpart of openfpm's vector_dist::reorder:
        "morton" along the Z-order curve over the box, "cell" row-major over a
        grid of about N^(1/dim) cells per dimension (neighbours then lie within a
        few cell rows of each other, which keeps column offsets small). Positions
        and properties are permuted in place; returns perm with new[i] =
        old[perm[i]]. Verlet lists and operators built before must be rebuilt
        (map counter)."""
        pos = self.pos_soa.t().cpu().numpy()
        if order == "morton":
            perm = morton_order(pos, self.box_lo, self.box_hi)
        elif order == "cell" or order.startswith("cell:"):
            perm = cell_order(pos, self.box_lo, self.box_hi, int(order[5:]) if order.startswith("cell:") else None)
        else:
            raise ValueError("order must be 'morton', 'cell' or 'cell:<cells per dimension>'")
        idx = torch.as_tensor(perm, device=self.device)
        self.pos_soa = self.pos_soa[:, idx].contiguous()
        for k, v in self.props.items():
            self.props[k] = v[..., idx].contiguous() if v.shape[-1] == perm.size else v
        self.map_ctr += 1
        return perm

    def getVerlet(self, rCut) -> "VerletList":
        return VerletList.build(self, rCut)

    def updateVerlet(self, vl: "VerletList", rCut) -> "VerletList":
        fresh = VerletList.build(self, rCut, domain=vl.domain)
        vl.__dict__.update(fresh.__dict__)
        vl._sell = None
        vl._slices = None
        vl._n_exc16 = None
        return vl


class ParticleSubset:
    """The particles ``ids`` of a ParticleSet, the counterpart of openfpm's
    ``vector_dist_subset``: the domain of two-set operators whose support is
    the whole set (src/DCPSE/DCPSE_op/tests/DCPSE_op_subset_test.cpp:113-141).
    Its Verlet list skips each particle's own index in the parent
    (VL_SKIP_REF_PART); fields over the subset are the parent's at ``ids``."""

    def __init__(self, parent: ParticleSet, ids):
        self.parent = parent
        ids_t = torch.as_tensor(np.asarray(ids) if not torch.is_tensor(ids) else ids, device=parent.device)
        if ids_t.dim() != 1:
            raise ValueError("ids must be one-dimensional")
        if ids_t.numel() and (int(ids_t.min()) < 0 or int(ids_t.max()) >= parent.size_local()):
            raise ValueError("subset ids out of range")
        self.ids = ids_t.to(torch.int32).contiguous()
        self._pos = None
        self._pos_ctr = None

    dims = property(lambda self: self.parent.dims)
    box_lo = property(lambda self: self.parent.box_lo)
    box_hi = property(lambda self: self.parent.box_hi)
    periodic_mask = property(lambda self: self.parent.periodic_mask)
    device = property(lambda self: self.parent.device)
    map_ctr = property(lambda self: self.parent.map_ctr)

    @property
    def pos_soa(self) -> torch.Tensor:
        """Positions of the subset (gathered again after the parent moved)."""
        if self._pos is None or self._pos_ctr != self.parent.map_ctr:
            self._pos = self.parent.pos_soa[:, self.ids.long()].contiguous()
            self._pos_ctr = self.parent.map_ctr
        return self._pos

    def getIds(self) -> torch.Tensor:
        return self.ids

    def size_local(self) -> int:
        return int(self.ids.numel())

    def getPos(self, p: int) -> np.ndarray:
        return self.pos_soa[:, p].cpu().numpy()

    def gather(self, f: torch.Tensor) -> torch.Tensor:
        """A field over the parent restricted to the subset (last axis)."""
        return f[..., self.ids.long()]

    def getVerlet(self, rCut) -> "VerletList":
        return VerletList.build(self.parent, rCut, domain=self)


def cell_order(pos: np.ndarray, box_lo, box_hi, cells: int | None = None) -> np.ndarray:
    """Permutation sorting (N, dim) points row-major (last dimension fastest)
    by their cell in a grid of `cells` per dimension (default about N^(1/dim))."""
    pos = np.asarray(pos, dtype=np.float64)
    n, dim = pos.shape
    cells = cells or max(1, int(round(n ** (1.0 / dim))))
    lo = np.asarray(box_lo, dtype=np.float64)
    hi = np.asarray(box_hi, dtype=np.float64)
    q = np.clip(np.floor((pos - lo) / (hi - lo) * cells), 0, cells - 1).astype(np.int64)
    key = np.zeros(n, np.int64)
    for d in range(dim):
        key = key * cells + q[:, d]
    return np.argsort(key, kind="stable")


def morton_order(pos: np.ndarray, box_lo, box_hi, bits: int | None = None) -> np.ndarray:
    """Permutation sorting (N, dim) points by their Morton code over the box."""
    pos = np.asarray(pos, dtype=np.float64)
    n, dim = pos.shape
    bits = bits or (21 if dim == 3 else 31)
    lo = np.asarray(box_lo, dtype=np.float64)
    hi = np.asarray(box_hi, dtype=np.float64)
    cells = (1 << bits) - 1
    q = np.clip(np.floor((pos - lo) / (hi - lo) * cells), 0, cells).astype(np.uint64)
    code = np.zeros(n, np.uint64)
    for b in range(bits):
        for d in range(dim):
            code |= ((q[:, d] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * dim + d)
    return np.argsort(code, kind="stable")


class VerletList:
    """Neighbour CSR in HBM: row_ptr int64[n+1], col int32[nnz]."""

    def __init__(self, row_ptr: torch.Tensor, col: torch.Tensor, n_cols: int, rcut, domain=None):
        self.row_ptr = row_ptr
        self.col = col
        self.n_rows = int(row_ptr.numel() - 1)
        self.n_cols = n_cols
        self.rcut = rcut
        self.domain = domain
        self._host = None
        self._sell = None
        self._slices = None
        self._n_exc16 = None
        self.map_ctr = -1

    @property
    def nnz(self) -> int:
        return int(self.col.numel())

    @classmethod
    def build(cls, support: ParticleSet, rCut, domain: "ParticleSet | ParticleSubset | None" = None) -> "VerletList":
        ctx = _lib.context(support.device.index or 0)
        dev = support.device
        dom = support if domain is None else domain
        if dom.dims != support.dims:
            raise ValueError("support and domain sets differ in dimension")
        rc = torch.as_tensor(rCut, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        per = int(rc.numel() > 1)
        if per and rc.numel() != dom.size_local():
            raise ValueError("per-particle rCut needs one value per domain particle")
        n_dom = dom.size_local()
        if n_dom == 0:  # no rows (an empty set has no position storage to pass on)
            vl = cls(torch.zeros(1, dtype=torch.int64, device=dev), torch.empty(0, dtype=torch.int32, device=dev),
                     support.size_local(), rc, domain)
            vl.map_ctr = support.map_ctr
            return vl
        row_ptr = torch.empty(n_dom + 1, dtype=torch.int64, device=dev)
        nnz = np.zeros(1, np.int64)
        pdom = None if domain is None else _lib.ptr(dom.pos_soa)
        lo = np.ascontiguousarray(support.box_lo)
        hi = np.ascontiguousarray(support.box_hi)
        if isinstance(domain, ParticleSubset):
            if domain.parent is not support:
                raise ValueError("a subset domain must be a subset of the support set")
            count, fill, dom_args = "dcpse_support_count_subset", "dcpse_support_fill_subset", (
                pdom, _lib.ptr(domain.ids), n_dom)
        else:
            count, fill, dom_args = "dcpse_support_count", "dcpse_support_fill", (pdom, n_dom)
        ctx.call(count, support.dims, _lib.ptr(support.pos_soa), support.size_local(),
                 *dom_args, _lib.ptr(rc), per, lo.ctypes.data, hi.ctypes.data,
                 support.periodic_mask, _lib.ptr(row_ptr), nnz.ctypes.data)
        col = torch.empty(max(int(nnz[0]), 1), dtype=torch.int32, device=dev)
        if int(nnz[0]) > 0:
            ctx.call(fill, support.dims, _lib.ptr(support.pos_soa), support.size_local(),
                     *dom_args, _lib.ptr(rc), per, lo.ctypes.data, hi.ctypes.data,
                     support.periodic_mask, _lib.ptr(row_ptr), _lib.ptr(col))
        col = col[: int(nnz[0])]
        vl = cls(row_ptr, col, support.size_local(), rc, domain)
        vl.map_ctr = support.map_ctr
        return vl

    # host mirror for per-key accessors (getNNPart / get), built on first use
    def _mirror(self):
        if self._host is None:
            self._host = (self.row_ptr.cpu().numpy(), self.col.cpu().numpy())
        return self._host

    def getNNPart(self, p: int) -> int:
        rp, _ = self._mirror()
        return int(rp[p + 1] - rp[p])

    def get(self, p: int, j: int) -> int:
        rp, col = self._mirror()
        return int(col[rp[p] + j])

    def row_lengths(self) -> torch.Tensor:
        return self.row_ptr[1:] - self.row_ptr[:-1]

    def sell_slices(self):
        """(slice_ptr, sell_row, n_slots): the slice structure of the
        sliced-ELL layout (what the 16-bit bundles need); built on first use."""
        if getattr(self, "_slices", None) is None:
            self._slices = sell_slices(self.row_ptr, self.n_rows)
        return self._slices

    def sell(self):
        """(slice_ptr, sell_row, sell_col): the sliced-ELL execution layout of
        this support with 32-bit columns, shared by every FP32 operator
        assembled on it; built on first use."""
        if getattr(self, "_sell", None) is None:
            slice_ptr, sell_row, n_slots = self.sell_slices()
            self._sell = (slice_ptr, sell_row, sell_cols(self.row_ptr, self.col, self.n_rows, slice_ptr, sell_row,
                                                         n_slots))
        return self._sell


def sell_slices(row_ptr: torch.Tensor, n_rows: int):
    """Slice structure of the sliced-ELL layout of a CSR: slices of 64 rows,
    rows ordered by decreasing length inside a slice (sell_row maps lane ->
    row, -1 = none); (slice_ptr, sell_row, n_slots)."""
    ctx = _lib.context(row_ptr.device.index or 0)
    dev = row_ptr.device
    n_slices = (n_rows + 63) // 64
    slice_ptr = torch.empty(n_slices + 1, dtype=torch.int64, device=dev)
    sell_row = torch.empty(max(64 * n_slices, 64), dtype=torch.int32, device=dev)
    n_slots = np.zeros(1, np.int64)
    ctx.call("dcpse_sell_count", n_rows, _lib.ptr(row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
             n_slots.ctypes.data)
    return slice_ptr, sell_row, int(n_slots[0])


def sell_cols(row_ptr: torch.Tensor, col: torch.Tensor, n_rows: int, slice_ptr, sell_row, n_slots: int):
    """The 32-bit columns of the sliced-ELL layout (padding slots included)."""
    sell_col = torch.empty(max(n_slots, 4), dtype=torch.int32, device=row_ptr.device)
    if n_rows:
        _lib.context(row_ptr.device.index or 0).call(
            "dcpse_sell_fill_cols", n_rows, _lib.ptr(row_ptr), _lib.ptr(col), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
            _lib.ptr(sell_col))
    return sell_col


def sell_layout(row_ptr: torch.Tensor, col: torch.Tensor, n_rows: int):
    """Sliced-ELL layout of a CSR: (slice_ptr, sell_row, sell_col), see
    sell_slices and sell_cols."""
    slice_ptr, sell_row, n_slots = sell_slices(row_ptr, n_rows)
    return slice_ptr, sell_row, sell_cols(row_ptr, col, n_rows, slice_ptr, sell_row, n_slots)


def sell_weights(row_ptr: torch.Tensor, W: torch.Tensor, layout, n_rows: int):
    """FP32 weights W (CSR order) in the sliced-ELL layout (slice_ptr, sell_row, sell_col)."""
    slice_ptr, sell_row, sell_col = layout
    ctx = _lib.context(row_ptr.device.index or 0)
    out = torch.empty(sell_col.numel(), dtype=torch.float32, device=row_ptr.device)
    if n_rows:
        ctx.call("dcpse_sell_fill_weights_f32", n_rows, _lib.ptr(row_ptr), _lib.ptr(W), _lib.ptr(slice_ptr),
                 _lib.ptr(sell_row), _lib.ptr(out))
    return out
