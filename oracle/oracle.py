"""numpy/ctypes front end of the CPU DC-PSE oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker or the timed CPU
baseline, never as the product path. The C++ source (dcpse_oracle.cpp) cites
the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return path


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_basis_size.argtypes = [C.c_int, _u32p, C.c_uint]
        L.orc_basis_size.restype = C.c_int
        L.orc_basis.argtypes = [C.c_int, _u32p, C.c_uint, _u32p, C.c_int]
        L.orc_basis.restype = C.c_int
        L.orc_monomial_eval.argtypes = [C.c_int, _u32p, C.c_uint, _f64p]
        L.orc_monomial_eval.restype = C.c_double
        L.orc_monomial_derivative.argtypes = [C.c_int, _u32p, C.c_uint, _u32p, _u32p]
        L.orc_monomial_derivative.restype = C.c_uint
        L.orc_rhs.argtypes = [C.c_int, _u32p, C.c_uint, _f64p]
        L.orc_rhs.restype = C.c_int
        L.orc_vandermonde_row.argtypes = [C.c_int, C.c_int, _u32p, _f64p, _f64p]
        L.orc_vandermonde_row.restype = None
        L.orc_support.argtypes = [C.c_int, _f64p, C.c_int64, _f64p, C.c_int64, C.c_int, C.c_void_p,
                                  _f64p, C.c_int, _f64p, _f64p, C.c_uint,
                                  _i64p, C.c_void_p, C.c_int]
        L.orc_support.restype = C.c_int64
        L.orc_assemble.argtypes = [C.c_int, _f64p, C.c_int64, _f64p, C.c_int64,
                                   _f64p, _f64p, C.c_uint, _i64p, _i32p,
                                   _u32p, C.c_uint, C.c_double, C.c_int,
                                   C.c_int64, C.c_int64,
                                   _f64p, _f64p, _f64p, _i32p, C.c_int]
        L.orc_assemble.restype = C.c_int
        L.orc_apply.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_uint,
                                _f64p, _f64p, _f64p, C.c_int]
        L.orc_apply.restype = None
        L.orc_value_nz.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_uint,
                                   _i64p, _i64p, _f64p]
        L.orc_value_nz.restype = None
        L.orc_max_threads.argtypes = []
        L.orc_max_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def basis(beta, order):
    """Monomial exponents (m, dim) in the reference's enumeration order."""
    beta = _u32(beta)
    dim = beta.size
    m = lib().orc_basis_size(dim, beta, order)
    out = np.zeros(m * dim, np.uint32)
    lib().orc_basis(dim, beta, order, out, m)
    return out.reshape(m, dim)


def monomial_eval(exps, x, scalar=1):
    exps = _u32(exps)
    return lib().orc_monomial_eval(exps.size, exps, scalar,
                                   np.ascontiguousarray(x, dtype=np.float64))


def monomial_derivative(exps, beta, scalar=1):
    exps = _u32(exps)
    out = np.zeros_like(exps)
    s = lib().orc_monomial_derivative(exps.size, exps, scalar, _u32(beta), out)
    return out, s


def rhs(beta, order):
    beta = _u32(beta)
    m = lib().orc_basis_size(beta.size, beta, order)
    b = np.zeros(m)
    lib().orc_rhs(beta.size, beta, order, b)
    return b


def vandermonde_row(exps, x):
    exps = np.ascontiguousarray(np.asarray(exps, dtype=np.uint32))
    m, dim = exps.shape
    row = np.zeros(m)
    lib().orc_vandermonde_row(dim, m, exps.ravel(), np.ascontiguousarray(x, dtype=np.float64), row)
    return row


def _soa(pos):
    """(N, dim) positions -> flat SoA float64 (dim*N)."""
    pos = np.asarray(pos, dtype=np.float64)
    return np.ascontiguousarray(pos.T).ravel()


def support(pos, rcut, box_lo, box_hi, periodic=0, pos_dom=None, nthreads=0, dom_ids=None):
    """Neighbour CSR (row_ptr int64[Nd+1], col int32[nnz]); rows ascending.

    ``rcut`` is a scalar or a per-domain-particle array. ``pos_dom``: a
    separate domain set (nothing skipped); ``dom_ids``: a domain that is the
    subset ``pos[dom_ids]`` of the support set, each row skipping its own
    particle.
    """
    pos = np.asarray(pos, dtype=np.float64)
    self_ids = None
    if dom_ids is not None:
        if pos_dom is not None:
            raise ValueError("give pos_dom or dom_ids, not both")
        self_ids = np.ascontiguousarray(dom_ids, dtype=np.int32)
        pos_dom = pos[self_ids]
    same = pos_dom is None
    pdom = pos if same else np.asarray(pos_dom, dtype=np.float64)
    dim = pos.shape[1]
    rc = np.atleast_1d(np.asarray(rcut, dtype=np.float64)).copy()
    per = int(rc.size > 1)
    sup, dom = _soa(pos), _soa(pdom)
    lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    hi = np.ascontiguousarray(box_hi, dtype=np.float64)
    nd = pdom.shape[0]
    row_ptr = np.zeros(nd + 1, np.int64)
    sp = None if self_ids is None else self_ids.ctypes.data_as(C.c_void_p)
    nnz = lib().orc_support(dim, sup, pos.shape[0], dom, nd, int(same), sp, rc, per, lo, hi,
                            int(periodic), row_ptr, None, nthreads)
    if nnz < 0:
        raise ValueError("periodic dimension has fewer than 3 cells; rCut too large for the box")
    col = np.zeros(max(nnz, 1), np.int32)
    lib().orc_support(dim, sup, pos.shape[0], dom, nd, int(same), sp, rc, per, lo, hi,
                      int(periodic), row_ptr, col.ctypes.data_as(C.c_void_p), nthreads)
    return row_ptr, col[:nnz]


class Assembly:
    """Result of one oracle assembly (one differential signature)."""

    def __init__(self, eps, eps_inv_pow, K, info, beta, order):
        self.eps, self.eps_inv_pow, self.K, self.info = eps, eps_inv_pow, K, info
        self.beta = tuple(int(b) for b in beta)
        self.order = order
        self.diff_order = int(sum(self.beta))


def assemble(pos, row_ptr, col, beta, order, box_lo, box_hi, periodic=0, h_over_eps=0.9,
             dense_ev=True, pos_dom=None, p_range=None, nthreads=0):
    pos = np.asarray(pos, dtype=np.float64)
    pdom = pos if pos_dom is None else np.asarray(pos_dom, dtype=np.float64)
    dim = pos.shape[1]
    nd = pdom.shape[0]
    nnz = int(row_ptr[-1])
    eps = np.zeros(nd)
    eip = np.zeros(nd)
    K = np.zeros(max(nnz, 1))
    info = np.zeros(nd, np.int32)
    pb, pe = (0, nd) if p_range is None else p_range
    lib().orc_assemble(dim, _soa(pos), pos.shape[0], _soa(pdom), nd,
                       np.ascontiguousarray(box_lo, dtype=np.float64),
                       np.ascontiguousarray(box_hi, dtype=np.float64), int(periodic),
                       np.ascontiguousarray(row_ptr, dtype=np.int64),
                       np.ascontiguousarray(col, dtype=np.int32) if nnz else np.zeros(1, np.int32),
                       _u32(beta), order, h_over_eps, int(dense_ev), pb, pe,
                       eps, eip, K, info, nthreads)
    return Assembly(eps, eip, K[:nnz], info, beta, order)


def apply(asm: Assembly, row_ptr, col, f, f_dom=None, nthreads=0):
    """Reference-form apply (Dcpse.hpp:467-499) in FP64."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    fd = f if f_dom is None else np.ascontiguousarray(f_dom, dtype=np.float64)
    out = np.zeros(fd.size)
    nnz = int(row_ptr[-1])
    lib().orc_apply(fd.size, np.ascontiguousarray(row_ptr, dtype=np.int64),
                    np.ascontiguousarray(col, dtype=np.int32) if nnz else np.zeros(1, np.int32),
                    np.ascontiguousarray(asm.K) if nnz else np.zeros(1),
                    asm.eps_inv_pow, asm.diff_order, f, fd, out, nthreads)
    return out


def value_nz(asm: Assembly, row_ptr, col, n_cols=None):
    """Assembled sparse operator (value_nz semantics) as a scipy CSR matrix."""
    import scipy.sparse as sp
    nd = asm.eps.size
    nnz = int(row_ptr[-1])
    rows = np.zeros(2 * nnz, np.int64)
    cols = np.zeros(2 * nnz, np.int64)
    vals = np.zeros(2 * nnz)
    lib().orc_value_nz(nd, np.ascontiguousarray(row_ptr, dtype=np.int64),
                       np.ascontiguousarray(col, dtype=np.int32), np.ascontiguousarray(asm.K),
                       asm.eps_inv_pow, asm.diff_order, rows, cols, vals)
    nc = nd if n_cols is None else n_cols
    return sp.csr_matrix((vals, (rows, cols)), shape=(nd, nc))


def max_threads():
    return lib().orc_max_threads()
