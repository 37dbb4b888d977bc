"""Explicit time integration on the device for DC-PSE right-hand sides.

Classic RK4 of df/dt = alpha * D f, the integration the reference runs with
boost::odeint's runge_kutta4 over its vector_dist state
(src/OdeIntegrators/OdeIntegrators.hpp:63-135,
src/OdeIntegrators/tests/OdeIntegratores_base_tests.cpp:478-577). Each stage
is one launch of the library (dcpse_stage_* / dcpse_rk4_*): the operator
gather with the stage's axpys fused into its epilogue, so the state stays in
HBM for the whole integration and a step costs four operator passes.
"""
from __future__ import annotations

import torch

from . import _lib
from .dcpse import Dcpse


def _dcpse(op) -> Dcpse:
    d = op.dcpse if hasattr(op, "dcpse") else op
    if not d.same_set:
        raise ValueError("time integration needs an operator whose support and domain sets are the same")
    return d


def rk4(op, f: torch.Tensor, dt: float, n_steps: int, alpha: float = 1.0, layout: str = "auto",
        work: torch.Tensor | None = None) -> torch.Tensor:
    """Advances f (N,) in place by n_steps RK4 steps of df/dt = alpha * D f and returns it.

    float32 state runs the FP32 kernels: layout "bundle16" (sliced-ELL with
    16-bit column offsets, the fewest bytes), "sell" (32-bit columns), "csr",
    or "auto": the 4 n_steps stage applies count towards the operator's
    ``dcpse.SELL_AFTER`` and run on "bundle16" once it is reached (its copy is
    then built; "sell" where too many neighbours fall outside 16-bit offsets,
    ``dcpse.BUNDLE16_MAX_EXC``), on "csr" before. float64 state runs the FP64 CSR kernels.
    ``work`` (3N of f's dtype) may be passed to avoid an allocation per call.
    """
    d = _dcpse(op)
    vl = d.verletList
    n = vl.n_rows
    if f.dim() != 1 or f.numel() != n or not f.is_contiguous():
        raise ValueError("f must be a contiguous (N,) tensor on the operator's particles")
    if f.dtype not in (torch.float32, torch.float64):
        raise TypeError("state must be float32 or float64")
    if work is None:
        work = torch.empty(3 * n, dtype=f.dtype, device=f.device)
    elif work.numel() < 3 * n or work.dtype != f.dtype or not work.is_contiguous():
        raise ValueError("work must be a contiguous tensor of 3N elements of f's dtype")
    ctx = _lib.context(f.device.index or 0)
    ctx.bind_current_stream()
    if f.dtype == torch.float32 and layout == "auto":
        d._n32["fwd"] += 4 * int(n_steps) - 1
        layout = ("bundle16" if d.bundle16_suits() else "sell") if d._count_apply32("fwd") else "csr"
    if f.dtype == torch.float64:
        ctx.call("dcpse_rk4_f64", n, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(d.weights(torch.float64)),
                 float(alpha), float(dt), int(n_steps), _lib.ptr(f), _lib.ptr(work))
    elif layout == "bundle16":
        slice_ptr = vl.sell_slices()[0]
        bundle, row16, exc_ptr, exc_col, exc_W, ne = d.sell_bundle16()
        ctx.call("dcpse_rk4_bundle16_f32", n, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(row16),
                 _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne, float(alpha),
                 float(dt), int(n_steps), _lib.ptr(f), _lib.ptr(work))
    elif layout == "sell":
        slice_ptr, sell_row, sell_col = vl.sell()
        ctx.call("dcpse_rk4_sell_f32", n, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                 _lib.ptr(sell_col), _lib.ptr(d.sell_weights()), float(alpha), float(dt), int(n_steps), _lib.ptr(f),
                 _lib.ptr(work))
    elif layout == "csr":
        ctx.call("dcpse_rk4_f32", n, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col), _lib.ptr(d.weights(torch.float32)),
                 float(alpha), float(dt), int(n_steps), _lib.ptr(f), _lib.ptr(work))
    else:
        raise ValueError(f"unknown layout {layout!r}")
    return f


def stage(op, u: torch.Tensor, alpha: float, y_a: torch.Tensor | None, ca: float, out_a: torch.Tensor | None,
          y_b: torch.Tensor | None, cb: float, out_b: torch.Tensor | None, layout: str = "auto"):
    """One fused stage: k = alpha * D u; out_a = y_a + ca k; out_b = y_b + cb k (outputs may be None).

    Layouts as for rk4; "auto" counts one apply towards ``dcpse.SELL_AFTER``."""
    d = _dcpse(op)
    vl = d.verletList
    ctx = _lib.context(u.device.index or 0)
    ctx.bind_current_stream()
    p = _lib.ptr
    if u.dtype == torch.float32 and layout == "auto":
        layout = ("bundle16" if d.bundle16_suits() else "sell") if d._count_apply32("fwd") else "csr"
    if u.dtype == torch.float64:
        ctx.call("dcpse_stage_f64", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col),
                 _lib.ptr(d.weights(torch.float64)), _lib.ptr(u), float(alpha), p(y_a), float(ca), p(out_a), p(y_b),
                 float(cb), p(out_b))
    elif layout == "bundle16":
        slice_ptr = vl.sell_slices()[0]
        bundle, row16, exc_ptr, exc_col, exc_W, ne = d.sell_bundle16()
        ctx.call("dcpse_stage_bundle16_f32", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(row16),
                 _lib.ptr(bundle), _lib.ptr(exc_ptr), _lib.ptr(exc_col), _lib.ptr(exc_W), ne, _lib.ptr(u),
                 float(alpha), p(y_a), float(ca), p(out_a), p(y_b), float(cb), p(out_b))
    elif layout == "sell":
        slice_ptr, sell_row, sell_col = vl.sell()
        ctx.call("dcpse_stage_sell_f32", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(slice_ptr), _lib.ptr(sell_row),
                 _lib.ptr(sell_col), _lib.ptr(d.sell_weights()), _lib.ptr(u), float(alpha), p(y_a), float(ca),
                 p(out_a), p(y_b), float(cb), p(out_b))
    elif layout == "csr":
        ctx.call("dcpse_stage_f32", vl.n_rows, _lib.ptr(vl.row_ptr), _lib.ptr(vl.col),
                 _lib.ptr(d.weights(torch.float32)), _lib.ptr(u), float(alpha), p(y_a), float(ca), p(out_a), p(y_b),
                 float(cb), p(out_b))
    else:
        raise ValueError(f"unknown layout {layout!r}")
