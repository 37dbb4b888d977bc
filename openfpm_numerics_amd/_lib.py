"""ctypes binding of libdcpse_hip.so (the C ABI declared in include/dcpse_hip.h).

The product path has no CPU fallback: if the library is missing or no GPU is
visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdcpse_hip.so")

DCPSE_OK = 0
DCPSE_ERR_RANGE = 4
DCPSE_INFO = {0: "ok", 1: "fewer neighbours than monomials", 2: "moment matrix not positive definite",
              3: "eps is zero (coincident support)", 4: "non-finite kernel weight"}
DCPSE_ASM_REFINE = 1
DCPSE_MAX_ROW = 1024

_vp = C.c_void_p
_i32 = C.c_int32
_i64 = C.c_int64
_u32 = C.c_uint32
_f32 = C.c_float
_f64 = C.c_double

# name -> (restype, argtypes); mirrors include/dcpse_hip.h one to one
SIGNATURES = {
    "dcpse_version": (_i32, []),
    "dcpse_ctx_create": (_i32, [C.POINTER(_vp), _i32]),
    "dcpse_ctx_destroy": (_i32, [_vp]),
    "dcpse_ctx_set_stream": (_i32, [_vp, _vp]),
    "dcpse_ctx_get_stream": (_vp, [_vp]),
    "dcpse_last_error": (C.c_char_p, [_vp]),
    "dcpse_sync": (_i32, [_vp]),
    "dcpse_set_apply_config": (_i32, [_vp, _i32, _i32, _i32, _i32]),
    "dcpse_set_assembly_config": (_i32, [_vp, _i32, _i32]),
    "dcpse_set_assembly_launch": (_i32, [_vp, _i32, _i32, _i32, _i32]),
    "dcpse_basis_size": (_i32, [_i32, _vp, _u32]),
    "dcpse_basis": (_i32, [_i32, _vp, _u32, _vp, _i32]),
    "dcpse_rhs": (_i32, [_i32, _vp, _u32, _vp, _i32]),
    "dcpse_support_count": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _i32, _vp, _vp, _u32, _vp, _vp]),
    "dcpse_support_fill": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _i32, _vp, _vp, _u32, _vp, _vp]),
    "dcpse_support_count_subset": (_i32, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp, _i32, _vp, _vp, _u32, _vp,
                                          _vp]),
    "dcpse_support_fill_subset": (_i32, [_vp, _i32, _vp, _i64, _vp, _vp, _i64, _vp, _i32, _vp, _vp, _u32, _vp,
                                         _vp]),
    "dcpse_assemble": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _u32, _vp, _vp,
                              _i32, _vp, _vp, _u32, _f64, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_assemble_multi": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _u32, _vp, _vp, _i32, _vp, _vp, _vp,
                                    _u32, _f64, _u32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_apply_f32": (_i32, [_vp, _i64, _vp, _vp, _i32, _vp, _vp, _i64, _i32, _vp, _i64, _vp, _i64,
                               _f32, _f32]),
    "dcpse_apply_f64": (_i32, [_vp, _i64, _vp, _vp, _i32, _vp, _vp, _i64, _i32, _vp, _i64, _vp, _i64,
                               _f64, _f64]),
    "dcpse_transpose_count": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "dcpse_transpose_fill_f32": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_transpose_fill_f64": (_i32, [_vp, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_apply_transpose_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _vp, _i64,
                                         _f32, _f32]),
    "dcpse_apply_transpose_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _i64, _i32, _vp, _i64,
                                         _f64, _f64]),
    "dcpse_operator_diagnostics": (_i32, [_vp, _i32, _vp, _i64, _vp, _i64, _vp, _vp, _u32, _vp, _vp, _i32, _vp, _vp,
                                          _u32, _f64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_axpby_f32": (_i32, [_vp, _i64, _f32, _vp, _f32, _vp]),
    "dcpse_axpby_f64": (_i32, [_vp, _i64, _f64, _vp, _f64, _vp]),
    "dcpse_sell_count": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "dcpse_sell_count_noalloc": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "dcpse_sell_fill_cols": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_sell_fill_weights_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_apply_sell_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _f32, _f32]),
    "dcpse_sell_fill_weights_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_apply_sell_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _f64, _f64]),
    "dcpse_apply_transpose_sell_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f32, _f32]),
    "dcpse_apply_transpose_bundle16_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                                  _f32, _f32]),
    "dcpse_set_sell_pipeline": (_i32, [_vp, _i32, _i32]),
    "dcpse_set_apply_accumulation": (_i32, [_vp, _i32]),
    "dcpse_sell_fill_bundle_f32": (_i32, [_vp, _i64, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "dcpse_apply_sell_bundle_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _f32, _f32]),
    "dcpse_sell_bundle16_count": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "dcpse_sell_fill_bundle16_f32": (_i32, [_vp, _i64, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "dcpse_apply_sell_bundle16_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                             _f32, _f32]),
    "dcpse_sell_fill_bundle16_f64": (_i32, [_vp, _i64, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "dcpse_apply_sell_bundle16_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp,
                                             _f64, _f64]),
    "dcpse_export_csr_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _f64, _vp, _vp, _vp]),
    "dcpse_export_csr_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _f64, _vp, _vp, _vp]),
    "dcpse_scale_rows_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp]),
    "dcpse_stage_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _f32, _vp, _f32, _vp, _vp, _f32, _vp]),
    "dcpse_stage_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _f64, _vp, _f64, _vp, _vp, _f64, _vp]),
    "dcpse_stage_sell_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _f32, _vp, _f32, _vp, _vp, _f32,
                                    _vp]),
    "dcpse_rk4_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _f32, _f64, _i64, _vp, _vp]),
    "dcpse_rk4_f64": (_i32, [_vp, _i64, _vp, _vp, _vp, _f64, _f64, _i64, _vp, _vp]),
    "dcpse_rk4_sell_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _f32, _f64, _i64, _vp, _vp]),
    "dcpse_stage_bundle16_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _f32, _vp, _f32,
                                        _vp, _vp, _f32, _vp]),
    "dcpse_rk4_bundle16_f32": (_i32, [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _f32, _f64, _i64, _vp,
                                      _vp]),
}

_LIB = None
_LOCK = threading.Lock()


class DcpseError(RuntimeError):
    pass


def load():
    """Load libdcpse_hip.so; raises if it has not been built."""
    global _LIB
    with _LOCK:
        if _LIB is None:
            if not os.path.exists(LIB_PATH):
                raise DcpseError(f"{LIB_PATH} is missing: build it with "
                                 "`python -c 'import __graft_entry__ as g; g.build()'`")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _LIB = lib
    return _LIB


def ptr(t) -> int | None:
    """Device/host address of a torch tensor or numpy array (None for None)."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


class Context:
    """One library context bound to a device and, per call, to torch's stream."""

    def __init__(self, device: int = 0):
        import torch

        if not torch.cuda.is_available():
            raise DcpseError("no GPU visible: the DC-PSE HIP path has no CPU fallback")
        self.lib = load()
        self.device = device
        h = C.c_void_p()
        st = self.lib.dcpse_ctx_create(C.byref(h), device)
        if st != DCPSE_OK:
            raise DcpseError(f"dcpse_ctx_create(device={device}) failed with status {st}")
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and getattr(self, "lib", None) is not None:
            try:
                self.lib.dcpse_ctx_destroy(h)
            except Exception:
                pass
            self.h = None

    def bind_current_stream(self):
        import torch

        s = torch.cuda.current_stream(self.device)
        self.lib.dcpse_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream))

    def check(self, status: int, what: str):
        if status != DCPSE_OK:
            msg = self.lib.dcpse_last_error(self.h)
            raise DcpseError(f"{what} failed (status {status}): {msg.decode() if msg else ''}")

    def call(self, name: str, *args):
        self.bind_current_stream()
        self.check(getattr(self.lib, name)(self.h, *args), name)


_CTX: dict[int, Context] = {}


def context(device: int | None = None) -> Context:
    import torch

    if device is None:
        device = torch.cuda.current_device() if torch.cuda.is_available() else 0
    ctx = _CTX.get(device)
    if ctx is None:
        ctx = Context(device)
        _CTX[device] = ctx
    return ctx


def ptr_array(tensors) -> C.Array:
    arr = (C.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = ptr(t)
    return arr
