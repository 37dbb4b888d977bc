"""MI355X-native DC-PSE operators (openfpm_numerics' Dcpse / Derivative_* / Laplacian path).

Assembly and application run as hand-written gfx950 kernels in
libdcpse_hip.so, reached through the C ABI of include/dcpse_hip.h. PyTorch
provides device memory and streams only.
"""
from ._lib import DcpseError, load as load_library
from .autograd import dcpse_apply
from .dcpse import (ADAPTIVE, AT_LEAST_N_PARTICLES, LOAD, RADIUS, Dcpse, apply_ops, basis, rhs,
                    set_apply_accumulation)
from .operators import (Derivative_G, Derivative_x, Derivative_xx, Derivative_xy, Derivative_xz, Derivative_y,
                        Derivative_yy, Derivative_yz, Derivative_z, Derivative_zz, Gradient, Laplacian, fused)
from .ode import rk4
from .particles import NON_PERIODIC, PERIODIC, ParticleSet, ParticleSubset, VerletList

__all__ = [
    "ADAPTIVE", "AT_LEAST_N_PARTICLES", "LOAD", "RADIUS", "NON_PERIODIC", "PERIODIC",
    "Dcpse", "DcpseError", "ParticleSet", "ParticleSubset", "VerletList", "apply_ops", "basis", "rhs", "dcpse_apply", "fused",
    "set_apply_accumulation",
    "Derivative_G", "Derivative_x", "Derivative_xx", "Derivative_xy", "Derivative_xz", "Derivative_y",
    "Derivative_yy", "Derivative_yz", "Derivative_z", "Derivative_zz", "Gradient", "Laplacian",
    "load_library", "rk4",
]
