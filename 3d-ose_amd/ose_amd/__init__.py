"""ose_amd -- MI355X-native one-step-extrapolation (OSE) acoustic time stepper.

The GPU path is libose_hip.so (hand-written gfx950 kernels behind the C-ABI
of include/ose_api.h); the host-side logic (wavelets, Bessel coefficients,
tapers, models, file formats) is libose_host.so (include/ose_host.h).  This
package is a thin ctypes view of both for tests and the benchmark.
"""
from ._lib import OseError  # noqa: F401
from .backend import BoundaryNet, Context, boundary_net_size, extent_is_supported, extent_kind, supported_extent  # noqa: F401
