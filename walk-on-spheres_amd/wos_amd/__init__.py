"""wos_amd -- MI355X walk-on-spheres engine (host side).

The solver itself is libwos_hip.so (HIP kernels for gfx950 behind the C ABI
of include/wos_hip.h); this package holds its ctypes binding, the equation
descriptors, the domain generators and the torch autograd wrapper.
"""
from . import capi, meshgen, problems  # noqa: F401
from .engine import Context, screened_epsilon  # noqa: F401

__version__ = "0.1.0"
