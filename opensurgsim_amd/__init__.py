"""MI355X-native Fem3D implicit-step path (OpenSurgSim Fem3DRepresentation drop-in).

The compute path is libfem_gpu.so (hand-written gfx950 HIP kernels behind the C ABI declared in
include/fem_gpu.h); this package holds its ctypes binding, synthetic mesh generators and the
Python mirror of the reference's representation interface.
"""
from . import meshgen  # noqa: F401

__all__ = ["meshgen", "femgpu"]
