"""talus_amd -- MI355X-native deformable tet step behind the talus API.

The product is libtalus_mi.so (gfx950 HIP kernels, host C++ orchestration and
the C ABI in include/talus_mi.h).  ``talus_amd.api`` binds that C ABI with
ctypes, ``talus_amd.world`` mirrors talus::World / PhysicalMaterial on top of
it, and ``talus_amd.scenes`` generates the seeded benchmark workloads.
"""
from . import api  # noqa: F401

__all__ = ["api"]
