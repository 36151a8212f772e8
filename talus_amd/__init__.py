"""talus_amd -- MI355X-native deformable tet step behind the talus API.

The product is libtalus_mi.so (gfx950 HIP kernels, host C++ orchestration and
the C ABI in include/talus_mi.h).  ``talus_amd.api`` binds that C ABI with
ctypes (PhysicalMaterial, the tet mesh builder, Context for the tet bodies'
deformable step, the strainstep hex lattices), ``talus_amd.adjoint`` wraps
the differentiable rollout in a torch.autograd.Function, and
``talus_amd.scenes`` generates the seeded benchmark workloads.  The
talus::World host binding is C++ (include/talus_mi.hpp).
"""
from . import api  # noqa: F401

__all__ = ["api"]
