"""ctypes binding of libtalus_mi.so (the C ABI declared in include/talus_mi.h).

Mirrors the host surface of the reference's deformable path: a
``PhysicalMaterial`` record (src/talus/material.h:12-29), a tet
``PreparedMesh`` built by ``build_lattice`` (deformable::buildMesh,
src/talus/deformable/solver/solver.cc:2185), bodies with a
``DeformableState`` (mesh.h:56-62) and ``Context.step`` for
``deformable::step`` (solver.h:451-462).  Errors raise ``TalusError`` with the
library's message; nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtalus_mi.so")

FP64 = 0
MIXED = 1
BLOCK_JACOBI = 0
FACTOR = 1
DETECT_TEARS = 1
RECORD_EVENTS = 2
CONVERGE = 4
DIRECT_LAUNCH = 8
KERNEL_TIMING = 16
FIXED_FRAME = 32
EVENT_YIELD = 0
EVENT_TEAR = 1
COROTATED = 0
NEO_HOOKEAN = 1


class TalusError(RuntimeError):
    pass


class PhysicalMaterial(C.Structure):
    """talus::PhysicalMaterial, same field order and defaults."""

    _fields_ = [
        ("friction", C.c_float), ("restitution", C.c_float),
        ("density", C.c_float), ("strength", C.c_float),
        ("youngs_modulus", C.c_float), ("poissons_ratio", C.c_float),
        ("stiffness_damping", C.c_float), ("yield_stress", C.c_float),
        ("hardening_modulus", C.c_float), ("elastic_model", C.c_uint32),
        ("hysteresis", C.c_float), ("damping_ratio", C.c_float),
        ("tear_strain", C.c_float), ("plastic_max", C.c_float),
        ("crush_stress", C.c_float), ("crush_limit", C.c_float),
    ]

    def __init__(self, **kw):
        super().__init__()
        defaults = dict(friction=0.5, restitution=0.0, density=1000.0, strength=math.inf,
                        youngs_modulus=0.0, poissons_ratio=0.3, stiffness_damping=0.0,
                        yield_stress=math.inf, hardening_modulus=0.0, elastic_model=COROTATED,
                        hysteresis=0.0, damping_ratio=math.inf, tear_strain=math.inf,
                        plastic_max=math.inf, crush_stress=math.inf, crush_limit=1.0)
        defaults.update(kw)
        for k, v in defaults.items():
            setattr(self, k, v)


class _Mesh(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint32), ("tets", C.c_uint32),
        ("rest_positions", C.POINTER(C.c_float)), ("node_mass", C.POINTER(C.c_float)),
        ("anchored", C.POINTER(C.c_uint8)), ("tet_nodes", C.POINTER(C.c_uint32)),
        ("rest_inverse", C.POINTER(C.c_float)), ("volume", C.POINTER(C.c_float)),
        ("tet_material", C.POINTER(C.c_uint32)),
    ]


class _Lattice(C.Structure):
    _fields_ = [
        ("resolution", C.c_uint32 * 3), ("voxel_size", C.c_float * 3),
        ("origin", C.c_float * 3), ("solid", C.POINTER(C.c_uint8)),
        ("cell_lane", C.POINTER(C.c_uint32)), ("sample_positions", C.POINTER(C.c_float)),
    ]


class Support(C.Structure):
    """SupportVolume: anchored iff |world_to_support * p| <= half_extents."""

    _fields_ = [("world_to_support", C.c_float * 12), ("half_extents", C.c_float * 3)]

    @staticmethod
    def box(center, half):
        s = Support()
        s.world_to_support[:] = [1, 0, 0, -center[0], 0, 1, 0, -center[1], 0, 0, 1, -center[2]]
        s.half_extents[:] = list(half)
        return s


class Energy(C.Structure):
    """talus_mi_energy (ledger.h Energy)."""

    _fields_ = [("kinetic", C.c_double), ("gravity", C.c_double), ("elastic", C.c_double)]


class RestBody(C.Structure):
    """talus_mi_rest_body (screen.h RestBody)."""

    _fields_ = [("total", C.c_double), ("centre", C.c_double * 3), ("inertia", C.c_double * 9),
                ("axes", C.c_double * 9), ("moments", C.c_double * 3), ("thickness", C.c_double)]


class Sdf(C.Structure):
    """talus_mi_sdf (SampledSdf without its brick summaries)."""

    _fields_ = [("origin", C.c_float * 3), ("voxel_size", C.c_float * 3), ("resolution", C.c_uint32 * 3),
                ("distances", C.POINTER(C.c_float)), ("materials", C.POINTER(C.c_uint8))]


class Pose(C.Structure):
    """talus_mi_pose: position and rotation quaternion w, x, y, z."""

    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 4)]


class SampleBox(C.Structure):
    _fields_ = [("minimum", C.c_uint32 * 3), ("extent", C.c_uint32 * 3)]


class Shape(C.Structure):
    """talus_mi_shape (ShapeProbe): a rigid partner's resident field placed
    by its centre of mass and rotation (w, x, y, z)."""

    _fields_ = [("field", C.c_uint32), ("center_of_mass", C.c_float * 3), ("rotation", C.c_float * 4),
                ("field_offset", C.c_float * 3), ("linear_velocity", C.c_float * 3),
                ("angular_velocity", C.c_float * 3)]


class ContactQuery(C.Structure):
    """talus_mi_contact_query: findContacts' dt, gravity, ContactGaps,
    voxel size and cull flag."""

    _fields_ = [("dt", C.c_float), ("gravity", C.c_float * 3), ("bias", C.c_float), ("threshold", C.c_float),
                ("voxel_size", C.c_float * 3), ("cull", C.c_int32)]


class SurfaceContact(C.Structure):
    """talus_mi_surface_contact (SurfaceContact)."""

    _fields_ = [("nodes", C.c_uint32 * 3), ("weights", C.c_float * 3), ("partner", C.c_int32),
                ("normal", C.c_float * 3), ("local_point", C.c_float * 3), ("gap", C.c_float),
                ("bias", C.c_float), ("friction", C.c_float), ("triangle", C.c_uint32)]


# numpy view of talus_mi_surface_contact
CONTACT_DTYPE = np.dtype([("nodes", "<u4", 3), ("weights", "<f4", 3), ("partner", "<i4"), ("normal", "<f4", 3),
                          ("local_point", "<f4", 3), ("gap", "<f4"), ("bias", "<f4"), ("friction", "<f4"),
                          ("triangle", "<u4")])
assert CONTACT_DTYPE.itemsize == C.sizeof(SurfaceContact)


def _sdf(field, keep):
    """talus_mi_sdf over a field dict (origin, voxel, res, distances,
    optional materials); arrays kept alive in ``keep``."""
    dist = np.ascontiguousarray(field["distances"], dtype=np.float32).reshape(-1)
    f = Sdf()
    f.origin[:] = [float(v) for v in field["origin"]]
    f.voxel_size[:] = [float(v) for v in field["voxel"]]
    f.resolution[:] = [int(v) for v in field["res"]]
    if dist.size != int(np.prod([int(v) for v in field["res"]])):
        raise TalusError("a field needs one distance per sample")
    f.distances = dist.ctypes.data_as(C.POINTER(C.c_float))
    keep.append(dist)
    if field.get("materials") is not None:
        mat = np.ascontiguousarray(field["materials"], dtype=np.uint8).reshape(-1)
        if mat.size != dist.size:
            raise TalusError("a field needs one material per sample")
        f.materials = mat.ctypes.data_as(C.POINTER(C.c_uint8))
        keep.append(mat)
    return f


class StepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float * 3), ("substeps", C.c_uint32),
        ("cg_iterations", C.c_uint32), ("factor_budget", C.c_uint32),
        ("tolerance", C.c_float), ("spectral_radius", C.c_float),
        ("preconditioner", C.c_uint32), ("flags", C.c_uint32),
        ("plastic_passes", C.c_uint32),
        ("refine_tolerance", C.c_float), ("refine_rounds", C.c_uint32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("bodies", C.c_uint32), ("nodes", C.c_uint32), ("elements", C.c_uint32),
        ("substeps", C.c_uint32), ("cg_iterations", C.c_uint32), ("cg_capped", C.c_uint32),
        ("solves", C.c_uint32), ("refactors", C.c_uint32), ("cg_worst_relative", C.c_double),
        ("elastic_energy", C.c_double), ("kinetic_energy", C.c_double),
        ("plastic_growth", C.c_uint32), ("yield_events", C.c_uint32),
        ("tear_events", C.c_uint32), ("split_nodes", C.c_uint32), ("non_finite", C.c_uint32),
        ("preconditioner_applications", C.c_uint32), ("device_seconds", C.c_double),
        ("plastic_evaluations", C.c_uint32), ("plastic_solves", C.c_uint32),
        ("plastic_backtracks", C.c_uint32), ("plastic_linear_reads", C.c_uint32),
        ("plastic_handoffs", C.c_uint32), ("plastic_fallbacks", C.c_uint32),
        ("plastic_rejections", C.c_uint32), ("plastic_capped", C.c_uint32),
        ("refinements", C.c_uint32), ("refine_unmet", C.c_uint32),
        ("refine_worst_relative", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FactorInfo(C.Structure):
    _fields_ = [("fronts", C.c_uint32), ("levels", C.c_uint32), ("ok", C.c_uint32),
                ("reserved", C.c_uint32), ("entries", C.c_uint64), ("flops", C.c_double),
                ("factor_seconds", C.c_double)]


class FactorKernel(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("launches", C.c_uint32), ("workgroups", C.c_uint64),
                ("mfma_flops", C.c_double), ("useful_flops", C.c_double), ("ms", C.c_double)]


FACTOR_KERNELS = {0: "assemble_pairs", 1: "extend_add", 2: "potrf", 3: "trsm", 4: "syrk", 5: "trtri",
                  6: "pack"}


class TuneReport(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("lanes_forward", C.c_uint32),
                ("lanes_backward", C.c_uint32), ("leaf", C.c_uint32),
                ("before_seconds", C.c_double), ("after_seconds", C.c_double)]


class KernelTiming(C.Structure):
    _fields_ = [("avg_ms", C.c_double), ("bytes", C.c_double), ("launches", C.c_uint32),
                ("reserved", C.c_uint32)]


class Event(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("body", C.c_uint32), ("tet", C.c_uint32),
                ("step", C.c_uint32)]


class _HexMesh(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint32), ("hexes", C.c_uint32),
        ("rest_positions", C.POINTER(C.c_float)), ("node_mass", C.POINTER(C.c_float)),
        ("anchored", C.POINTER(C.c_uint8)), ("hex_nodes", C.POINTER(C.c_uint32)),
        ("volume", C.POINTER(C.c_float)), ("hex_material", C.POINTER(C.c_uint32)),
        ("voxel_size", C.c_float * 3),
    ]


class HexElement(C.Structure):
    """makeElement's outputs (StrainElement, strainstep/element.h:40-73)."""

    _fields_ = [
        ("gradient", C.c_double * 24), ("inverse_mass", C.c_double * 8),
        ("rest_hourglass", C.c_double * 12), ("bend", C.c_double), ("size", C.c_double),
        ("stiffness", C.c_double * 18), ("rate", C.c_double * 18),
        ("softness", C.c_double * 18), ("cap", C.c_double * 18),
        ("inverse", C.c_double * 324), ("volume_stiffness", C.c_double),
        ("volume_rate", C.c_double), ("volume_softness", C.c_double),
        ("volume_cap", C.c_double), ("volumetric", C.c_uint32), ("reserved", C.c_uint32),
    ]

    def arr(self, name):
        return np.ctypeslib.as_array(getattr(self, name)).copy()


class LatticeParams(C.Structure):
    _fields_ = [("dt", C.c_double), ("gravity", C.c_double * 3),
                ("softstep_substeps", C.c_uint32), ("flags", C.c_uint32)]


class LatticeReport(C.Structure):
    _fields_ = [("lattices", C.c_uint32), ("substeps", C.c_uint32), ("elements", C.c_uint32),
                ("rebuilds", C.c_uint32), ("element_solves", C.c_uint64), ("device_ms", C.c_double),
                ("inverses", C.c_uint32), ("solve_launches", C.c_uint32),
                ("update_launches", C.c_uint32), ("refreshed", C.c_uint32),
                ("solve_ms", C.c_double), ("update_ms", C.c_double),
                ("torn", C.c_uint32), ("node_copies", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every entry point of include/talus_mi.h with its ctypes signature
_P = C.POINTER
_SIGNATURES = {
    "talus_mi_create": (C.c_int, [_P(C.c_void_p), C.c_int, C.c_uint32]),
    "talus_mi_destroy": (C.c_int, [C.c_void_p]),
    "talus_mi_last_error": (C.c_char_p, []),
    "talus_mi_default_params": (None, [_P(StepParams)]),
    "talus_mi_build_lattice": (C.c_int, [_P(_Lattice), _P(PhysicalMaterial), C.c_uint32,
                                         _P(Support), C.c_uint32, _P(_Mesh)]),
    "talus_mi_mesh_free": (None, [_P(_Mesh)]),
    "talus_mi_free": (None, [C.c_void_p]),
    "talus_mi_tear_mesh": (C.c_int, [_P(_Mesh), _P(PhysicalMaterial), C.c_uint32, _P(C.c_double),
                                     _P(C.c_uint32), C.c_uint32, _P(_Mesh), _P(_P(C.c_uint32)),
                                     _P(C.c_uint32)]),
    "talus_mi_body_add": (C.c_int, [C.c_void_p, _P(_Mesh), _P(PhysicalMaterial), C.c_uint32,
                                    _P(C.c_uint32)]),
    "talus_mi_body_info": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), _P(C.c_uint32)]),
    "talus_mi_state_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                     _P(C.c_double)]),
    "talus_mi_state_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                     _P(C.c_double)]),
    "talus_mi_rotations_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_rotations_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_frames_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_frames_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_plastic_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_plastic_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_topology_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), _P(C.c_double),
                                        _P(C.c_double), _P(C.c_uint8)]),
    "talus_mi_step": (C.c_int, [C.c_void_p, _P(StepParams), _P(Report)]),
    "talus_mi_events": (C.c_int, [C.c_void_p, _P(Event), C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_body_report": (C.c_int, [C.c_void_p, C.c_uint32, _P(Report)]),
    "talus_mi_field_add": (C.c_int, [C.c_void_p, _P(Sdf), _P(C.c_uint32)]),
    "talus_mi_field_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(Sdf)]),
    "talus_mi_field_lipschitz": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_float)]),
    "talus_mi_find_contacts": (C.c_int, [C.c_void_p, C.c_uint32, _P(Shape), C.c_uint32, _P(ContactQuery),
                                         _P(C.c_float), _P(SurfaceContact), C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_eval_forces": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                       _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
    "talus_mi_apply_system": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                        _P(C.c_double), _P(C.c_double)]),
    "talus_mi_apply_system_stressed": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), C.c_double, C.c_double,
                                                 _P(C.c_double), _P(C.c_double)]),
    "talus_mi_apply_system_damped": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), C.c_double, C.c_double,
                                               C.c_double, _P(C.c_double), _P(C.c_double)]),
    "talus_mi_apply_system_noalloc": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                                C.c_void_p, C.c_void_p]),
    "talus_mi_time_kernel": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _P(KernelTiming)]),
    "talus_mi_factor_profile": (C.c_int, [C.c_void_p, _P(FactorKernel), C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_body_frame_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_factor_solve": (C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double,
                                        _P(C.c_double), _P(C.c_double)]),
    "talus_mi_factor_info_get": (C.c_int, [C.c_void_p, _P(FactorInfo)]),
    "talus_mi_tune": (C.c_int, [C.c_void_p, C.c_uint32, _P(TuneReport)]),
    "talus_mi_lame_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
    "talus_mi_lame_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
    "talus_mi_rest_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_rollout": (C.c_int, [C.c_void_p, _P(StepParams), C.c_uint32, C.c_uint32, _P(Report)]),
    "talus_mi_rollout_backward": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                            C.c_double, _P(C.c_double), _P(C.c_double), _P(C.c_double),
                                            _P(C.c_double), _P(C.c_double), _P(C.c_double)]),
    "talus_mi_plan_get": (C.c_int, [C.c_void_p, _P(C.c_uint8), C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_plan_set": (C.c_int, [C.c_void_p, _P(C.c_uint8), C.c_uint32]),
    "talus_mi_set_streams": (C.c_int, [C.c_void_p, C.c_uint32]),
    "talus_mi_mesh_surface": (C.c_int, [C.c_void_p, C.c_void_p]),
    "talus_mi_hex_mesh_surface": (C.c_int, [C.c_void_p, C.c_void_p]),
    "talus_mi_lattice_surface_mesh": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "talus_mi_surface_thickness": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_uint32, _P(C.c_uint32),
                                             _P(C.c_float)]),
    "talus_mi_surface_get": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "talus_mi_surface_free": (None, [C.c_void_p]),
    "talus_mi_stream": (C.c_void_p, [C.c_void_p]),
    "talus_mi_build_hex_lattice": (C.c_int, [_P(_Lattice), _P(PhysicalMaterial), C.c_uint32,
                                             _P(Support), C.c_uint32, _P(_HexMesh)]),
    "talus_mi_hex_mesh_free": (None, [_P(_HexMesh)]),
    "talus_mi_hex_element_make": (C.c_int, [_P(C.c_double), _P(C.c_double), C.c_double, C.c_float,
                                            C.c_float, C.c_double, _P(HexElement)]),
    "talus_mi_lattice_default_params": (None, [_P(LatticeParams)]),
    "talus_mi_lattice_add": (C.c_int, [C.c_void_p, _P(_HexMesh), _P(PhysicalMaterial), C.c_uint32,
                                       _P(C.c_double), _P(C.c_double), _P(C.c_uint32)]),
    "talus_mi_lattice_set_substeps": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "talus_mi_lattice_state_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
    "talus_mi_lattice_state_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
    "talus_mi_lattice_impulses_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), _P(C.c_double),
                                                C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_lattice_colours_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint8), _P(C.c_uint8)]),
    "talus_mi_lattice_step": (C.c_int, [C.c_void_p, _P(LatticeParams), _P(LatticeReport)]),
    "talus_mi_lattice_topology_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), _P(C.c_uint32),
                                                _P(C.c_float), _P(C.c_uint8)]),
    "talus_mi_lattice_tear": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_lattice_cell_stretch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, _P(C.c_double),
                                                _P(C.c_double)]),
    "talus_mi_lattice_pieces": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32)]),
    "talus_mi_lattice_snh_terms": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                             _P(C.c_double), _P(C.c_double)]),
    "talus_mi_lattice_energy": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), C.c_void_p]),
    "talus_mi_lattice_rest_body": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
    "talus_mi_lattice_bake_field": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                              C.c_void_p]),
    "talus_mi_lattice_screen": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), C.c_uint32, _P(C.c_double),
                                          _P(C.c_double), _P(C.c_double), _P(C.c_uint32)]),
    "talus_mi_lattice_lowest_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double)]),
    "talus_mi_lattice_surface_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_uint32), C.c_uint32,
                                               _P(C.c_uint32)]),
    "talus_mi_lattice_plastic_get": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double),
                                               _P(C.c_uint32)]),
    "talus_mi_tune_leaf": (C.c_int, [C.c_void_p, _P(C.c_uint32), C.c_uint32, C.c_uint32, _P(C.c_double),
                                     _P(TuneReport)]),
    "talus_mi_set_leaf": (C.c_int, [C.c_void_p, C.c_uint32]),
    "talus_mi_leaf_get": (C.c_int, [C.c_void_p, _P(C.c_uint32)]),
    "talus_mi_lattice_plastic_set": (C.c_int, [C.c_void_p, C.c_uint32, _P(C.c_double), _P(C.c_double)]),
}

_lib = None


def lib():
    """Loads the in-tree libtalus_mi.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TalusError(f"{LIB_PATH} is missing: run `python -m talus_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().talus_mi_last_error()
        raise TalusError(f"status {rc}: {msg.decode() if msg else ''}")


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _materials(materials):
    """ctypes array (or pointer into a numpy record array with the
    PhysicalMaterial layout) of the material lanes."""
    if isinstance(materials, np.ndarray):
        if materials.dtype.itemsize != C.sizeof(PhysicalMaterial):
            raise ValueError("material records must have the PhysicalMaterial layout")
        arr = np.ascontiguousarray(materials)
        ptr = arr.ctypes.data_as(C.POINTER(PhysicalMaterial))
        ptr._keep = arr  # keep the buffer alive with the pointer
        return ptr
    arr = (PhysicalMaterial * max(len(materials), 1))()
    for i, m in enumerate(materials):
        arr[i] = m
    return arr


def build_lattice(resolution, voxel_size, origin=(0.0, 0.0, 0.0), materials=(), supports=(),
                  sample_positions=None, cell_lane=None, solid=None):
    """Kuhn tet mesh of a structured lattice; returns a dict of numpy arrays
    (rest_positions, node_mass, anchored, tet_nodes, rest_inverse, volume, tet_material)."""
    L = lib()
    lat = _Lattice()
    lat.resolution[:] = [int(r) for r in resolution]
    lat.voxel_size[:] = [float(v) for v in voxel_size]
    lat.origin[:] = [float(v) for v in origin]
    keep = []
    if sample_positions is not None:
        sp = np.ascontiguousarray(sample_positions, dtype=np.float32).reshape(-1)
        keep.append(sp)
        lat.sample_positions = sp.ctypes.data_as(C.POINTER(C.c_float))
    if cell_lane is not None:
        cl = np.ascontiguousarray(cell_lane, dtype=np.uint32)
        keep.append(cl)
        lat.cell_lane = cl.ctypes.data_as(C.POINTER(C.c_uint32))
    if solid is not None:
        so = np.ascontiguousarray(solid, dtype=np.uint8)
        keep.append(so)
        lat.solid = so.ctypes.data_as(C.POINTER(C.c_uint8))
    mats = _materials(materials if isinstance(materials, np.ndarray) else list(materials))
    sups = (Support * max(len(supports), 1))()
    for i, s in enumerate(supports):
        sups[i] = s
    out = _Mesh()
    _check(L.talus_mi_build_lattice(C.byref(lat), mats, len(materials), sups, len(supports),
                                    C.byref(out)))
    n, t = out.nodes, out.tets
    try:
        mesh = {
            "rest_positions": np.ctypeslib.as_array(out.rest_positions, (3 * n,)).reshape(n, 3).copy(),
            "node_mass": np.ctypeslib.as_array(out.node_mass, (n,)).copy(),
            "anchored": np.ctypeslib.as_array(out.anchored, (n,)).copy(),
            "tet_nodes": np.ctypeslib.as_array(out.tet_nodes, (4 * t,)).reshape(t, 4).copy(),
            "rest_inverse": np.ctypeslib.as_array(out.rest_inverse, (9 * t,)).reshape(t, 3, 3).copy(),
            "volume": np.ctypeslib.as_array(out.volume, (t,)).copy(),
            "tet_material": np.ctypeslib.as_array(out.tet_material, (t,)).copy(),
        }
    finally:
        L.talus_mi_mesh_free(C.byref(out))
    return mesh


def _mesh_in(mesh, keep):
    rest = np.ascontiguousarray(mesh["rest_positions"], dtype=np.float32).reshape(-1)
    anchored = np.ascontiguousarray(mesh["anchored"], dtype=np.uint8)
    tets = np.ascontiguousarray(mesh["tet_nodes"], dtype=np.uint32).reshape(-1)
    lanes = np.ascontiguousarray(mesh["tet_material"], dtype=np.uint32)
    keep += [rest, anchored, tets, lanes]
    m = _Mesh()
    m.nodes = anchored.size
    m.tets = lanes.size
    m.rest_positions = rest.ctypes.data_as(C.POINTER(C.c_float))
    m.anchored = anchored.ctypes.data_as(C.POINTER(C.c_uint8))
    m.tet_nodes = tets.ctypes.data_as(C.POINTER(C.c_uint32))
    m.tet_material = lanes.ctypes.data_as(C.POINTER(C.c_uint32))
    return m


def _mesh_out(out):
    n, t = out.nodes, out.tets
    return {
        "rest_positions": np.ctypeslib.as_array(out.rest_positions, (3 * n,)).reshape(n, 3).copy(),
        "node_mass": np.ctypeslib.as_array(out.node_mass, (n,)).copy(),
        "anchored": np.ctypeslib.as_array(out.anchored, (n,)).copy(),
        "tet_nodes": np.ctypeslib.as_array(out.tet_nodes, (4 * t,)).reshape(t, 4).copy(),
        "rest_inverse": np.ctypeslib.as_array(out.rest_inverse, (9 * t,)).reshape(t, 3, 3).copy(),
        "volume": np.ctypeslib.as_array(out.volume, (t,)).copy(),
        "tet_material": np.ctypeslib.as_array(out.tet_material, (t,)).copy(),
    }


def tear_mesh(mesh, materials, x, torn):
    """Tear topology of a host mesh at positions x (talus_mi_tear_mesh);
    returns (new mesh dict, origin of every new node)."""
    L = lib()
    keep = []
    m = _mesh_in(mesh, keep)
    mats = _materials(materials if isinstance(materials, np.ndarray) else list(materials))
    x = _f64(x)
    torn = np.ascontiguousarray(torn, dtype=np.uint32)
    out = _Mesh()
    origin = C.POINTER(C.c_uint32)()
    added = C.c_uint32()
    _check(L.talus_mi_tear_mesh(C.byref(m), mats, len(materials), _dp(x),
                                torn.ctypes.data_as(C.POINTER(C.c_uint32)), torn.size, C.byref(out),
                                C.byref(origin), C.byref(added)))
    try:
        new = _mesh_out(out)
        org = np.ctypeslib.as_array(origin, (max(added.value, 1),))[: added.value].copy()
    finally:
        L.talus_mi_mesh_free(C.byref(out))
        L.talus_mi_free(origin)
    return new, org


class _Surface(C.Structure):
    _fields_ = [("triangles", C.c_uint32), ("edges", C.c_uint32), ("nodes", C.c_uint32),
                ("tets", C.c_uint32), ("triangle_nodes", C.POINTER(C.c_uint32)),
                ("triangle_tet", C.POINTER(C.c_uint32)), ("triangle_lane", C.POINTER(C.c_uint32)),
                ("edge_nodes", C.POINTER(C.c_uint32)), ("surface_nodes", C.POINTER(C.c_uint32)),
                ("crack_faces", C.POINTER(C.c_uint8)), ("clusters", C.c_uint32),
                ("cluster_faces", C.POINTER(C.c_uint32)), ("cluster_starts", C.POINTER(C.c_uint32))]


def _surface_out(s: _Surface) -> dict:
    def arr(ptr, n):
        return np.ctypeslib.as_array(ptr, (max(n, 1),))[:n].copy()

    return {"triangles": arr(s.triangle_nodes, 3 * s.triangles).reshape(-1, 3),
            "triangle_tet": arr(s.triangle_tet, s.triangles),
            "triangle_lane": arr(s.triangle_lane, s.triangles),
            "edges": arr(s.edge_nodes, 2 * s.edges).reshape(-1, 2),
            "nodes": arr(s.surface_nodes, s.nodes),
            "crack_faces": arr(s.crack_faces, s.tets),
            "cluster_faces": arr(s.cluster_faces, s.triangles),
            "cluster_starts": arr(s.cluster_starts, s.clusters + 1 if s.triangles else 0)}


def surface_thickness(field, rest, triangles):
    """surfaceThickness of a surface's triangles over a lattice field (a dict
    of origin, voxel, res, distances) at rest positions (talus_mi_surface_thickness)."""
    dist = np.ascontiguousarray(field["distances"], dtype=np.float32).reshape(-1)
    f = Sdf()
    f.origin[:] = [float(v) for v in field["origin"]]
    f.voxel_size[:] = [float(v) for v in field["voxel"]]
    f.resolution[:] = [int(v) for v in field["res"]]
    f.distances = dist.ctypes.data_as(_P(C.c_float))
    r = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
    t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1)
    out = np.zeros(t.size // 3, dtype=np.float32)
    _check(lib().talus_mi_surface_thickness(C.byref(f), _dp(r), t.size // 3, t.ctypes.data_as(_P(C.c_uint32)),
                                            out.ctypes.data_as(_P(C.c_float))))
    return out


def _hex_mesh_in(mesh, keep) -> _HexMesh:
    """a talus_mi_hex_mesh over the dict's arrays (kept alive in `keep`)"""
    keep.update({
        "rest": np.ascontiguousarray(mesh["rest_positions"], dtype=np.float32).reshape(-1),
        "mass": np.ascontiguousarray(mesh["node_mass"], dtype=np.float32),
        "anchored": np.ascontiguousarray(mesh["anchored"], dtype=np.uint8),
        "hex": np.ascontiguousarray(mesh["hex_nodes"], dtype=np.uint32).reshape(-1),
        "volume": np.ascontiguousarray(mesh["volume"], dtype=np.float32),
        "lane": np.ascontiguousarray(mesh["hex_material"], dtype=np.uint32),
    })
    m = _HexMesh()
    m.nodes = keep["mass"].size
    m.hexes = keep["volume"].size
    m.rest_positions = keep["rest"].ctypes.data_as(C.POINTER(C.c_float))
    m.node_mass = keep["mass"].ctypes.data_as(C.POINTER(C.c_float))
    m.anchored = keep["anchored"].ctypes.data_as(C.POINTER(C.c_uint8))
    m.hex_nodes = keep["hex"].ctypes.data_as(C.POINTER(C.c_uint32))
    m.volume = keep["volume"].ctypes.data_as(C.POINTER(C.c_float))
    m.hex_material = keep["lane"].ctypes.data_as(C.POINTER(C.c_uint32))
    m.voxel_size[:] = [float(v) for v in mesh["voxel_size"]]
    return m


def hex_mesh_surface(mesh) -> dict:
    """collectSurface's hex faces with clusters of a host hex mesh
    (talus_mi_hex_mesh_surface); triangle_tet holds the cell."""
    L = lib()
    keep = {}
    m = _hex_mesh_in(mesh, keep)
    out = _Surface()
    _check(L.talus_mi_hex_mesh_surface(C.byref(m), C.byref(out)))
    try:
        return _surface_out(out)
    finally:
        L.talus_mi_surface_free(C.byref(out))


def mesh_surface(mesh) -> dict:
    """Boundary surface and crack faces of a host mesh (talus_mi_mesh_surface)."""
    L = lib()
    keep = []
    m = _mesh_in(mesh, keep)
    out = _Surface()
    _check(L.talus_mi_mesh_surface(C.byref(m), C.byref(out)))
    try:
        return _surface_out(out)
    finally:
        L.talus_mi_surface_free(C.byref(out))


def default_params(**kw) -> StepParams:
    p = StepParams()
    lib().talus_mi_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = list(v)
        else:
            setattr(p, k, v)
    return p


def build_hex_lattice(resolution, voxel_size, origin=(0.0, 0.0, 0.0), materials=(), supports=(),
                      sample_positions=None, cell_lane=None, solid=None):
    """Hex PreparedMesh of a structured lattice (buildMesh with hex elements);
    returns a dict of numpy arrays (rest_positions, node_mass, anchored,
    hex_nodes, volume, hex_material, voxel_size)."""
    L = lib()
    lat = _Lattice()
    lat.resolution[:] = [int(r) for r in resolution]
    lat.voxel_size[:] = [float(v) for v in voxel_size]
    lat.origin[:] = [float(v) for v in origin]
    keep = []
    for name, arr, ct in (("sample_positions", sample_positions, (np.float32, C.c_float)),
                          ("cell_lane", cell_lane, (np.uint32, C.c_uint32)),
                          ("solid", solid, (np.uint8, C.c_uint8))):
        if arr is not None:
            a = np.ascontiguousarray(arr, dtype=ct[0]).reshape(-1)
            keep.append(a)
            setattr(lat, name, a.ctypes.data_as(C.POINTER(ct[1])))
    mats = _materials(materials if isinstance(materials, np.ndarray) else list(materials))
    sups = (Support * max(len(supports), 1))()
    for i, s in enumerate(supports):
        sups[i] = s
    out = _HexMesh()
    _check(L.talus_mi_build_hex_lattice(C.byref(lat), mats, len(materials), sups, len(supports),
                                        C.byref(out)))
    n, h = out.nodes, out.hexes
    try:
        mesh = {
            "rest_positions": np.ctypeslib.as_array(out.rest_positions, (3 * n,)).reshape(n, 3).copy(),
            "node_mass": np.ctypeslib.as_array(out.node_mass, (n,)).copy(),
            "anchored": np.ctypeslib.as_array(out.anchored, (n,)).copy(),
            "hex_nodes": np.ctypeslib.as_array(out.hex_nodes, (8 * h,)).reshape(h, 8).copy(),
            "volume": np.ctypeslib.as_array(out.volume, (h,)).copy(),
            "hex_material": np.ctypeslib.as_array(out.hex_material, (h,)).copy(),
            "voxel_size": np.array(list(out.voxel_size), dtype=np.float32),
        }
    finally:
        L.talus_mi_hex_mesh_free(C.byref(out))
    return mesh


def hex_element_make(corners, inverse_mass, volume, youngs_modulus, poissons_ratio, h) -> HexElement:
    """makeElement of one hex from its 8 rest corners (cornerRest,
    elementMaterial, makeElement); raises when the hex builds no element."""
    c = _f64(corners, 24)
    w = _f64(inverse_mass, 8)
    e = HexElement()
    _check(lib().talus_mi_hex_element_make(_dp(c), _dp(w), float(volume), float(youngs_modulus),
                                           float(poissons_ratio), float(h), C.byref(e)))
    return e


def lattice_params(**kw) -> LatticeParams:
    p = LatticeParams()
    lib().talus_mi_lattice_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity[:] = [float(g) for g in v]
        else:
            setattr(p, k, v)
    return p


class Context:
    """One device context: bodies, device buffers and a HIP stream."""

    def __init__(self, device: int = 0, precision: int = FP64):
        h = C.c_void_p()
        _check(lib().talus_mi_create(C.byref(h), device, precision))
        self._h = h
        self.precision = precision
        self.bodies = []  # (nodes, tets)
        self.lattices = []  # (nodes, hexes)

    def close(self):
        if getattr(self, "_h", None):
            lib().talus_mi_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add_body(self, mesh, materials, use_mesh_mass=False) -> int:
        """Adds a tet body (copied); masses are rho V / 4 in FP64 unless
        use_mesh_mass passes the mesh's node_mass."""
        rest = np.ascontiguousarray(mesh["rest_positions"], dtype=np.float32).reshape(-1)
        anchored = np.ascontiguousarray(mesh["anchored"], dtype=np.uint8)
        tets = np.ascontiguousarray(mesh["tet_nodes"], dtype=np.uint32).reshape(-1)
        lanes = np.ascontiguousarray(mesh["tet_material"], dtype=np.uint32)
        m = _Mesh()
        m.nodes = anchored.size
        m.tets = lanes.size
        m.rest_positions = rest.ctypes.data_as(C.POINTER(C.c_float))
        m.anchored = anchored.ctypes.data_as(C.POINTER(C.c_uint8))
        m.tet_nodes = tets.ctypes.data_as(C.POINTER(C.c_uint32))
        m.tet_material = lanes.ctypes.data_as(C.POINTER(C.c_uint32))
        if use_mesh_mass:
            mass = np.ascontiguousarray(mesh["node_mass"], dtype=np.float32)
            m.node_mass = mass.ctypes.data_as(C.POINTER(C.c_float))
        mats = _materials(materials if isinstance(materials, np.ndarray) else list(materials))
        body = C.c_uint32()
        _check(lib().talus_mi_body_add(self._h, C.byref(m), mats, len(materials), C.byref(body)))
        self.bodies.append((int(m.nodes), int(m.tets)))
        return int(body.value)

    def info(self, body):
        n, t = C.c_uint32(), C.c_uint32()
        _check(lib().talus_mi_body_info(self._h, body, C.byref(n), C.byref(t)))
        return int(n.value), int(t.value)

    def set_state(self, body, x=None, v=None, a=None):
        n, _ = self.info(body)
        x, v, a = _f64(x, 3 * n), _f64(v, 3 * n), _f64(a, 3 * n)
        _check(lib().talus_mi_state_set(self._h, body, _dp(x), _dp(v), _dp(a)))

    def state(self, body):
        n, _ = self.info(body)
        x, v, a = (np.empty(3 * n) for _ in range(3))
        _check(lib().talus_mi_state_get(self._h, body, _dp(x), _dp(v), _dp(a)))
        return x.reshape(n, 3), v.reshape(n, 3), a.reshape(n, 3)

    def rotations(self, body):
        _, t = self.info(body)
        R = np.empty(9 * t)
        _check(lib().talus_mi_rotations_get(self._h, body, _dp(R)))
        return R.reshape(t, 3, 3)

    def set_rotations(self, body, R):
        _, t = self.info(body)
        R = _f64(R, 9 * t)
        _check(lib().talus_mi_rotations_set(self._h, body, _dp(R)))

    def frames(self, body):
        _, t = self.info(body)
        W = np.empty(9 * t)
        _check(lib().talus_mi_frames_get(self._h, body, _dp(W)))
        return W.reshape(t, 3, 3)

    def set_frames(self, body, W):
        _, t = self.info(body)
        W = _f64(W, 9 * t)
        _check(lib().talus_mi_frames_set(self._h, body, _dp(W)))

    def plastic(self, body):
        _, t = self.info(body)
        p = np.empty(10 * t)
        _check(lib().talus_mi_plastic_get(self._h, body, _dp(p)))
        return p.reshape(t, 10)

    def set_plastic(self, body, p):
        _, t = self.info(body)
        p = _f64(p, 10 * t)
        _check(lib().talus_mi_plastic_set(self._h, body, _dp(p)))

    def topology(self, body):
        n, t = self.info(body)
        tets = np.empty(4 * t, dtype=np.uint32)
        rest = np.empty(3 * n)
        mass = np.empty(n)
        anchored = np.empty(n, dtype=np.uint8)
        _check(lib().talus_mi_topology_get(self._h, body, tets.ctypes.data_as(C.POINTER(C.c_uint32)),
                                           _dp(rest), _dp(mass),
                                           anchored.ctypes.data_as(C.POINTER(C.c_uint8))))
        return {"tet_nodes": tets.reshape(t, 4), "rest": rest.reshape(n, 3), "mass": mass,
                "anchored": anchored}

    def surface(self, body) -> dict:
        """Boundary surface and crack faces of a body's current mesh."""
        out = _Surface()
        _check(lib().talus_mi_surface_get(self._h, body, C.byref(out)))
        try:
            return _surface_out(out)
        finally:
            lib().talus_mi_surface_free(C.byref(out))

    def set_streams(self, streams: int):
        """Split the bodies over `streams` concurrently stepped engines."""
        _check(lib().talus_mi_set_streams(self._h, streams))

    def step(self, params: StepParams | None = None, **kw) -> dict:
        p = params if params is not None else default_params(**kw)
        r = Report()
        _check(lib().talus_mi_step(self._h, C.byref(p), C.byref(r)))
        return r.as_dict()

    def body_report(self, body) -> dict:
        """the last step's report of one body (talus_mi_body_report)"""
        r = Report()
        _check(lib().talus_mi_body_report(self._h, body, C.byref(r)))
        return r.as_dict()

    def field_add(self, field) -> int:
        """Copies a sampled field (dict: origin, voxel, res, distances,
        optional materials) to the device for contact queries
        (talus_mi_field_add); returns its id."""
        keep = []
        f = _sdf(field, keep)
        i = C.c_uint32()
        _check(lib().talus_mi_field_add(self._h, C.byref(f), C.byref(i)))
        return i.value

    def field_set(self, fid, field):
        """Replaces a resident field's samples (talus_mi_field_set)."""
        keep = []
        f = _sdf(field, keep)
        _check(lib().talus_mi_field_set(self._h, fid, C.byref(f)))

    def field_lipschitz(self, fid) -> float:
        """probeLipschitz of a resident field (talus_mi_field_lipschitz)."""
        out = C.c_float()
        _check(lib().talus_mi_field_lipschitz(self._h, fid, C.byref(out)))
        return out.value

    def find_contacts(self, body, shapes, dt, bias=0.0, threshold=0.0, cull=True, gravity=(0.0, 0.0, 0.0),
                      voxel=(1.0, 1.0, 1.0), thickness=None, capacity=None):
        """findContacts of a body's surface at its state (talus_mi_find_contacts).
        shapes: dicts with field (a field_add id) and optional center,
        rotation (w, x, y, z), offset, linear, angular.  Returns the contacts
        as a CONTACT_DTYPE array and how many there are in all."""
        arr = (Shape * max(1, len(shapes)))()
        for i, sh in enumerate(shapes):
            s = arr[i]
            s.field = int(sh["field"])
            s.center_of_mass[:] = [float(v) for v in sh.get("center", (0, 0, 0))]
            s.rotation[:] = [float(v) for v in sh.get("rotation", (1, 0, 0, 0))]
            s.field_offset[:] = [float(v) for v in sh.get("offset", (0, 0, 0))]
            s.linear_velocity[:] = [float(v) for v in sh.get("linear", (0, 0, 0))]
            s.angular_velocity[:] = [float(v) for v in sh.get("angular", (0, 0, 0))]
        q = ContactQuery()
        q.dt, q.bias, q.threshold, q.cull = float(dt), float(bias), float(threshold), int(bool(cull))
        q.gravity[:] = [float(v) for v in gravity]
        q.voxel_size[:] = [float(v) for v in voxel]
        th = None
        if thickness is not None:
            th = np.ascontiguousarray(thickness, dtype=np.float32).reshape(-1)
            if th.size != self.surface(body)["triangles"].shape[0]:
                raise TalusError("thickness needs one value per surface triangle")
        thp = None if th is None else th.ctypes.data_as(_P(C.c_float))
        found = C.c_uint32()
        cap = 4096 if capacity is None else int(capacity)
        while True:
            out = (SurfaceContact * max(1, cap))()
            _check(lib().talus_mi_find_contacts(self._h, body, arr, len(shapes), C.byref(q), thp, out, cap,
                                                C.byref(found)))
            if capacity is not None or found.value <= cap:
                break
            cap = found.value
        n = min(cap, found.value)
        return np.frombuffer(bytes(out), dtype=CONTACT_DTYPE)[:n].copy(), found.value

    def events(self, capacity=1 << 20):
        buf = (Event * capacity)()
        n = C.c_uint32()
        _check(lib().talus_mi_events(self._h, buf, capacity, C.byref(n)))
        return [(e.kind, e.body, e.tet, e.step) for e in buf[: n.value]]

    def eval_forces(self, body, x):
        n, t = self.info(body)
        x = _f64(x, 3 * n)
        tf, te, nf, R = np.empty(12 * t), np.empty(t), np.empty(3 * n), np.empty(9 * t)
        _check(lib().talus_mi_eval_forces(self._h, body, _dp(x), _dp(tf), _dp(te), _dp(nf), _dp(R)))
        return tf.reshape(t, 4, 3), te, nf.reshape(n, 3), R.reshape(t, 3, 3)

    def apply_system(self, body, inertia, step, d):
        n, _ = self.info(body)
        d = _f64(d, 3 * n)
        y = np.empty(3 * n)
        _check(lib().talus_mi_apply_system(self._h, body, inertia, step, _dp(d), _dp(y)))
        return y.reshape(n, 3)

    def apply_system_damped(self, body, x, inertia, step, damp, d):
        """The operator with the stiffness-damping term at positions x
        (talus_mi_apply_system_damped)."""
        n, _ = self.info(body)
        x, d = _f64(x, 3 * n), _f64(d, 3 * n)
        y = np.empty(3 * n)
        _check(lib().talus_mi_apply_system_damped(self._h, body, _dp(x), inertia, step, damp, _dp(d), _dp(y)))
        return y.reshape(n, 3)

    def apply_system_stressed(self, body, x, inertia, step, d):
        """The operator with the initial-stress term at positions x
        (talus_mi_apply_system_stressed)."""
        n, _ = self.info(body)
        x, d = _f64(x, 3 * n), _f64(d, 3 * n)
        y = np.empty(3 * n)
        _check(lib().talus_mi_apply_system_stressed(self._h, body, _dp(x), inertia, step, _dp(d), _dp(y)))
        return y.reshape(n, 3)

    def apply_system_noalloc(self, body, inertia, step, d_ptr: int, y_ptr: int):
        """The operator on device vectors (addresses of 3*nodes doubles),
        enqueued on the context's stream with no allocation; y is ready once
        that stream has synchronized (talus_mi_apply_system_noalloc)."""
        _check(lib().talus_mi_apply_system_noalloc(self._h, body, inertia, step, C.c_void_p(d_ptr),
                                                   C.c_void_p(y_ptr)))

    def time_kernel(self, kind, repeats=20) -> dict:
        """Per-launch HIP-event time and algorithmic bytes of one kernel
        (kinds: 0 assembly, 1 operator element, 2 operator gather, 3/4 solve
        forward/backward lane shape, 5/6 rows shape, 7/8 split shape)."""
        t = KernelTiming()
        _check(lib().talus_mi_time_kernel(self._h, kind, repeats, C.byref(t)))
        return {"avg_ms": t.avg_ms, "bytes": t.bytes, "launches": t.launches}

    def factor_solve(self, body, inertia, step, r):
        n, _ = self.info(body)
        r = _f64(r, 3 * n)
        z = np.empty(3 * n)
        _check(lib().talus_mi_factor_solve(self._h, body, inertia, step, _dp(r), _dp(z)))
        return z.reshape(n, 3)

    def body_frame(self, body):
        """bodyRotation of the body at its current positions (3x3)."""
        R = np.empty(9)
        _check(lib().talus_mi_body_frame_get(self._h, body, _dp(R)))
        return R.reshape(3, 3)

    def factor_profile(self) -> dict:
        """One factorization with HIP events around every launch: per kernel
        kind its launches, workgroups, MFMA-issued and useful FP64 flops and
        summed time (talus_mi_factor_profile)."""
        out = (FactorKernel * 16)()
        n = C.c_uint32()
        _check(lib().talus_mi_factor_profile(self._h, out, 16, C.byref(n)))
        return {FACTOR_KERNELS[out[i].kind]: {k: getattr(out[i], k) for k, _ in FactorKernel._fields_ if k != "kind"}
                for i in range(n.value)}

    def factor_info(self) -> dict:
        i = FactorInfo()
        _check(lib().talus_mi_factor_info_get(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in FactorInfo._fields_ if k != "reserved"}

    def set_lame(self, body, lam, mu):
        """Per-tet Lame parameters in FP64 (override the lanes' E and nu)."""
        _, t = self.info(body)
        lam, mu = _f64(lam, t), _f64(mu, t)
        _check(lib().talus_mi_lame_set(self._h, body, _dp(lam), _dp(mu)))

    def lame(self, body):
        _, t = self.info(body)
        lam, mu = np.empty(t), np.empty(t)
        _check(lib().talus_mi_lame_get(self._h, body, _dp(lam), _dp(mu)))
        return lam, mu

    def set_rest(self, body, rest):
        """FP64 rest positions; rest inverse, volumes and masses follow."""
        n, _ = self.info(body)
        rest = _f64(rest, 3 * n)
        _check(lib().talus_mi_rest_set(self._h, body, _dp(rest)))

    def rollout(self, steps, every=1, params: StepParams | None = None, **kw) -> dict:
        """`steps` converged steps, checkpointing every `every` steps."""
        p = params if params is not None else default_params(**kw)
        r = Report()
        _check(lib().talus_mi_rollout(self._h, C.byref(p), steps, every, C.byref(r)))
        return r.as_dict()

    def rollout_backward(self, body, x_bar, v_bar=None, tolerance=1e-12) -> dict:
        """Gradients of the last rollout for one body: dL/dlambda, dL/dmu
        (per tet), dL/drest, dL/dx0, dL/dv0, dL/da0 (per node)."""
        n, t = self.info(body)
        xb = _f64(x_bar, 3 * n)
        vb = _f64(v_bar, 3 * n) if v_bar is not None else None
        out = {"lambda": np.empty(t), "mu": np.empty(t)}
        out.update({k: np.empty(3 * n) for k in ("rest", "x0", "v0", "a0")})
        _check(lib().talus_mi_rollout_backward(self._h, body, _dp(xb), _dp(vb) if vb is not None else None,
                                               tolerance, _dp(out["lambda"]), _dp(out["mu"]),
                                               _dp(out["rest"]), _dp(out["x0"]), _dp(out["v0"]),
                                               _dp(out["a0"])))
        for k in ("rest", "x0", "v0", "a0"):
            out[k] = out[k].reshape(n, 3)
        return out

    def tune(self, repeats=3) -> dict:
        """Autotunes the preconditioner solve's kernel shapes per level."""
        r = TuneReport()
        _check(lib().talus_mi_tune(self._h, repeats, C.byref(r)))
        return {k: getattr(r, k) for k, _ in TuneReport._fields_}

    def tune_leaf(self, candidates=(16, 24, 32, 48, 64, 96), repeats=3) -> dict:
        """Autotunes the supernode leaf size (talus_mi_tune_leaf): the
        chosen leaf's tune report plus each candidate's application time."""
        cand = (C.c_uint32 * len(candidates))(*candidates)
        times = (C.c_double * len(candidates))()
        r = TuneReport()
        _check(lib().talus_mi_tune_leaf(self._h, cand, len(candidates), repeats, times, C.byref(r)))
        out = {k: getattr(r, k) for k, _ in TuneReport._fields_}
        out["candidates"] = {int(c): t for c, t in zip(candidates, times)}
        return out

    def set_leaf(self, leaf: int):
        """Supernode leaf size of the block Cholesky preconditioner
        (talus_mi_set_leaf); rebuilds at the next call, keeps the state."""
        _check(lib().talus_mi_set_leaf(self._h, leaf))

    def leaf(self) -> int:
        out = C.c_uint32()
        _check(lib().talus_mi_leaf_get(self._h, C.byref(out)))
        return out.value

    def plan(self) -> np.ndarray:
        n = C.c_uint32()
        _check(lib().talus_mi_plan_get(self._h, None, 0, C.byref(n)))
        p = np.zeros(n.value, dtype=np.uint8)
        if n.value:
            _check(lib().talus_mi_plan_get(self._h, p.ctypes.data_as(_P(C.c_uint8)), n.value, C.byref(n)))
        return p

    def set_plan(self, plan):
        p = np.ascontiguousarray(plan, dtype=np.uint8)
        _check(lib().talus_mi_plan_set(self._h, p.ctypes.data_as(_P(C.c_uint8)), p.size))

    def stream(self) -> int:
        return int(lib().talus_mi_stream(self._h) or 0)

    # ---- strainstep hex lattices ------------------------------------------

    def add_lattice(self, mesh, materials, offset=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0)) -> int:
        """addLattice: a hex lattice body at rest + offset moving at velocity."""
        keep = {}
        m = _hex_mesh_in(mesh, keep)
        mats = _materials(materials if isinstance(materials, np.ndarray) else list(materials))
        off, vel = _f64(offset, 3), _f64(velocity, 3)
        out = C.c_uint32()
        _check(lib().talus_mi_lattice_add(self._h, C.byref(m), mats, len(materials), _dp(off), _dp(vel),
                                          C.byref(out)))
        self.lattices.append((int(m.nodes), int(m.hexes)))
        return int(out.value)

    def set_lattice_substeps(self, lattice, substeps):
        _check(lib().talus_mi_lattice_set_substeps(self._h, lattice, int(substeps)))

    def lattice_nodes(self, lattice) -> int:
        """the lattice's current node count (tearing adds node copies)"""
        n = C.c_uint32()
        _check(lib().talus_mi_lattice_topology_get(self._h, lattice, C.byref(n), None, None, None))
        return int(n.value)

    def lattice_topology(self, lattice):
        """(hex nodes [hexes, 8], node masses, crack-face bits per hex)"""
        h = self.lattices[lattice][1]
        n = self.lattice_nodes(lattice)
        hx = np.zeros(8 * h, dtype=np.uint32)
        mass = np.zeros(n, dtype=np.float32)
        crack = np.zeros(h, dtype=np.uint8)
        _check(lib().talus_mi_lattice_topology_get(self._h, lattice, None, hx.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                   mass.ctypes.data_as(C.POINTER(C.c_float)),
                                                   crack.ctypes.data_as(C.POINTER(C.c_uint8))))
        return hx.reshape(-1, 8), mass, crack

    def lattice_tear(self, lattice, cells) -> int:
        """tearCells at the lattice's current positions: the node copies made"""
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        out = C.c_uint32()
        _check(lib().talus_mi_lattice_tear(self._h, lattice, c.ctypes.data_as(C.POINTER(C.c_uint32)), c.size,
                                           C.byref(out)))
        return int(out.value)

    def lattice_cell_stretch(self, lattice, cell):
        s, d = C.c_double(), np.zeros(3)
        _check(lib().talus_mi_lattice_cell_stretch(self._h, lattice, int(cell), C.byref(s), _dp(d)))
        return s.value, d

    def lattice_lowest(self, lattice) -> float:
        """the lowest angular frequency the lattice's elements are damped for"""
        out = C.c_double()
        _check(lib().talus_mi_lattice_lowest_get(self._h, lattice, C.byref(out)))
        return float(out.value)

    def lattice_surface(self, lattice):
        """surfaceNodes of the lattice, ascending"""
        n = C.c_uint32()
        _check(lib().talus_mi_lattice_surface_get(self._h, lattice, None, 0, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=np.uint32)
        _check(lib().talus_mi_lattice_surface_get(self._h, lattice, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                  n.value, C.byref(n)))
        return out[: n.value]

    def lattice_snh_terms(self, lattice, x=None):
        """(cell energies, node gradients (n, 3), node Hessian blocks (n, 3, 3))
        of the FP64 neo-Hookean hex at x (default: the current positions)
        (talus_mi_lattice_snh_terms)."""
        n = self.lattice_nodes(lattice)
        cells = self.lattices[lattice][1]
        xp = None
        if x is not None:
            x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
            assert x.size == 3 * n
            xp = _dp(x)
        e, g, h = np.empty(cells), np.empty(3 * n), np.empty(9 * n)
        _check(lib().talus_mi_lattice_snh_terms(self._h, lattice, xp, _dp(e), _dp(g), _dp(h)))
        return e, g.reshape(n, 3), h.reshape(n, 3, 3)

    def lattice_energy(self, lattice, gravity=(0.0, -9.81, 0.0)):
        """latticeEnergy: dict of kinetic, gravity, elastic (talus_mi_lattice_energy)."""
        out = Energy()
        g = np.ascontiguousarray(gravity, dtype=np.float64)
        _check(lib().talus_mi_lattice_energy(self._h, lattice, _dp(g), C.byref(out)))
        return {"kinetic": out.kinetic, "gravity": out.gravity, "elastic": out.elastic}

    def lattice_rest_body(self, lattice):
        """restBodyOf (talus_mi_lattice_rest_body)."""
        rb = RestBody()
        _check(lib().talus_mi_lattice_rest_body(self._h, lattice, C.byref(rb)))
        return {"total": rb.total, "centre": np.array(rb.centre[:]),
                "inertia": np.array(rb.inertia[:]).reshape(3, 3), "axes": np.array(rb.axes[:]).reshape(3, 3),
                "moments": np.array(rb.moments[:]), "thickness": rb.thickness}

    def lattice_bake_field(self, lattice, rest, current=None, position=(0.0, 0.0, 0.0),
                           rotation=(1.0, 0.0, 0.0, 0.0), carve=False):
        """bakeDeformedField of the lattice at its current positions
        (talus_mi_lattice_bake_field).  Fields are dicts of origin, voxel,
        res (x fastest) and distances (float32); returns (the new current
        distances, (box minimum, box extent))."""
        def sdf(field, dist):
            f = Sdf()
            f.origin[:] = [float(v) for v in field["origin"]]
            f.voxel_size[:] = [float(v) for v in field["voxel"]]
            f.resolution[:] = [int(v) for v in field["res"]]
            f.distances = dist.ctypes.data_as(_P(C.c_float))
            f.materials = None
            return f
        rd = np.ascontiguousarray(rest["distances"], dtype=np.float32).reshape(-1)
        cur = rest if current is None else current
        cd = np.array(cur["distances"], dtype=np.float32).reshape(-1)
        pose = Pose()
        pose.position[:] = [float(v) for v in position]
        pose.rotation[:] = [float(v) for v in rotation]
        box = SampleBox()
        rs, cs = sdf(rest, rd), sdf(cur, cd)
        _check(lib().talus_mi_lattice_bake_field(self._h, lattice, C.byref(rs), C.byref(pose), int(carve),
                                                 C.byref(cs), C.byref(box)))
        return cd, (np.array(box.minimum[:], dtype=np.uint32), np.array(box.extent[:], dtype=np.uint32))

    def lattice_screen(self, lattice, forces, held=()):
        """factorScreen + screenLoad: (displacement (n, 3), dent, iterations)
        (talus_mi_lattice_screen)."""
        n = self.lattice_nodes(lattice)
        f = np.ascontiguousarray(forces, dtype=np.float64).reshape(-1)
        assert f.size == 3 * n
        hl = np.ascontiguousarray(np.asarray(held, dtype=np.uint32).reshape(-1))
        hp = hl.ctypes.data_as(_P(C.c_uint32)) if hl.size else None
        u = np.empty(3 * n)
        dent = C.c_double()
        its = C.c_uint32()
        _check(lib().talus_mi_lattice_screen(self._h, lattice, hp, hl.size, _dp(f), _dp(u), C.byref(dent),
                                             C.byref(its)))
        return u.reshape(n, 3), dent.value, int(its.value)

    def lattice_surface_mesh(self, lattice) -> dict:
        """the hex surface of the lattice's current mesh (talus_mi_lattice_surface_mesh)"""
        out = _Surface()
        _check(lib().talus_mi_lattice_surface_mesh(self._h, lattice, C.byref(out)))
        try:
            return _surface_out(out)
        finally:
            lib().talus_mi_surface_free(C.byref(out))

    def lattice_pieces(self, lattice) -> int:
        out = C.c_uint32()
        _check(lib().talus_mi_lattice_pieces(self._h, lattice, C.byref(out)))
        return int(out.value)

    def lattice_state(self, lattice):
        n = self.lattice_nodes(lattice)
        x, v = np.empty(3 * n), np.empty(3 * n)
        _check(lib().talus_mi_lattice_state_get(self._h, lattice, _dp(x), _dp(v)))
        return x.reshape(n, 3), v.reshape(n, 3)

    def set_lattice_state(self, lattice, x=None, v=None):
        n = self.lattice_nodes(lattice)
        x, v = _f64(x, 3 * n), _f64(v, 3 * n)
        _check(lib().talus_mi_lattice_state_set(self._h, lattice, _dp(x), _dp(v)))

    def lattice_impulses(self, lattice):
        """(cells, impulses[elements, 19]) of the lattice's elements."""
        count = C.c_uint32()
        _check(lib().talus_mi_lattice_impulses_get(self._h, lattice, None, None, 0, C.byref(count)))
        k = count.value
        cells = np.zeros(max(k, 1), dtype=np.uint32)
        imp = np.zeros(19 * max(k, 1))
        _check(lib().talus_mi_lattice_impulses_get(self._h, lattice, cells.ctypes.data_as(C.POINTER(C.c_uint32)),
                                                   _dp(imp), k, C.byref(count)))
        return cells[:k], imp[: 19 * k].reshape(-1, 19)

    def lattice_colours(self, lattice):
        n, h = self.lattice_nodes(lattice), self.lattices[lattice][1]
        nc = np.zeros(n, dtype=np.uint8)
        hc = np.zeros(h, dtype=np.uint8)
        _check(lib().talus_mi_lattice_colours_get(self._h, lattice, nc.ctypes.data_as(C.POINTER(C.c_uint8)),
                                                  hc.ctypes.data_as(C.POINTER(C.c_uint8))))
        return nc, hc

    def lattice_plastic(self, lattice):
        """(P[8 hexes, 3, 3], plastic strain[8 hexes]) per Gauss point, or None
        before the body first flows."""
        h = self.lattices[lattice][1]
        P, a, present = np.zeros(72 * h), np.zeros(8 * h), C.c_uint32()
        _check(lib().talus_mi_lattice_plastic_get(self._h, lattice, _dp(P), _dp(a), C.byref(present)))
        return (P.reshape(-1, 3, 3), a) if present.value else None

    def set_lattice_plastic(self, lattice, P=None, hardening=None):
        h = self.lattices[lattice][1]
        P, a = _f64(P, 72 * h), _f64(hardening, 8 * h)
        _check(lib().talus_mi_lattice_plastic_set(self._h, lattice, _dp(P), _dp(a)))

    def lattice_step(self, params: LatticeParams | None = None, **kw) -> dict:
        """solveStrainstep over every lattice of the context."""
        p = params if params is not None else lattice_params(**kw)
        r = LatticeReport()
        _check(lib().talus_mi_lattice_step(self._h, C.byref(p), C.byref(r)))
        return r.as_dict()
