"""FP32 CPU oracle of contact generation (contact_oracle.c): findContacts of
a deformable surface against rigid partners' sampled fields, with the float
field probe, the closest point on a triangle and the shape's motion.

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``: imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
Fields are the dicts of ``oracle.field``; a shape is a dict with ``field``
and optional center, rotation (w, x, y, z), offset, linear, angular and
lipschitz (default: the field's bound, ``field_lipschitz``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .field import Sdf, _Field, lib as _fieldlib

_ready = False

_f3 = C.c_float * 3


class Shape(C.Structure):
    _fields_ = [("field", C.POINTER(Sdf)), ("center", _f3), ("rotation", C.c_float * 4), ("offset", _f3),
                ("linear", _f3), ("angular", _f3), ("lipschitz", C.c_float)]


class SurfaceContact(C.Structure):
    _fields_ = [("nodes", C.c_uint32 * 3), ("weights", _f3), ("partner", C.c_int32), ("normal", _f3),
                ("local_point", _f3), ("gap", C.c_float), ("bias", C.c_float), ("friction", C.c_float),
                ("triangle", C.c_uint32)]


class ContactMesh(C.Structure):
    _fields_ = [("nodes", C.c_uint32), ("triangles", C.c_uint32), ("lanes", C.c_uint32), ("clusters", C.c_uint32),
                ("tri", C.POINTER(C.c_uint32)), ("lane", C.POINTER(C.c_uint32)),
                ("thickness", C.POINTER(C.c_float)), ("cluster_faces", C.POINTER(C.c_uint32)),
                ("cluster_starts", C.POINTER(C.c_uint32)), ("anchored", C.POINTER(C.c_uint8)),
                ("voxel", _f3), ("friction", C.POINTER(C.c_float))]


class Stats(C.Structure):
    _fields_ = [("clusters_far", C.c_uint32), ("faces_apart", C.c_uint32), ("faces_far", C.c_uint32),
                ("faces_probed", C.c_uint32), ("candidates", C.c_uint32), ("duplicates", C.c_uint32)]


# numpy view of SurfaceContact (68 bytes, no padding)
CONTACT_DTYPE = np.dtype([("nodes", "<u4", 3), ("weights", "<f4", 3), ("partner", "<i4"), ("normal", "<f4", 3),
                          ("local_point", "<f4", 3), ("gap", "<f4"), ("bias", "<f4"), ("friction", "<f4"),
                          ("triangle", "<u4")])
assert CONTACT_DTYPE.itemsize == C.sizeof(SurfaceContact)


def lib():
    global _ready
    L = _fieldlib()
    if not _ready:
        fp = C.POINTER(C.c_float)
        L.orc_find_contacts.argtypes = [C.POINTER(ContactMesh), fp, fp, fp, C.POINTER(Shape), C.c_uint32,
                                        C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(SurfaceContact),
                                        C.c_uint32, C.POINTER(Stats)]
        L.orc_find_contacts.restype = C.c_uint32
        L.orc_field_lipschitz.argtypes = [C.POINTER(Sdf)]
        L.orc_field_lipschitz.restype = C.c_float
        L.orc_field_probe.argtypes = [C.POINTER(Sdf), fp, fp, fp]
        L.orc_field_probe_distance.argtypes = [C.POINTER(Sdf), fp]
        L.orc_field_probe_distance.restype = C.c_float
        L.orc_field_material.argtypes = [C.POINTER(Sdf), fp]
        L.orc_field_material.restype = C.c_uint32
        L.orc_barycentric.argtypes = [fp, fp, fp]
        L.orc_closest_on_triangle.argtypes = [C.POINTER(Shape), C.c_float, fp, fp, fp, fp, fp]
        L.orc_advance_shape.argtypes = [C.POINTER(Shape), C.c_float, fp, fp]
        _ready = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def field_lipschitz(field):
    f = _Field(field)
    return float(lib().orc_field_lipschitz(C.byref(f.s)))


def field_probe(field, point):
    f = _Field(field)
    p, d, n = _f32(point, 3), np.zeros(1, np.float32), np.zeros(3, np.float32)
    lib().orc_field_probe(C.byref(f.s), _fp(p), _fp(d), _fp(n))
    return np.float32(d[0]), n


def field_probe_distance(field, point):
    f = _Field(field)
    return np.float32(lib().orc_field_probe_distance(C.byref(f.s), _fp(_f32(point, 3))))


def field_material(field, point):
    f = _Field(field)
    return int(lib().orc_field_material(C.byref(f.s), _fp(_f32(point, 3))))


def barycentric(triangle, point):
    t, p, w = _f32(triangle, 9), _f32(point, 3), np.zeros(3, np.float32)
    lib().orc_barycentric(_fp(t), _fp(p), _fp(w))
    return w


class _Shapes:
    """ctypes array of Shape structs over the given dicts (keeps fields alive)."""

    def __init__(self, shapes):
        self.fields = []
        self.arr = (Shape * max(1, len(shapes)))()
        for i, sh in enumerate(shapes):
            f = _Field(sh["field"])
            self.fields.append(f)
            s = self.arr[i]
            s.field = C.pointer(f.s)
            s.center[:] = [float(v) for v in sh.get("center", (0, 0, 0))]
            s.rotation[:] = [float(v) for v in sh.get("rotation", (1, 0, 0, 0))]
            s.offset[:] = [float(v) for v in sh.get("offset", (0, 0, 0))]
            s.linear[:] = [float(v) for v in sh.get("linear", (0, 0, 0))]
            s.angular[:] = [float(v) for v in sh.get("angular", (0, 0, 0))]
            lip = sh.get("lipschitz")
            s.lipschitz = float(lib().orc_field_lipschitz(C.byref(f.s)) if lip is None else lip)


def closest_on_triangle(shape, triangle, elapsed=0.0):
    sh = _Shapes([shape])
    t = _f32(triangle, 9)
    w, p, d, n = (np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(1, np.float32),
                  np.zeros(3, np.float32))
    lib().orc_closest_on_triangle(C.byref(sh.arr[0]), float(elapsed), _fp(t), _fp(w), _fp(p), _fp(d), _fp(n))
    return w, p, np.float32(d[0]), n


def advance_shape(shape, elapsed):
    sh = _Shapes([shape])
    c, q = np.zeros(3, np.float32), np.zeros(4, np.float32)
    lib().orc_advance_shape(C.byref(sh.arr[0]), float(elapsed), _fp(c), _fp(q))
    return c, q


def find_contacts(surface, position, velocity, shapes, dt, bias=0.0, threshold=0.0, cull=True, gravity=(0, 0, 0),
                  anchored=None, thickness=None, friction=None, voxel=(1, 1, 1), capacity=None, stats=False):
    """findContacts over ``surface`` (dict: triangles (T x 3), optional
    triangle_lane, cluster_faces, cluster_starts) with float32 positions and
    velocities (N x 3).  Returns the contacts as a CONTACT_DTYPE array (and
    the culling stats when asked)."""
    L = lib()
    pos = _f32(position)
    vel = _f32(velocity, pos.size)
    n = pos.size // 3
    tri = np.ascontiguousarray(surface["triangles"], dtype=np.uint32).reshape(-1)
    T = tri.size // 3
    keep = [tri]
    m = ContactMesh()
    m.nodes, m.triangles = n, T
    m.tri = tri.ctypes.data_as(C.POINTER(C.c_uint32))
    if surface.get("triangle_lane") is not None:
        lane = np.ascontiguousarray(surface["triangle_lane"], dtype=np.uint32).reshape(-1)
        assert lane.size == T
        keep.append(lane)
        m.lane = lane.ctypes.data_as(C.POINTER(C.c_uint32))
    if thickness is not None:
        th = _f32(thickness, T)
        keep.append(th)
        m.thickness = _fp(th)
    if surface.get("cluster_faces") is not None:
        cf = np.ascontiguousarray(surface["cluster_faces"], dtype=np.uint32).reshape(-1)
        cs = np.ascontiguousarray(surface["cluster_starts"], dtype=np.uint32).reshape(-1)
        keep += [cf, cs]
        m.cluster_faces = cf.ctypes.data_as(C.POINTER(C.c_uint32))
        m.cluster_starts = cs.ctypes.data_as(C.POINTER(C.c_uint32))
        m.clusters = max(0, cs.size - 1)
    anc = np.zeros(n, np.uint8) if anchored is None else np.ascontiguousarray(anchored, dtype=np.uint8).reshape(-1)
    assert anc.size == n
    keep.append(anc)
    m.anchored = anc.ctypes.data_as(C.POINTER(C.c_uint8))
    m.voxel[:] = [float(v) for v in voxel]
    fr = np.zeros(1, np.float32) if friction is None else _f32(friction)
    keep.append(fr)
    m.lanes = 0 if friction is None else fr.size
    m.friction = _fp(fr)
    sh = _Shapes(shapes)
    g = _f32(gravity, 3)
    st = Stats()
    total = L.orc_find_contacts(C.byref(m), _fp(pos), _fp(vel), _fp(g), sh.arr, len(shapes), float(dt), float(bias),
                                float(threshold), int(bool(cull)), None, 0, C.byref(st))
    cap = total if capacity is None else capacity
    out = (SurfaceContact * max(1, cap))()
    total2 = L.orc_find_contacts(C.byref(m), _fp(pos), _fp(vel), _fp(g), sh.arr, len(shapes), float(dt),
                                 float(bias), float(threshold), int(bool(cull)), out, cap, C.byref(st))
    assert total2 == total
    arr = np.frombuffer(bytes(out), dtype=CONTACT_DTYPE)[:min(cap, total)].copy()
    if stats:
        return arr, {k: int(getattr(st, k)) for k, _ in Stats._fields_}, total
    return arr
