"""FP64 CPU oracle of the deformed-field re-bake of a hex lattice
(field_oracle.c): the sampled SDF probe, probeDeformed and
bakeDeformedField.

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``: imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
A field is a dict: origin (3), voxel (3), res (3, x fastest), distances
(float32, res[2] x res[1] x res[0] flattened x-fastest), materials (uint8).
Meshes are the dicts of ``oracle.hex.build_hex_lattice``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _dp
from .hex import HexMesh
from .screen import _Mesh, lib as _screenlib

_ready = False


class Sdf(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("voxel", C.c_float * 3), ("res", C.c_uint32 * 3),
                ("dist", C.POINTER(C.c_float)), ("mat", C.POINTER(C.c_uint8))]


def lib():
    global _ready
    L = _screenlib()
    if not _ready:
        dp = C.POINTER(C.c_double)
        u8p = C.POINTER(C.c_uint8)
        u32p = C.POINTER(C.c_uint32)
        L.orc_sdf_probe.argtypes = [C.POINTER(Sdf), dp]
        L.orc_sdf_probe.restype = C.c_double
        L.orc_probe_deformed.argtypes = [C.POINTER(HexMesh), u8p, C.POINTER(Sdf), dp, dp, dp, u32p, dp]
        L.orc_probe_deformed.restype = C.c_int
        L.orc_bake_deformed.argtypes = [C.POINTER(HexMesh), u8p, dp, dp, C.POINTER(Sdf), C.POINTER(Sdf), C.c_int,
                                        u32p, u32p]
        L.orc_bake_deformed.restype = C.c_int
        L.orc_sdf_distance.argtypes = [C.POINTER(Sdf), C.POINTER(C.c_float)]
        L.orc_sdf_distance.restype = C.c_float
        L.orc_surface_thickness.argtypes = [C.POINTER(Sdf), dp, C.c_uint32, u32p, C.POINTER(C.c_float)]
        L.orc_hex_surface.argtypes = [C.POINTER(HexMesh), u32p, u32p, u32p, u32p, u32p]
        L.orc_hex_surface.restype = C.c_uint32
        _ready = True
    return L


class _Field:
    def __init__(self, field):
        self.dist = np.ascontiguousarray(field["distances"], dtype=np.float32).reshape(-1)
        self.mat = np.ascontiguousarray(field.get("materials", np.zeros(self.dist.size, np.uint8)),
                                        dtype=np.uint8).reshape(-1)
        s = Sdf()
        s.origin[:] = [float(v) for v in field["origin"]]
        s.voxel[:] = [float(v) for v in field["voxel"]]
        s.res[:] = [int(v) for v in field["res"]]
        assert self.dist.size == int(np.prod(field["res"]))
        s.dist = self.dist.ctypes.data_as(C.POINTER(C.c_float))
        s.mat = self.mat.ctypes.data_as(C.POINTER(C.c_uint8))
        self.s = s


def _crack(mesh, crack):
    if crack is None:
        return None, None
    c = np.ascontiguousarray(crack, dtype=np.uint8)
    assert c.size == len(mesh["hex"])
    return c, c.ctypes.data_as(C.POINTER(C.c_uint8))


def sdf_probe(field, point):
    f = _Field(field)
    p = np.ascontiguousarray(point, dtype=np.float64)
    return lib().orc_sdf_probe(C.byref(f.s), _dp(p))


def sdf_distance(field, point):
    f = _Field(field)
    p = np.ascontiguousarray(point, dtype=np.float32)
    return lib().orc_sdf_distance(C.byref(f.s), p.ctypes.data_as(C.POINTER(C.c_float)))


def hex_surface(mesh):
    """collectSurface's hex faces: triangles, their cells and lanes, edges
    and the clusters of clusterSurface"""
    from . import cluster_surface
    mm = _Mesh(mesh)
    H = len(mesh["hex"])
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    tri = np.zeros(36 * H + 3, np.uint32)
    cell = np.zeros(12 * H + 1, np.uint32)
    lane = np.zeros(12 * H + 1, np.uint32)
    edges = np.zeros(72 * H + 2, np.uint32)
    ne = C.c_uint32()
    nt = lib().orc_hex_surface(C.byref(mm.m), u32(tri), u32(cell), u32(lane), u32(edges), C.byref(ne))
    faces, starts = cluster_surface(tri[: 3 * nt], mesh["rest"])
    return {"triangles": tri[: 3 * nt].reshape(-1, 3).copy(), "triangle_tet": cell[:nt].copy(),
            "triangle_lane": lane[:nt].copy(), "edges": edges[: 2 * ne.value].reshape(-1, 2).copy(),
            "cluster_faces": faces, "cluster_starts": starts}


def surface_thickness(field, rest, triangles):
    f = _Field(field)
    r = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
    t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1)
    out = np.zeros(t.size // 3, dtype=np.float32)
    lib().orc_surface_thickness(C.byref(f.s), _dp(r), t.size // 3, t.ctypes.data_as(C.POINTER(C.c_uint32)),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def probe_deformed(mesh, field, position, point, crack=None):
    """(distance, element, local) or None"""
    mm, f = _Mesh(mesh), _Field(field)
    keep, cp = _crack(mesh, crack)
    pos = np.ascontiguousarray(position, dtype=np.float64).reshape(-1)
    pt = np.ascontiguousarray(point, dtype=np.float64)
    d, e, loc = C.c_double(), C.c_uint32(), np.empty(3)
    ok = lib().orc_probe_deformed(C.byref(mm.m), cp, C.byref(f.s), _dp(pos), _dp(pt), C.byref(d), C.byref(e),
                                  _dp(loc))
    return (d.value, int(e.value), loc) if ok else None


def bake_deformed(mesh, rest, position, model=None, current=None, crack=None, carve=False):
    """(new current distances, (box minimum, box extent)); model is the 3x4 [R | t]"""
    mm, fr = _Mesh(mesh), _Field(rest)
    fc = _Field(current if current is not None else rest)
    fc.dist = fc.dist.copy()
    fc.s.dist = fc.dist.ctypes.data_as(C.POINTER(C.c_float))
    keep, cp = _crack(mesh, crack)
    pos = np.ascontiguousarray(position, dtype=np.float64).reshape(-1)
    M = np.ascontiguousarray(np.eye(3, 4) if model is None else model, dtype=np.float64).reshape(-1)
    lo, ext = np.zeros(3, np.uint32), np.zeros(3, np.uint32)
    ok = lib().orc_bake_deformed(C.byref(mm.m), cp, _dp(pos), _dp(M), C.byref(fr.s), C.byref(fc.s), int(carve),
                                 lo.ctypes.data_as(C.POINTER(C.c_uint32)), ext.ctypes.data_as(C.POINTER(C.c_uint32)))
    if not ok:
        raise ValueError("the current field must share the rest field's lattice")
    return fc.dist.copy(), (lo, ext)
