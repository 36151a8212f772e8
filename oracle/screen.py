"""FP64 CPU oracle of the FP64 neo-Hookean hex, the lattice energy ledger and
the screen factor (screen_oracle.c).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``: imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
Meshes are the dicts of ``oracle.hex.build_hex_lattice``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _dp
from .hex import HexMesh, _f64, lib as _hexlib

_ready = False


class RestBody(C.Structure):
    _fields_ = [("total", C.c_double), ("centre", C.c_double * 3), ("inertia", C.c_double * 9),
                ("axes", C.c_double * 9), ("moments", C.c_double * 3), ("thickness", C.c_double)]


def lib():
    global _ready
    L = _hexlib()
    if not _ready:
        dp = C.POINTER(C.c_double)
        u32p = C.POINTER(C.c_uint32)
        L.orc_snh_moduli.argtypes = [C.c_float, C.c_float, dp, dp]
        L.orc_snh_cell_energy.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double]
        L.orc_snh_cell_energy.restype = C.c_double
        L.orc_snh_node_terms.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, C.c_uint32, dp, dp]
        L.orc_snh_rest_stiffness.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_snh_volume_ratio.argtypes = [dp, dp, dp]
        L.orc_snh_volume_ratio.restype = C.c_double
        L.orc_lattice_energy.argtypes = [C.POINTER(HexMesh), dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_rest_body_of.argtypes = [C.POINTER(HexMesh), C.POINTER(RestBody)]
        L.orc_screen_load.argtypes = [C.POINTER(HexMesh), dp, dp, u32p, C.c_uint32, dp, dp, dp]
        L.orc_screen_load.restype = C.c_int
        _ready = True
    return L


def moduli(youngs, poisson):
    """(shear, shifted) of moduliOf."""
    s, t = C.c_double(), C.c_double()
    lib().orc_snh_moduli(youngs, poisson, C.byref(s), C.byref(t))
    return s.value, t.value


def cell_energy(deformed, rest, half, volume, shear, shifted):
    d, r, h = _f64(deformed, 24), _f64(rest, 24), _f64(half, 3)
    return lib().orc_snh_cell_energy(_dp(d), _dp(r), _dp(h), volume, shear, shifted)


def node_terms(deformed, rest, half, volume, shear, shifted, corner):
    d, r, h = _f64(deformed, 24), _f64(rest, 24), _f64(half, 3)
    g, H = np.empty(3), np.empty(9)
    lib().orc_snh_node_terms(_dp(d), _dp(r), _dp(h), volume, shear, shifted, corner, _dp(g), _dp(H))
    return g, H.reshape(3, 3)


def rest_stiffness(half, volume, shear, shifted):
    h = _f64(half, 3)
    K = np.empty(576)
    lib().orc_snh_rest_stiffness(_dp(h), volume, shear, shifted, _dp(K))
    return K.reshape(24, 24)


def volume_ratio(deformed, rest, half):
    d, r, h = _f64(deformed, 24), _f64(rest, 24), _f64(half, 3)
    return lib().orc_snh_volume_ratio(_dp(d), _dp(r), _dp(h))


class _Mesh:
    """an orc_hex_mesh over numpy arrays kept alive with it"""

    def __init__(self, mesh):
        self.keep = {
            "rest": _f64(mesh["rest"]),
            "mass": np.ascontiguousarray(mesh["mass"], dtype=np.float32),
            "anchored": np.ascontiguousarray(mesh["anchored"], dtype=np.uint8),
            "hex": np.ascontiguousarray(mesh["hex"], dtype=np.uint32).reshape(-1),
            "volume": np.ascontiguousarray(mesh["volume"], dtype=np.float32),
            "lane": np.ascontiguousarray(mesh["lane"], dtype=np.uint32),
        }
        m = HexMesh()
        m.nodes = self.keep["mass"].size
        m.hexes = self.keep["volume"].size
        m.rest = _dp(self.keep["rest"])
        m.mass = self.keep["mass"].ctypes.data_as(C.POINTER(C.c_float))
        m.anchored = self.keep["anchored"].ctypes.data_as(C.POINTER(C.c_uint8))
        m.hex = self.keep["hex"].ctypes.data_as(C.POINTER(C.c_uint32))
        m.volume = self.keep["volume"].ctypes.data_as(C.POINTER(C.c_float))
        m.lane = self.keep["lane"].ctypes.data_as(C.POINTER(C.c_uint32))
        m.voxel[:] = [float(v) for v in mesh["voxel"]]
        self.m = m


def cell_moduli(mesh, materials):
    """per cell (shear, shifted) from the cell's material lane, a lane past
    the table taking the last one (makeLattice, bodyset.cc:63-77)"""
    shear = np.empty(len(mesh["lane"]))
    shifted = np.empty_like(shear)

    def field(rec, name):
        return float(rec[name]) if isinstance(rec, np.void) else float(getattr(rec, name))

    for i, lane in enumerate(mesh["lane"]):
        rec = materials[min(int(lane), len(materials) - 1)]
        shear[i], shifted[i] = moduli(field(rec, "youngs_modulus"), field(rec, "poissons_ratio"))
    return shear, shifted


def lattice_energy(mesh, shear, shifted, x, v, gravity):
    """(kinetic, gravity, elastic) of latticeEnergy"""
    mm = _Mesh(mesh)
    s, t = _f64(shear), _f64(shifted)
    xx, vv, g = _f64(x), _f64(v), _f64(gravity, 3)
    out = [C.c_double() for _ in range(3)]
    lib().orc_lattice_energy(C.byref(mm.m), _dp(s), _dp(t), _dp(xx), _dp(vv), _dp(g),
                             *[C.byref(o) for o in out])
    return tuple(o.value for o in out)


def rest_body(mesh):
    mm = _Mesh(mesh)
    rb = RestBody()
    lib().orc_rest_body_of(C.byref(mm.m), C.byref(rb))
    return {"total": rb.total, "centre": np.array(rb.centre[:]),
            "inertia": np.array(rb.inertia[:]).reshape(3, 3), "axes": np.array(rb.axes[:]).reshape(3, 3),
            "moments": np.array(rb.moments[:]), "thickness": rb.thickness}


def screen_load(mesh, shear, shifted, forces, held=()):
    """(displacement (n, 3), dent) of factorScreen + screenLoad"""
    mm = _Mesh(mesh)
    s, t = _f64(shear), _f64(shifted)
    f = _f64(forces)
    hl = np.ascontiguousarray(np.asarray(held, dtype=np.uint32).reshape(-1))
    hp = hl.ctypes.data_as(C.POINTER(C.c_uint32)) if hl.size else None
    u = np.empty(f.size)
    dent = C.c_double()
    ok = lib().orc_screen_load(C.byref(mm.m), _dp(s), _dp(t), hp, hl.size, _dp(f), _dp(u), C.byref(dent))
    if not ok:
        raise RuntimeError("oracle screen factor failed (mesh too large for the dense factor?)")
    return u.reshape(-1, 3), dent.value
