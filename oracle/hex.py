"""FP64 CPU oracle for the strainstep hex-lattice solve (hex_oracle.c).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle``: imported by tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import Lattice, Material, Support, _dp, lib as _lib, materials_array

ROWS = 18


class Hex(C.Structure):
    """orc_hex: one 18-row hex element (strainstep/element.h:40-73)."""

    _fields_ = [
        ("cell", C.c_uint32), ("nodes", C.c_uint32 * 8),
        ("gradient", C.c_double * 24), ("inv_mass", C.c_double * 8),
        ("rest_hg", C.c_double * 12), ("bend", C.c_double), ("size", C.c_double),
        ("stiffness", C.c_double * ROWS), ("rate", C.c_double * ROWS),
        ("softness", C.c_double * ROWS), ("cap", C.c_double * ROWS),
        ("inverse", C.c_double * (ROWS * ROWS)),
        ("vol_stiffness", C.c_double), ("vol_rate", C.c_double),
        ("vol_softness", C.c_double), ("vol_cap", C.c_double), ("volumetric", C.c_int),
        ("impulse", C.c_double * ROWS), ("vol_impulse", C.c_double),
        ("quat", C.c_double * 4), ("turn", C.c_double * 9),
        ("strain", C.c_double * ROWS), ("vol_gradient", C.c_double * 24),
        ("coupling", C.c_double * ROWS), ("image", C.c_double * ROWS),
        ("gain", C.c_double), ("rhs", C.c_double), ("schur", C.c_double),
    ]

    def arr(self, name):
        return np.ctypeslib.as_array(getattr(self, name)).copy()


class HexMesh(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint32), ("hexes", C.c_uint32),
        ("rest", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_float)),
        ("anchored", C.POINTER(C.c_uint8)), ("hex", C.POINTER(C.c_uint32)),
        ("volume", C.POINTER(C.c_float)), ("lane", C.POINTER(C.c_uint32)),
        ("voxel", C.c_float * 3),
    ]


_ready = False


def lib():
    global _ready
    L = _lib()
    if not _ready:
        dp = C.POINTER(C.c_double)
        u32p = C.POINTER(C.c_uint32)
        hp = C.POINTER(Hex)
        L.orc_hourglass_sign.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_hourglass_sign.restype = C.c_double
        L.orc_hex_corner_rest.argtypes = [dp, dp, dp]
        L.orc_hex_material.argtypes = [C.c_float, C.c_float, dp, dp]
        L.orc_hex_cell_frequency.argtypes = [C.c_float, C.c_float, C.c_float, C.c_double]
        L.orc_hex_cell_frequency.restype = C.c_double
        L.orc_hex_make.argtypes = [hp, C.c_uint32, u32p, dp, dp, dp, C.c_double, C.c_double,
                                   C.c_double, C.c_double]
        L.orc_hex_make.restype = C.c_int
        L.orc_rotation_of.argtypes = [dp, dp, C.c_uint32, dp]
        L.orc_quat_matrix.argtypes = [dp, dp]
        L.orc_hex_update.argtypes = [hp, dp]
        L.orc_hex_apply.argtypes = [hp, dp, dp, C.c_double]
        L.orc_hex_solve.argtypes = [hp, dp]
        L.orc_build_hex_lattice.argtypes = [C.POINTER(Lattice), C.POINTER(Material), C.c_uint32,
                                            C.POINTER(Support), C.c_uint32, C.POINTER(HexMesh)]
        L.orc_build_hex_lattice.restype = C.c_int
        L.orc_hex_mesh_free.argtypes = [C.POINTER(HexMesh)]
        L.orc_lattice_create.argtypes = [C.POINTER(HexMesh), C.POINTER(Material), C.c_uint32, dp, dp]
        L.orc_lattice_create.restype = C.c_void_p
        L.orc_lattice_free.argtypes = [C.c_void_p]
        L.orc_lattice_set_substeps.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_lattice_get_state.argtypes = [C.c_void_p, dp, dp]
        L.orc_lattice_set_state.argtypes = [C.c_void_p, dp, dp]
        L.orc_lattice_colours.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]
        L.orc_lattice_elements.argtypes = [C.c_void_p]
        L.orc_lattice_elements.restype = C.c_uint32
        L.orc_lattice_impulses.argtypes = [C.c_void_p, u32p, dp]
        L.orc_lattice_substeps.argtypes = [C.c_void_p, C.c_double, C.c_uint32]
        L.orc_lattice_substeps.restype = C.c_uint32
        L.orc_lattice_step.argtypes = [C.c_void_p, C.c_double, dp, C.c_uint32, C.c_uint32]
        L.orc_lattice_step.restype = C.c_uint32
        L.orc_hex_shape_gradients.argtypes = [dp, dp, dp]
        L.orc_hex_return_map.argtypes = [dp, C.c_double, C.c_double, dp, dp, dp]
        L.orc_hex_return_map.restype = C.c_int
        L.orc_hex_fitted_jacobian.argtypes = [dp, dp, C.c_uint32, dp]
        L.orc_hex_fit_holds.argtypes = [dp, dp]
        L.orc_hex_fit_holds.restype = C.c_int
        L.orc_lattice_has_plastic.argtypes = [C.c_void_p]
        L.orc_lattice_has_plastic.restype = C.c_uint32
        L.orc_lattice_get_plastic.argtypes = [C.c_void_p, dp, dp]
        L.orc_lattice_set_plastic.argtypes = [C.c_void_p, dp, dp]
        L.orc_lattice_flow.argtypes = [C.c_void_p]
        L.orc_lattice_flow.restype = C.c_double
        L.orc_lattice_plastic_rest.argtypes = [C.c_void_p, C.c_uint32, dp, dp, dp]
        L.orc_lattice_write_fitted.argtypes = [C.c_void_p, C.c_uint32, dp, dp]
        u32p, u8p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)
        L.orc_lattice_cell_stretch.argtypes = [C.c_void_p, C.c_uint32, dp, dp]
        L.orc_lattice_overstretched.argtypes = [C.c_void_p, u32p]
        L.orc_lattice_overstretched.restype = C.c_uint32
        L.orc_lattice_tear.argtypes = [C.c_void_p, u32p, C.c_uint32]
        L.orc_lattice_tear.restype = C.c_uint32
        L.orc_lattice_surface.argtypes = [C.c_void_p, u32p]
        L.orc_lattice_surface.restype = C.c_uint32
        L.orc_lattice_pieces.argtypes = [C.c_void_p]
        L.orc_lattice_pieces.restype = C.c_uint32
        L.orc_lattice_topology.argtypes = [C.c_void_p, u32p, C.POINTER(C.c_float), u8p]
        L.orc_lattice_topology.restype = C.c_uint32
        L.orc_sym_eigen3.argtypes = [dp, dp, dp]
        L.orc_lattice_lowest_angular.argtypes = [C.c_void_p]
        L.orc_lattice_lowest_angular.restype = C.c_double
        L.orc_lattice_lowest.argtypes = [C.c_void_p]
        L.orc_lattice_lowest.restype = C.c_double
        _ready = True
    return L


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    if n is not None:
        assert a.size == n, (a.size, n)
    return a


def hourglass_sign(mode, corner):
    return lib().orc_hourglass_sign(mode, corner)


def corner_signs(corner):
    return np.array([1.0 if corner & 1 else -1.0, 1.0 if corner & 2 else -1.0,
                     1.0 if corner & 4 else -1.0])


def corner_rest(corners):
    """cornerRest: (shape 3x3, hourglass 4x3) of a hex from its 8 corners."""
    c = _f64(corners, 24)
    shape = np.zeros(9)
    hg = np.zeros(12)
    lib().orc_hex_corner_rest(_dp(c), _dp(shape), _dp(hg))
    return shape.reshape(3, 3), hg.reshape(4, 3)


def material(youngs, poisson):
    s, f = C.c_double(), C.c_double()
    lib().orc_hex_material(youngs, poisson, C.byref(s), C.byref(f))
    return s.value, f.value


def cell_frequency(youngs, poisson, density, cell):
    return lib().orc_hex_cell_frequency(youngs, poisson, density, cell)


def make(corners, inv_mass, volume, youngs, poisson, h, nodes=range(8), cell=0):
    """makeElement from corner rest positions; returns a Hex or None."""
    shape, hg = corner_rest(corners)
    shear, first = material(youngs, poisson)
    e = Hex()
    nd = (C.c_uint32 * 8)(*[int(n) for n in nodes])
    w = _f64(inv_mass, 8)
    sh = _f64(shape, 9)
    hgv = _f64(hg, 12)
    ok = lib().orc_hex_make(C.byref(e), cell, nd, _dp(w), _dp(sh), _dp(hgv), volume, shear, first, h)
    return e if ok else None


def rotation_of(F, guess=(1.0, 0.0, 0.0, 0.0), iterations=16):
    F = _f64(F, 9)
    g = _f64(guess, 4)
    out = np.zeros(4)
    lib().orc_rotation_of(_dp(F), _dp(g), iterations, _dp(out))
    return out


def quat_matrix(q):
    q = _f64(q, 4)
    R = np.zeros(9)
    lib().orc_quat_matrix(_dp(q), _dp(R))
    return R.reshape(3, 3)


def quat_from_axis_angle(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])


def set_rotation(e, q):
    for k in range(4):
        e.quat[k] = float(q[k])
    R = quat_matrix(q).reshape(-1)
    for k in range(9):
        e.turn[k] = float(R[k])


def update(e, positions):
    p = _f64(positions)
    lib().orc_hex_update(C.byref(e), _dp(p))


def solve(e, velocities):
    """solveElement; velocities (n, 3) float64 updated in place."""
    assert velocities.dtype == np.float64 and velocities.flags.c_contiguous
    lib().orc_hex_solve(C.byref(e), _dp(velocities.reshape(-1)))


def apply(e, velocities, impulse, vol_impulse):
    assert velocities.dtype == np.float64 and velocities.flags.c_contiguous
    imp = _f64(impulse, ROWS)
    lib().orc_hex_apply(C.byref(e), _dp(velocities.reshape(-1)), _dp(imp), float(vol_impulse))


def shape_gradients(at, half):
    """shapeGradients: trilinear shape gradients (8, 3) at reference point `at`."""
    a, hh = _f64(at, 3), _f64(half, 3)
    g = np.zeros(24)
    lib().orc_hex_shape_gradients(_dp(a), _dp(hh), _dp(g))
    return g.reshape(8, 3)


def return_map(elastic, shear, yield_stress):
    """returnMap: (mapped elastic F, trial stress, plastic strain) or None."""
    E = _f64(elastic, 9)
    out = np.zeros(9)
    t, st = C.c_double(), C.c_double()
    ok = lib().orc_hex_return_map(_dp(E), shear, yield_stress, _dp(out), C.byref(t), C.byref(st))
    return (out.reshape(3, 3), t.value, st.value) if ok else None


def fitted_jacobian(shape, hourglass, point):
    sh, hg = _f64(shape, 9), _f64(hourglass, 12)
    J = np.zeros(9)
    lib().orc_hex_fitted_jacobian(_dp(sh), _dp(hg), point, _dp(J))
    return J.reshape(3, 3)


def fit_holds(shape, hourglass):
    sh, hg = _f64(shape, 9), _f64(hourglass, 12)
    return bool(lib().orc_hex_fit_holds(_dp(sh), _dp(hg)))


def build_hex_lattice(res, voxel, origin=(0.0, 0.0, 0.0), materials=(), supports=(),
                      sample_pos=None, cell_lane=None, solid=None):
    """Hex mesh of a structured lattice; returns a dict of numpy arrays."""
    L = lib()
    lat = Lattice()
    lat.res[:] = [int(r) for r in res]
    lat.voxel[:] = [float(v) for v in voxel]
    lat.origin[:] = [float(v) for v in origin]
    keep = []
    if sample_pos is not None:
        sp = np.ascontiguousarray(sample_pos, dtype=np.float32).reshape(-1)
        keep.append(sp)
        lat.sample_pos = sp.ctypes.data_as(C.POINTER(C.c_float))
    if cell_lane is not None:
        cl = np.ascontiguousarray(cell_lane, dtype=np.uint32)
        keep.append(cl)
        lat.cell_lane = cl.ctypes.data_as(C.POINTER(C.c_uint32))
    if solid is not None:
        so = np.ascontiguousarray(solid, dtype=np.uint8)
        keep.append(so)
        lat.solid = so.ctypes.data_as(C.POINTER(C.c_uint8))
    mats = materials_array(materials if isinstance(materials, np.ndarray) else list(materials))
    sups = (Support * max(len(supports), 1))()
    for i, s in enumerate(supports):
        sups[i] = s
    out = HexMesh()
    rc = L.orc_build_hex_lattice(C.byref(lat), mats, len(materials), sups, len(supports), C.byref(out))
    if rc != 0:
        raise RuntimeError("oracle hex mesh build failed")
    n, h = out.nodes, out.hexes
    mesh = {
        "rest": np.ctypeslib.as_array(out.rest, (3 * n,)).reshape(n, 3).copy(),
        "mass": np.ctypeslib.as_array(out.mass, (n,)).copy(),
        "anchored": np.ctypeslib.as_array(out.anchored, (n,)).copy(),
        "hex": np.ctypeslib.as_array(out.hex, (8 * h,)).reshape(h, 8).copy(),
        "volume": np.ctypeslib.as_array(out.volume, (h,)).copy(),
        "lane": np.ctypeslib.as_array(out.lane, (h,)).copy(),
        "voxel": np.array(list(out.voxel), dtype=np.float32),
    }
    L.orc_hex_mesh_free(C.byref(out))
    return mesh


class LatticeBody:
    """One strainstep lattice body stepped by the oracle."""

    def __init__(self, mesh, materials, offset=(0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0)):
        L = lib()
        keep = {
            "rest": _f64(mesh["rest"]),
            "mass": np.ascontiguousarray(mesh["mass"], dtype=np.float32),
            "anchored": np.ascontiguousarray(mesh["anchored"], dtype=np.uint8),
            "hex": np.ascontiguousarray(mesh["hex"], dtype=np.uint32).reshape(-1),
            "volume": np.ascontiguousarray(mesh["volume"], dtype=np.float32),
            "lane": np.ascontiguousarray(mesh["lane"], dtype=np.uint32),
        }
        m = HexMesh()
        m.nodes = keep["mass"].size
        m.hexes = keep["volume"].size
        m.rest = _dp(keep["rest"])
        m.mass = keep["mass"].ctypes.data_as(C.POINTER(C.c_float))
        m.anchored = keep["anchored"].ctypes.data_as(C.POINTER(C.c_uint8))
        m.hex = keep["hex"].ctypes.data_as(C.POINTER(C.c_uint32))
        m.volume = keep["volume"].ctypes.data_as(C.POINTER(C.c_float))
        m.lane = keep["lane"].ctypes.data_as(C.POINTER(C.c_uint32))
        m.voxel[:] = [float(v) for v in mesh["voxel"]]
        mats = materials_array(materials if isinstance(materials, np.ndarray) else list(materials))
        off = _f64(offset, 3)
        vel = _f64(velocity, 3)
        self._h = L.orc_lattice_create(C.byref(m), mats, len(materials), _dp(off), _dp(vel))
        if not self._h:
            raise ValueError("a lattice needs its material lanes")
        self.hexes = int(m.hexes)

    @property
    def nodes(self):
        """the current node count (tearing adds node copies)"""
        return int(lib().orc_lattice_topology(self._h, None, None, None))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_lattice_free(self._h)
            self._h = None

    def set_substeps(self, n):
        lib().orc_lattice_set_substeps(self._h, int(n))

    def state(self):
        x = np.zeros(3 * self.nodes)
        v = np.zeros(3 * self.nodes)
        lib().orc_lattice_get_state(self._h, _dp(x), _dp(v))
        return x.reshape(-1, 3), v.reshape(-1, 3)

    def set_state(self, x=None, v=None):
        f = lambda z: None if z is None else _f64(z, 3 * self.nodes)
        x, v = f(x), f(v)
        lib().orc_lattice_set_state(self._h, _dp(x), _dp(v))

    def colours(self):
        n = np.zeros(self.nodes, dtype=np.uint8)
        h = np.zeros(max(self.hexes, 1), dtype=np.uint8)
        lib().orc_lattice_colours(self._h, n.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  h.ctypes.data_as(C.POINTER(C.c_uint8)))
        return n, h[: self.hexes]

    def impulses(self):
        k = lib().orc_lattice_elements(self._h)
        cells = np.zeros(max(k, 1), dtype=np.uint32)
        imp = np.zeros(19 * max(k, 1))
        lib().orc_lattice_impulses(self._h, cells.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(imp))
        return cells[:k], imp[: 19 * k].reshape(-1, 19)

    def substeps(self, dt=1.0 / 60.0, softstep_substeps=4):
        return int(lib().orc_lattice_substeps(self._h, dt, softstep_substeps))

    def step(self, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), softstep_substeps=4, colored=False):
        g = _f64(gravity, 3)
        return int(lib().orc_lattice_step(self._h, dt, _dp(g), softstep_substeps, 1 if colored else 0))

    # ---- plasticity -------------------------------------------------------

    def has_plastic(self):
        return bool(lib().orc_lattice_has_plastic(self._h))

    def plastic(self):
        """(P[8 hexes, 3, 3], plastic strain[8 hexes]) per Gauss point, or None."""
        if not self.has_plastic():
            return None
        P = np.zeros(72 * self.hexes)
        a = np.zeros(8 * self.hexes)
        lib().orc_lattice_get_plastic(self._h, _dp(P), _dp(a))
        return P.reshape(-1, 3, 3), a

    def set_plastic(self, P=None, hardening=None):
        P = None if P is None else _f64(P, 72 * self.hexes)
        a = None if hardening is None else _f64(hardening, 8 * self.hexes)
        lib().orc_lattice_set_plastic(self._h, _dp(P), _dp(a))

    def flow(self):
        return float(lib().orc_lattice_flow(self._h))

    def plastic_rest(self, cell):
        shape, hg, vol = np.zeros(9), np.zeros(12), C.c_double()
        lib().orc_lattice_plastic_rest(self._h, cell, _dp(shape), _dp(hg), C.byref(vol))
        return shape.reshape(3, 3), hg.reshape(4, 3), vol.value

    def write_fitted(self, cell, shape, hourglass):
        sh, hg = _f64(shape, 9), _f64(hourglass, 12)
        lib().orc_lattice_write_fitted(self._h, cell, _dp(sh), _dp(hg))

    def cell_stretch(self, cell):
        """(stretch, direction) of cellStretch"""
        s, d = C.c_double(), np.zeros(3)
        lib().orc_lattice_cell_stretch(self._h, int(cell), C.byref(s), _dp(d))
        return s.value, d

    def overstretched(self):
        out = np.zeros(max(self.hexes, 1), dtype=np.uint32)
        n = lib().orc_lattice_overstretched(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out[:n]

    def tear(self, cells):
        """tearCells: the node copies made"""
        c = np.ascontiguousarray(cells, dtype=np.uint32)
        return int(lib().orc_lattice_tear(self._h, c.ctypes.data_as(C.POINTER(C.c_uint32)), c.size))

    def pieces(self):
        return int(lib().orc_lattice_pieces(self._h))

    def surface(self):
        """surfaceNodes, ascending"""
        n = int(lib().orc_lattice_surface(self._h, None))
        out = np.zeros(max(n, 1), dtype=np.uint32)
        lib().orc_lattice_surface(self._h, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out[:n]

    def lowest_angular(self):
        """lowestAngular of the undamped elements on the current rests"""
        return float(lib().orc_lattice_lowest_angular(self._h))

    def lowest(self):
        """the lowest frequency the elements were last damped for"""
        return float(lib().orc_lattice_lowest(self._h))

    def topology(self):
        """(hex nodes [hexes, 8], node masses, crack-face bits per cell)"""
        hx = np.zeros(8 * self.hexes, dtype=np.uint32)
        mass = np.zeros(self.nodes, dtype=np.float32)
        crack = np.zeros(self.hexes, dtype=np.uint8)
        lib().orc_lattice_topology(self._h, hx.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   mass.ctypes.data_as(C.POINTER(C.c_float)),
                                   crack.ctypes.data_as(C.POINTER(C.c_uint8)))
        return hx.reshape(-1, 8), mass, crack


def sym_eigen3(A):
    """(eigenvalues, eigenvector columns) by the oracle's cyclic Jacobi"""
    w, V = np.zeros(3), np.zeros(9)
    lib().orc_sym_eigen3(_dp(_f64(A, 9)), _dp(w), _dp(V))
    return w, V.reshape(3, 3)
