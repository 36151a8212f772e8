"""FP64 CPU oracle for the Talus deformable tet step.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py``; the product package
``talus_amd`` never imports it.  The C library (``talus_oracle.c``) restates
the reference algorithm; the file header there cites the reference lines each
function follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class Material(C.Structure):
    """Field order of talus::PhysicalMaterial (src/talus/material.h:12-29)."""

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


class Mesh(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint32), ("tets", C.c_uint32),
        ("rest", C.POINTER(C.c_double)), ("mass", C.POINTER(C.c_double)),
        ("anchored", C.POINTER(C.c_uint8)), ("tet", C.POINTER(C.c_uint32)),
        ("dminv", C.POINTER(C.c_double)), ("volume", C.POINTER(C.c_double)),
        ("lane", C.POINTER(C.c_uint32)),
    ]


class Support(C.Structure):
    _fields_ = [("world_to_support", C.c_float * 12), ("half", C.c_float * 3)]


class Lattice(C.Structure):
    _fields_ = [
        ("res", C.c_uint32 * 3), ("voxel", C.c_float * 3),
        ("origin", C.c_float * 3), ("solid", C.POINTER(C.c_uint8)),
        ("cell_lane", C.POINTER(C.c_uint32)),
        ("sample_pos", C.POINTER(C.c_float)),
    ]


class Settings(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float * 3),
        ("substeps", C.c_uint32), ("cg_iterations", C.c_uint32),
        ("tolerance", C.c_float), ("spectral_radius", C.c_float),
        ("threads", C.c_uint32), ("converge", C.c_uint32),
        ("plastic_passes", C.c_uint32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("substeps", C.c_uint32), ("cg_iterations", C.c_uint32),
        ("cg_capped", C.c_uint32), ("solves", C.c_uint32),
        ("cg_worst_relative", C.c_double), ("elastic_energy", C.c_double),
        ("kinetic_energy", C.c_double), ("plastic_growth", C.c_uint32),
        ("yield_events", C.c_uint32), ("non_finite", C.c_uint32),
        ("plastic_evaluations", C.c_uint32), ("plastic_solves", C.c_uint32),
        ("plastic_backtracks", C.c_uint32), ("plastic_linear_reads", C.c_uint32),
        ("plastic_handoffs", C.c_uint32), ("plastic_fallbacks", C.c_uint32),
        ("plastic_rejections", C.c_uint32), ("plastic_capped", C.c_uint32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.orc_build_lattice.argtypes = [C.POINTER(Lattice), C.POINTER(Material), C.c_uint32,
                                        C.POINTER(Support), C.c_uint32, C.POINTER(Mesh)]
        L.orc_mesh_free.argtypes = [C.POINTER(Mesh)]
        L.orc_lame.argtypes = [C.c_float, C.c_float, dp, dp]
        L.orc_det3.argtypes = [dp]
        L.orc_det3.restype = C.c_double
        L.orc_polar.argtypes = [dp, dp]
        L.orc_element_rotation.argtypes = [dp, dp]
        L.orc_snh_tet.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, dp, dp, dp]
        L.orc_corotated_tet.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, dp,
                                        C.c_double, C.c_double, dp, dp, dp, dp,
                                        C.POINTER(C.c_uint32)]
        L.orc_hencky_flow.argtypes = [dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_hencky_flow.restype = C.c_int
        L.orc_stretch.argtypes = [dp, dp]
        L.orc_stretch.restype = C.c_double
        L.orc_eigen_sym3.argtypes = [dp, dp, dp]
        L.orc_body_create.argtypes = [C.POINTER(Mesh), C.POINTER(Material), C.c_uint32]
        L.orc_body_create.restype = C.c_void_p
        L.orc_body_free.argtypes = [C.c_void_p]
        L.orc_body_set_state.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_body_get_state.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_body_get_rotations.argtypes = [C.c_void_p, dp]
        L.orc_body_set_rotations.argtypes = [C.c_void_p, dp]
        L.orc_body_get_frames.argtypes = [C.c_void_p, dp]
        L.orc_body_set_frames.argtypes = [C.c_void_p, dp]
        L.orc_body_get_plastic.argtypes = [C.c_void_p, dp]
        L.orc_body_set_plastic.argtypes = [C.c_void_p, dp]
        L.orc_step.argtypes = [C.c_void_p, C.POINTER(Settings), C.POINTER(Report)]
        L.orc_j2_point.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_j2_point.restype = C.c_int
        L.orc_j2_tangent.argtypes = [dp, C.c_double, C.c_double, dp, dp]
        L.orc_tangent_apply.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_stressed_apply.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_damped_apply.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_tangent_apply.restype = C.c_uint32
        L.orc_eval_forces.argtypes = [C.c_void_p, dp, C.c_uint32, dp, dp, dp, dp]
        L.orc_apply_system.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, dp, C.c_uint32]
        L.orc_tet_stretch.argtypes = [C.c_void_p, dp]
        L.orc_stretch_tets.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, dp, dp, dp]
        L.orc_yield_flags.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        u32p = C.POINTER(C.c_uint32)
        L.orc_surface.argtypes = [C.c_uint32, C.c_uint32, u32p, u32p, dp, u32p, u32p, u32p, u32p, u32p,
                                  u32p, u32p, C.POINTER(C.c_uint8)]
        L.orc_surface.restype = C.c_uint32
        L.orc_cluster_surface.argtypes = [u32p, C.c_uint32, dp, u32p, u32p]
        L.orc_cluster_surface.restype = C.c_uint32
        L.orc_tear_topology.argtypes = [C.c_uint32, C.c_uint32, u32p, dp, dp, C.POINTER(C.c_uint8),
                                        C.c_uint32, u32p, C.c_uint32, u32p]
        L.orc_tear_topology.restype = C.c_uint32
        L.orc_body_tear.argtypes = [C.c_void_p, u32p, C.c_uint32]
        L.orc_body_tear.restype = C.c_uint32
        L.orc_body_nodes.argtypes = [C.c_void_p]
        L.orc_body_nodes.restype = C.c_uint32
        L.orc_body_topology.argtypes = [C.c_void_p, u32p, dp, dp, C.POINTER(C.c_uint8)]
        L.orc_body_set_lame.argtypes = [C.c_void_p, dp, dp]
        L.orc_body_get_lame.argtypes = [C.c_void_p, dp, dp]
        L.orc_body_set_rest.argtypes = [C.c_void_p, dp]
        L.orc_body_set_rest.restype = C.c_int
        L.orc_rollout.argtypes = [C.c_void_p, C.POINTER(Settings), C.c_uint32, C.c_uint32]
        L.orc_rollout.restype = C.c_int
        L.orc_rollout_backward.argtypes = [C.c_void_p, dp, dp, dp, dp, dp, dp, dp, dp, C.c_double]
        L.orc_rollout_backward.restype = C.c_int
        L.orc_hessian_apply.argtypes = [C.c_void_p, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def materials_array(materials):
    """ctypes array of Material, or a pointer into a numpy record array with
    the same 64-byte layout."""
    if isinstance(materials, np.ndarray):
        assert materials.dtype.itemsize == C.sizeof(Material)
        arr = np.ascontiguousarray(materials)
        ptr = arr.ctypes.data_as(C.POINTER(Material))
        ptr._keep = arr
        return ptr
    arr = (Material * max(len(materials), 1))()
    for i, m in enumerate(materials):
        arr[i] = m
    return arr


def build_lattice(res, voxel, origin=(0.0, 0.0, 0.0), materials=(), supports=(),
                  sample_pos=None, cell_lane=None, solid=None):
    """Kuhn tet mesh of a structured lattice; returns a dict of numpy arrays."""
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
    out = Mesh()
    rc = L.orc_build_lattice(C.byref(lat), mats, len(materials), sups, len(supports), C.byref(out))
    if rc != 0:
        raise RuntimeError("oracle mesh build failed")
    n, t = out.nodes, out.tets
    mesh = {
        "rest": np.ctypeslib.as_array(out.rest, (3 * n,)).reshape(n, 3).copy(),
        "mass": np.ctypeslib.as_array(out.mass, (n,)).copy(),
        "anchored": np.ctypeslib.as_array(out.anchored, (n,)).copy(),
        "tet": np.ctypeslib.as_array(out.tet, (4 * t,)).reshape(t, 4).copy(),
        "dminv": np.ctypeslib.as_array(out.dminv, (9 * t,)).reshape(t, 3, 3).copy(),
        "volume": np.ctypeslib.as_array(out.volume, (t,)).copy(),
        "lane": np.ctypeslib.as_array(out.lane, (t,)).copy(),
    }
    L.orc_mesh_free(C.byref(out))
    return mesh


def _mesh_struct(mesh, keep):
    m = Mesh()
    arrs = {
        "rest": np.ascontiguousarray(mesh["rest"], dtype=np.float64).reshape(-1),
        "mass": np.ascontiguousarray(mesh["mass"], dtype=np.float64),
        "anchored": np.ascontiguousarray(mesh["anchored"], dtype=np.uint8),
        "tet": np.ascontiguousarray(mesh["tet"], dtype=np.uint32).reshape(-1),
        "dminv": np.ascontiguousarray(mesh["dminv"], dtype=np.float64).reshape(-1),
        "volume": np.ascontiguousarray(mesh["volume"], dtype=np.float64),
        "lane": np.ascontiguousarray(mesh["lane"], dtype=np.uint32),
    }
    keep.append(arrs)
    m.nodes = arrs["mass"].shape[0]
    m.tets = arrs["volume"].shape[0]
    m.rest = arrs["rest"].ctypes.data_as(C.POINTER(C.c_double))
    m.mass = arrs["mass"].ctypes.data_as(C.POINTER(C.c_double))
    m.anchored = arrs["anchored"].ctypes.data_as(C.POINTER(C.c_uint8))
    m.tet = arrs["tet"].ctypes.data_as(C.POINTER(C.c_uint32))
    m.dminv = arrs["dminv"].ctypes.data_as(C.POINTER(C.c_double))
    m.volume = arrs["volume"].ctypes.data_as(C.POINTER(C.c_double))
    m.lane = arrs["lane"].ctypes.data_as(C.POINTER(C.c_uint32))
    return m


class Body:
    """One deformable tet body stepped by the oracle."""

    def __init__(self, mesh, materials):
        L = lib()
        keep = []
        ms = _mesh_struct(mesh, keep)
        mats = materials_array(materials if isinstance(materials, np.ndarray) else list(materials))
        self._h = L.orc_body_create(C.byref(ms), mats, len(materials))
        self.nodes = int(ms.nodes)
        self.tets = int(ms.tets)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_body_free(self._h)
            self._h = None

    def set_state(self, x=None, v=None, a=None):
        f = lambda z: None if z is None else np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
        x, v, a = f(x), f(v), f(a)
        lib().orc_body_set_state(self._h, _dp(x), _dp(v), _dp(a))

    def state(self):
        x = np.empty(3 * self.nodes)
        v = np.empty(3 * self.nodes)
        a = np.empty(3 * self.nodes)
        lib().orc_body_get_state(self._h, _dp(x), _dp(v), _dp(a))
        return x.reshape(-1, 3), v.reshape(-1, 3), a.reshape(-1, 3)

    def rotations(self):
        R = np.empty(9 * self.tets)
        lib().orc_body_get_rotations(self._h, _dp(R))
        return R.reshape(-1, 3, 3)

    def set_rotations(self, R):
        R = np.ascontiguousarray(R, dtype=np.float64).reshape(-1)
        lib().orc_body_set_rotations(self._h, _dp(R))

    def frames(self):
        W = np.empty(9 * self.tets)
        lib().orc_body_get_frames(self._h, _dp(W))
        return W.reshape(-1, 3, 3)

    def set_frames(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1)
        lib().orc_body_set_frames(self._h, _dp(W))

    def plastic(self):
        p = np.empty(10 * self.tets)
        lib().orc_body_get_plastic(self._h, _dp(p))
        return p.reshape(-1, 10)

    def set_plastic(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        lib().orc_body_set_plastic(self._h, _dp(p))

    def step(self, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), substeps=1, cg_iterations=64,
             tolerance=1e-3, spectral_radius=0.5, threads=1, converge=False, plastic_passes=40):
        s = Settings()
        s.converge = 1 if converge else 0
        s.plastic_passes = plastic_passes
        s.dt = dt
        s.gravity[:] = list(gravity)
        s.substeps = substeps
        s.cg_iterations = cg_iterations
        s.tolerance = tolerance
        s.spectral_radius = spectral_radius
        s.threads = threads
        r = Report()
        lib().orc_step(self._h, C.byref(s), C.byref(r))
        return {k: getattr(r, k) for k, _ in Report._fields_}

    def eval_forces(self, x, threads=1):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        tf = np.empty(12 * self.tets)
        te = np.empty(self.tets)
        nf = np.empty(3 * self.nodes)
        R = np.empty(9 * self.tets)
        lib().orc_eval_forces(self._h, _dp(x), threads, _dp(tf), _dp(te), _dp(nf), _dp(R))
        return tf.reshape(-1, 4, 3), te, nf.reshape(-1, 3), R.reshape(-1, 3, 3)

    def apply_system(self, inertia, step, d, threads=1):
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        y = np.empty_like(d)
        lib().orc_apply_system(self._h, inertia, step, _dp(d), _dp(y), threads)
        return y.reshape(-1, 3)

    def stretch(self):
        s = np.empty(self.tets)
        lib().orc_tet_stretch(self._h, _dp(s))
        return s

    def tear(self, torn):
        torn = np.ascontiguousarray(np.sort(np.asarray(torn, dtype=np.uint32)))
        added = lib().orc_body_tear(self._h, torn.ctypes.data_as(C.POINTER(C.c_uint32)), torn.size)
        self.nodes = int(lib().orc_body_nodes(self._h))
        return int(added)

    def topology(self):
        n = self.nodes
        tet = np.empty(4 * self.tets, dtype=np.uint32)
        rest = np.empty(3 * n)
        mass = np.empty(n)
        anchored = np.empty(n, dtype=np.uint8)
        lib().orc_body_topology(self._h, tet.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(rest), _dp(mass),
                                anchored.ctypes.data_as(C.POINTER(C.c_uint8)))
        return {"tet": tet.reshape(-1, 4), "rest": rest.reshape(-1, 3), "mass": mass, "anchored": anchored}

    def set_lame(self, lam, mu):
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        lib().orc_body_set_lame(self._h, _dp(lam), _dp(mu))

    def lame(self):
        lam, mu = np.empty(self.tets), np.empty(self.tets)
        lib().orc_body_get_lame(self._h, _dp(lam), _dp(mu))
        return lam, mu

    def set_rest(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
        if lib().orc_body_set_rest(self._h, _dp(X)) != 0:
            raise ValueError("a tet degenerates or inverts at these rest positions")

    def hessian_apply(self, pos, d):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        y = np.empty_like(d)
        lib().orc_hessian_apply(self._h, _dp(pos), _dp(d), _dp(y))
        return y.reshape(-1, 3)

    def rollout(self, steps, every=1, dt=1.0 / 60.0, gravity=(0.0, -9.81, 0.0), substeps=1,
                cg_iterations=20000, tolerance=1e-13, spectral_radius=0.5, threads=1):
        s = Settings()
        s.dt = dt
        s.gravity[:] = list(gravity)
        s.substeps = substeps
        s.cg_iterations = cg_iterations
        s.tolerance = tolerance
        s.spectral_radius = spectral_radius
        s.threads = threads
        s.converge = 1
        if lib().orc_rollout(self._h, C.byref(s), steps, every) != 0:
            raise RuntimeError("rollout became non-finite")

    def rollout_backward(self, x_bar, v_bar=None, tolerance=1e-13):
        n, t = self.nodes, self.tets
        xb = np.ascontiguousarray(x_bar, dtype=np.float64).reshape(-1)
        vb = np.zeros(3 * n) if v_bar is None else np.ascontiguousarray(v_bar, dtype=np.float64).reshape(-1)
        out = {k: np.empty(t) for k in ("lambda", "mu")}
        out.update({k: np.empty(3 * n) for k in ("X", "x0", "v0", "a0")})
        rc = lib().orc_rollout_backward(self._h, _dp(xb), _dp(vb), _dp(out["lambda"]), _dp(out["mu"]),
                                        _dp(out["X"]), _dp(out["x0"]), _dp(out["v0"]), _dp(out["a0"]),
                                        tolerance)
        if rc != 0:
            raise RuntimeError(f"rollout backward failed ({rc})")
        for k in ("X", "x0", "v0", "a0"):
            out[k] = out[k].reshape(-1, 3)
        return out

    def tangent_apply(self, x, d):
        """(K_tangent d, loading tets) at positions x, as a Newton attempt
        assembles it (loading margin 1e-3, consistent tangent bound)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        y = np.empty_like(d)
        n = lib().orc_tangent_apply(self._h, _dp(x), _dp(d), _dp(y))
        return y.reshape(-1, 3), int(n)

    def stressed_apply(self, x, d):
        """(R K R^T + initialStress) d at positions x (h = 1, no inertia)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        y = np.empty_like(d)
        lib().orc_stressed_apply(self._h, _dp(x), _dp(d), _dp(y))
        return y.reshape(-1, 3)

    def damped_apply(self, x, inertia, step, damp, d):
        """inertia M d + step^2 K d + damp sum tau D d at positions x."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1)
        y = np.empty_like(d)
        lib().orc_damped_apply(self._h, _dp(x), inertia, step, damp, _dp(d), _dp(y))
        return y.reshape(-1, 3)

    def yield_flags(self):
        f = np.empty(self.tets, dtype=np.uint8)
        lib().orc_yield_flags(self._h, f.ctypes.data_as(C.POINTER(C.c_uint8)))
        return f


# ---- per-element kernels, for closed-form and finite-difference checks ----

def lame(E, nu):
    lam, mu = C.c_double(), C.c_double()
    lib().orc_lame(E, nu, C.byref(lam), C.byref(mu))
    return lam.value, mu.value


def snh_tet(u, dminv, volume, lam, mu, plastic_offset=None):
    u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
    dm = np.ascontiguousarray(dminv, dtype=np.float64).reshape(-1)
    ep = None if plastic_offset is None else np.ascontiguousarray(plastic_offset, dtype=np.float64).reshape(-1)
    f = np.empty(12)
    e = C.c_double()
    lib().orc_snh_tet(_dp(u), _dp(dm), volume, lam, mu, _dp(ep), _dp(f), C.byref(e))
    return f.reshape(4, 3), e.value


def corotated_tet(x, X, dminv, volume, lam, mu, plastic=None, yield_stress=np.inf, hardening=0.0):
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1)
    dm = np.ascontiguousarray(dminv, dtype=np.float64).reshape(-1)
    pl = np.zeros(10) if plastic is None else np.ascontiguousarray(plastic, dtype=np.float64).reshape(-1)
    R = np.empty(9)
    f = np.empty(12)
    e = C.c_double()
    nxt = np.empty(10)
    y = C.c_uint32()
    lib().orc_corotated_tet(_dp(x), _dp(X), _dp(dm), volume, lam, mu, _dp(pl), yield_stress, hardening,
                            _dp(R), _dp(f), C.byref(e), _dp(nxt), C.byref(y))
    return {"R": R.reshape(3, 3), "force": f.reshape(4, 3), "energy": e.value,
            "plastic": nxt, "yielding": y.value}


def j2_point(strain, lam, mu, yield_stress, hardening, plastic=None):
    """(stress, [theta, thetaBar, normal(9)], returned) of the J2 return map."""
    e = np.ascontiguousarray(strain, dtype=np.float64).reshape(9)
    p = np.zeros(10) if plastic is None else np.ascontiguousarray(plastic, dtype=np.float64)
    out, lin = np.empty(9), np.empty(11)
    r = lib().orc_j2_point(_dp(e), _dp(p), lam, mu, yield_stress, hardening, _dp(out), _dp(lin))
    return out.reshape(3, 3), lin, bool(r)


def j2_tangent(lin, lam, mu, e):
    """tangentOf: the consistent tangent's stress for the strain e."""
    lin = np.ascontiguousarray(lin, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(9)
    out = np.empty(9)
    lib().orc_j2_tangent(_dp(lin), lam, mu, _dp(e), _dp(out))
    return out.reshape(3, 3)


def polar(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    R = np.empty(9)
    lib().orc_polar(_dp(F), _dp(R))
    return R.reshape(3, 3)


def element_rotation(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    R = np.empty(9)
    lib().orc_element_rotation(_dp(F), _dp(R))
    return R.reshape(3, 3)


def hencky_flow(F, mu, yield_stress, hardening, plastic_max=np.inf, offset=None, alpha=0.0):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    ep = np.zeros(9) if offset is None else np.ascontiguousarray(offset, dtype=np.float64).reshape(-1).copy()
    al = C.c_double(alpha)
    flowed = lib().orc_hencky_flow(_dp(F), mu, yield_stress, hardening, plastic_max, _dp(ep), C.byref(al))
    return bool(flowed), ep.reshape(3, 3), al.value


def stretch(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1)
    d = np.empty(3)
    s = lib().orc_stretch(_dp(F), _dp(d))
    return s, d


def stretch_tets(tet, rest, x):
    """Per-tet stretch (overstretched's measure) of connectivity `tet` at
    positions `x`, with the rest inverse rebuilt from `rest`."""
    tet = np.ascontiguousarray(tet, dtype=np.uint32).reshape(-1, 4)
    rest = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1, 3)
    x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    if tet.size and int(tet.max()) >= min(rest.shape[0], x.shape[0]):
        raise ValueError("node index out of range")
    out = np.empty(tet.shape[0])
    lib().orc_stretch_tets(tet.ctypes.data_as(C.POINTER(C.c_uint32)), tet.shape[0], _dp(rest), _dp(x),
                           _dp(out))
    return out


def surface(tet, rest, lane=None):
    """Boundary surface and crack faces of a tet mesh given as arrays: the
    same dict as the library's mesh_surface."""
    tet = np.ascontiguousarray(tet, dtype=np.uint32).reshape(-1, 4)
    rest = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1, 3)
    T, N = tet.shape[0], rest.shape[0]
    if T and int(tet.max()) >= N:
        raise ValueError("node index out of range")
    lane = None if lane is None else np.ascontiguousarray(lane, dtype=np.uint32)
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    tri = np.zeros(12 * T + 1, dtype=np.uint32)
    tri_tet = np.zeros(4 * T + 1, dtype=np.uint32)
    tri_lane = np.zeros(4 * T + 1, dtype=np.uint32)
    edges = np.zeros(24 * T + 2, dtype=np.uint32)
    surf = np.zeros(N + 1, dtype=np.uint32)
    crack = np.zeros(T + 1, dtype=np.uint8)
    ne, ns = C.c_uint32(), C.c_uint32()
    nt = lib().orc_surface(N, T, u32(tet), None if lane is None else u32(lane), _dp(rest), u32(tri),
                           u32(tri_tet), u32(tri_lane), u32(edges), C.byref(ne), u32(surf), C.byref(ns),
                           crack.ctypes.data_as(C.POINTER(C.c_uint8)))
    faces, starts = cluster_surface(tri[: 3 * nt], rest)
    return {"triangles": tri[: 3 * nt].reshape(-1, 3).copy(), "triangle_tet": tri_tet[:nt].copy(),
            "triangle_lane": tri_lane[:nt].copy(), "edges": edges[: 2 * ne.value].reshape(-1, 2).copy(),
            "nodes": surf[: ns.value].copy(), "crack_faces": crack[:T].copy(),
            "cluster_faces": faces, "cluster_starts": starts}


def cluster_surface(triangles, rest):
    """(faces, starts) of clusterSurface"""
    tri = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1)
    rest = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
    n = tri.size // 3
    faces = np.zeros(n + 1, dtype=np.uint32)
    starts = np.zeros(n // 32 + 3, dtype=np.uint32)
    u32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    c = lib().orc_cluster_surface(u32(tri), n, _dp(rest), u32(faces), u32(starts))
    return faces[:n].copy(), starts[: c + 1 if n else 0].copy()


def eigen_sym3(A):
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(-1)
    w = np.empty(3)
    V = np.empty(9)
    lib().orc_eigen_sym3(_dp(A), _dp(w), _dp(V))
    return w, V.reshape(3, 3)


def tear_topology(tet, rest, x, anchored, torn):
    """Tear topology on plain arrays; returns (tet, rest, x, anchored, origin)."""
    tet = np.ascontiguousarray(tet, dtype=np.uint32).reshape(-1).copy()
    nodes = np.asarray(rest).reshape(-1, 3).shape[0]
    torn = np.ascontiguousarray(np.sort(np.asarray(torn, dtype=np.uint32)))
    cap = nodes + 64 * (torn.size + 1)
    r = np.zeros(3 * cap)
    r[: 3 * nodes] = np.asarray(rest, dtype=np.float64).reshape(-1)
    xx = np.zeros(3 * cap)
    xx[: 3 * nodes] = np.asarray(x, dtype=np.float64).reshape(-1)
    an = np.zeros(cap, dtype=np.uint8)
    an[:nodes] = np.asarray(anchored, dtype=np.uint8)
    origin = np.zeros(cap, dtype=np.uint32)
    u32 = C.POINTER(C.c_uint32)
    n2 = lib().orc_tear_topology(nodes, tet.size // 4, tet.ctypes.data_as(u32), _dp(r), _dp(xx),
                                 an.ctypes.data_as(C.POINTER(C.c_uint8)), cap,
                                 torn.ctypes.data_as(u32), torn.size, origin.ctypes.data_as(u32))
    if n2 == 0:
        raise RuntimeError("tear topology overflow")
    return (tet.reshape(-1, 4), r[: 3 * n2].reshape(-1, 3), xx[: 3 * n2].reshape(-1, 3), an[:n2],
            origin[: n2 - nodes])
