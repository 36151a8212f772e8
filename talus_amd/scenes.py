"""Seeded synthetic workloads for the deformable tet step.

Each function returns a ``Scene``: structured-lattice inputs for the Kuhn
tet builder (resolution, voxel size, optional jittered sample positions,
support volumes) plus materials and per-tet material lanes.  Every number is
drawn from ``numpy.random.default_rng(seed)`` so CPU and GPU runs see
identical inputs.

Workloads (sizes in hexes; each hex is split into 6 tets):
  cantilever   0.8 x 0.2 x 0.2 m, 10x4x4 -> 960 tets, x=0 face pinned,
               E=1e6 Pa, nu=0.3, rho=1000, neo-Hookean, 60 steps at 1/60 s
  cube         1 m cube, 40x40x40 -> 384,000 tets, bottom face pinned,
               per-tet E = 5e6 exp(0.3 N(0,1)), nu ~ U[0.25, 0.4],
               nodes jittered by U[-0.02 h, 0.02 h] per axis, 240 steps
  slab         2 x 1 x 0.1 m, 82x41x5 -> 100,860 tets, yield 2e5 Pa,
               hardening 1e5 Pa, tear strain 0.35, E=1e6, nu=0.3; x-ends
               pinned, a 0.4 x 0.4 m box driven into the top face at 1 m/s
  bodies       512 cuboid bodies, tet counts log-uniform in [200, 20000],
               E log-uniform [1e5, 1e8], nu ~ U[0.2, 0.45], yield on with
               p = 0.5 (yield stress 1e-3 E), x=0 face pinned, 4 substeps
  beam         1 x 0.1 x 0.1 m, 103x9x9 -> 50,058 tets, materials as cube,
               x=0 face pinned, 60 steps, loss = |x_tip - target|^2
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import api

MATERIAL_DTYPE = np.dtype([
    ("friction", "<f4"), ("restitution", "<f4"), ("density", "<f4"), ("strength", "<f4"),
    ("youngs_modulus", "<f4"), ("poissons_ratio", "<f4"), ("stiffness_damping", "<f4"),
    ("yield_stress", "<f4"), ("hardening_modulus", "<f4"), ("elastic_model", "<u4"),
    ("hysteresis", "<f4"), ("damping_ratio", "<f4"), ("tear_strain", "<f4"),
    ("plastic_max", "<f4"), ("crush_stress", "<f4"), ("crush_limit", "<f4"),
])


def material_array(n: int, **kw) -> np.ndarray:
    """n PhysicalMaterial records with the reference defaults, overridden by kw
    (scalars or arrays of length n)."""
    m = np.zeros(n, dtype=MATERIAL_DTYPE)
    defaults = dict(friction=0.5, restitution=0.0, density=1000.0, strength=np.inf,
                    youngs_modulus=0.0, poissons_ratio=0.3, stiffness_damping=0.0,
                    yield_stress=np.inf, hardening_modulus=0.0, elastic_model=api.COROTATED,
                    hysteresis=0.0, damping_ratio=np.inf, tear_strain=np.inf,
                    plastic_max=np.inf, crush_stress=np.inf, crush_limit=1.0)
    defaults.update(kw)
    for k, v in defaults.items():
        m[k] = v
    return m


def materials_list(arr: np.ndarray):
    return [api.PhysicalMaterial(**{k: arr[k][i].item() for k in MATERIAL_DTYPE.names})
            for i in range(arr.shape[0])]


@dataclasses.dataclass
class Body:
    resolution: tuple
    voxel: tuple
    origin: tuple = (0.0, 0.0, 0.0)
    sample_positions: np.ndarray | None = None  # (samples, 3) float32
    supports: list = dataclasses.field(default_factory=list)  # [(center, half)]
    materials: np.ndarray | None = None  # MATERIAL_DTYPE
    per_tet: bool = False  # materials[e] belongs to tet e (lane = tet index)

    def lattice_args(self):
        return dict(resolution=self.resolution, voxel_size=self.voxel, origin=self.origin,
                    sample_positions=self.sample_positions)


@dataclasses.dataclass
class Scene:
    name: str
    bodies: list
    steps: int
    dt: float = 1.0 / 60.0
    substeps: int = 1
    gravity: tuple = (0.0, -9.81, 0.0)
    extra: dict = dataclasses.field(default_factory=dict)


def _box_support(center, half):
    return (tuple(float(c) for c in center), tuple(float(h) for h in half))


def cantilever() -> Scene:
    mats = material_array(1, youngs_modulus=1e6, poissons_ratio=0.3, density=1000.0,
                          elastic_model=api.NEO_HOOKEAN)
    body = Body(resolution=(11, 5, 5), voxel=(0.08, 0.05, 0.05),
                supports=[_box_support((0.0, 0.1, 0.1), (0.004, 1.0, 1.0))], materials=mats)
    return Scene("cantilever", [body], steps=60)


def _jittered(resolution, voxel, rng, fraction):
    rx, ry, rz = resolution
    iz, iy, ix = np.meshgrid(np.arange(rz), np.arange(ry), np.arange(rx), indexing="ij")
    base = np.stack([ix * voxel[0], iy * voxel[1], iz * voxel[2]], axis=-1).reshape(-1, 3)
    jitter = rng.uniform(-fraction, fraction, size=base.shape) * np.asarray(voxel)
    return (base + jitter).astype(np.float32)


def _lognormal_materials(n, rng, **kw):
    E = 5e6 * np.exp(0.3 * rng.standard_normal(n))
    nu = rng.uniform(0.25, 0.4, n)
    return material_array(n, youngs_modulus=E, poissons_ratio=nu, density=1000.0,
                          elastic_model=api.NEO_HOOKEAN, **kw)


def cube(cells: int = 40, seed: int = 2024) -> Scene:
    rng = np.random.default_rng(seed)
    h = 1.0 / cells
    res = (cells + 1,) * 3
    pos = _jittered(res, (h, h, h), rng, 0.02)
    tets = 6 * cells ** 3
    mats = _lognormal_materials(tets, rng)
    body = Body(resolution=res, voxel=(h, h, h), sample_positions=pos,
                supports=[_box_support((0.5, 0.0, 0.5), (1.0, 0.05 * h, 1.0))],
                materials=mats, per_tet=True)
    return Scene(f"cube{cells}", [body], steps=240)


def hex_cube(cells: int = 40, seed: int = 2024) -> dict:
    """The cube as a strainstep hex lattice: cells^3 hexes of a unit cube,
    nodes jittered by up to 2 % of a cell, every hex its own material lane
    (E = 5e6 exp(0.3 N(0,1)) Pa, nu ~ U[0.25, 0.4], rho 1000), bottom face
    held.  Keyword arguments of api.build_hex_lattice plus "materials"."""
    rng = np.random.default_rng(seed)
    h = 1.0 / cells
    res = (cells + 1,) * 3
    pos = _jittered(res, (h, h, h), rng, 0.02)
    n = cells ** 3
    E = 5e6 * np.exp(0.3 * rng.standard_normal(n))
    nu = rng.uniform(0.25, 0.4, n)
    mats = material_array(n, youngs_modulus=E, poissons_ratio=nu, density=1000.0)
    return dict(resolution=res, voxel_size=(h, h, h), origin=(0.0, 0.0, 0.0), materials=mats,
                supports=[api.Support.box((0.5, 0.0, 0.5), (1.0, 0.05 * h, 1.0))], sample_positions=pos,
                cell_lane=np.arange(n, dtype=np.uint32))


def slab(cells=(82, 41, 5), youngs_modulus=1e6, held_footprint=True) -> Scene:
    size = (2.0, 1.0, 0.1)
    voxel = tuple(size[k] / cells[k] for k in range(3))
    mats = material_array(1, youngs_modulus=youngs_modulus, poissons_ratio=0.3, density=1000.0,
                          elastic_model=api.NEO_HOOKEAN, yield_stress=2e5,
                          hardening_modulus=1e5, tear_strain=0.35)
    # the box footprint on the top face is held kinematically: its nodes are
    # anchored and driven down at 1 m/s (slab_drive)
    driver = dict(center=(1.0, 0.5), half=(0.2, 0.2), top=size[2], speed=1.0)
    ends = [_box_support((0.0, 0.5, 0.05), (0.2 * voxel[0], 1.0, 1.0)),
            _box_support((2.0, 0.5, 0.05), (0.2 * voxel[0], 1.0, 1.0))]
    if held_footprint:
        ends.append(_box_support((driver["center"][0], driver["center"][1], size[2]),
                                 (driver["half"][0] + 0.25 * voxel[0], driver["half"][1] + 0.25 * voxel[1],
                                  0.2 * voxel[2])))
    body = Body(resolution=tuple(c + 1 for c in cells), voxel=voxel, supports=ends,
                materials=mats)
    return Scene("slab", [body], steps=180, gravity=(0.0, 0.0, -9.81), extra={"driver": driver})


def slab_driven_nodes(scene: Scene, rest: np.ndarray, anchored: np.ndarray) -> np.ndarray:
    """Indices of the anchored top-face nodes under the crushing box."""
    d = scene.extra["driver"]
    v = scene.bodies[0].voxel
    inside = ((np.abs(rest[:, 0] - d["center"][0]) <= d["half"][0] + 0.5 * v[0])
              & (np.abs(rest[:, 1] - d["center"][1]) <= d["half"][1] + 0.5 * v[1])
              & (np.abs(rest[:, 2] - d["top"]) <= 0.2 * v[2]))
    held = np.asarray(anchored) != 0
    return np.flatnonzero(inside & held) if held[inside].any() else np.flatnonzero(inside)


def slab_drive(scene: Scene, rest_z: np.ndarray, step: int) -> np.ndarray:
    """Heights of the driven nodes at the start of `step`: the box moves down
    at the driver speed (m/s) from the top face."""
    return rest_z - scene.extra["driver"]["speed"] * scene.dt * step


def bodies(count: int = 512, seed: int = 7, lo: int = 200, hi: int = 20000, cell: float = 0.05) -> Scene:
    rng = np.random.default_rng(seed)
    out = []
    targets = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    E = np.exp(rng.uniform(math.log(1e5), math.log(1e8), count))
    nu = rng.uniform(0.2, 0.45, count)
    plastic = rng.uniform(0.0, 1.0, count) < 0.5
    for b in range(count):
        hexes = max(targets[b] / 6.0, 2.0)
        c = max(1, int(round((hexes / 2.0) ** (1.0 / 3.0))))
        a = max(2, int(round(hexes / (c * c))))
        res = (a + 1, c + 1, c + 1)
        kw = dict(youngs_modulus=E[b], poissons_ratio=nu[b], density=1000.0,
                  elastic_model=api.NEO_HOOKEAN)
        if plastic[b]:
            kw.update(yield_stress=1e-3 * E[b], hardening_modulus=1e-2 * E[b])
        mats = material_array(1, **kw)
        out.append(Body(resolution=res, voxel=(cell, cell, cell),
                        supports=[_box_support((0.0, 0.0, 0.0), (0.2 * cell, 1e3, 1e3))],
                        materials=mats))
    return Scene("bodies", out, steps=120, substeps=4)


def hex_bodies(count: int = 512, seed: int = 7, lo: int = 200, hi: int = 20000, cell: float = 0.05) -> list:
    """The batch of uneven cuboid bodies (the same sizes, E and nu as
    bodies(): a body of n tets is a cuboid of about n / 6 cells) as strainstep
    hex lattices, elastic, each held on its x = 0 face.  A list of keyword
    arguments of api.build_hex_lattice plus "materials"."""
    rng = np.random.default_rng(seed)
    out = []
    targets = np.exp(rng.uniform(math.log(lo), math.log(hi), count))
    E = np.exp(rng.uniform(math.log(1e5), math.log(1e8), count))
    nu = rng.uniform(0.2, 0.45, count)
    rng.uniform(0.0, 1.0, count)  # the tet batch's plasticity draw, kept so the sizes match
    for b in range(count):
        hexes = max(targets[b] / 6.0, 2.0)
        c = max(1, int(round((hexes / 2.0) ** (1.0 / 3.0))))
        a = max(2, int(round(hexes / (c * c))))
        out.append(dict(resolution=(a + 1, c + 1, c + 1), voxel_size=(cell, cell, cell), origin=(0.0, 0.0, 0.0),
                        materials=material_array(1, youngs_modulus=E[b], poissons_ratio=nu[b], density=1000.0),
                        supports=[api.Support.box((0.0, 0.0, 0.0), (0.2 * cell, 1e3, 1e3))]))
    return out


def beam(cells=(103, 9, 9), seed: int = 5) -> Scene:
    rng = np.random.default_rng(seed)
    size = (1.0, 0.1, 0.1)
    voxel = tuple(size[k] / cells[k] for k in range(3))
    res = tuple(c + 1 for c in cells)
    tets = 6 * cells[0] * cells[1] * cells[2]
    mats = _lognormal_materials(tets, rng)
    target = np.array([1.0, -0.05, 0.05]) + rng.uniform(-0.01, 0.01, 3)
    body = Body(resolution=res, voxel=voxel,
                supports=[_box_support((0.0, 0.05, 0.05), (0.2 * voxel[0], 1.0, 1.0))],
                materials=mats, per_tet=True)
    return Scene("beam", [body], steps=60, extra={"target": target})


def build_mesh(body: Body):
    """Product mesh (dict of arrays) and the material records for a body."""
    sups = [api.Support.box(c, h) for c, h in body.supports]
    mats = body.materials
    lane_mats = mats if not body.per_tet else mats[:1]
    mesh = api.build_lattice(materials=materials_list(lane_mats), supports=sups,
                             **body.lattice_args())
    if body.per_tet:
        t = mesh["tet_material"].shape[0]
        if mats.shape[0] != t:
            raise ValueError(f"per-tet materials: {mats.shape[0]} records for {t} tets")
        mesh["tet_material"] = np.arange(t, dtype=np.uint32)
    return mesh, mats


# ------------------------------------------------ rigid partners' fields

def sampled_field(fn, lo, hi, voxel, materials=None) -> dict:
    """A sampled field (dict: origin, voxel, res, distances, materials) of
    fn(x, y, z) on the lattice from lo with spacing voxel (float32
    coordinates) up to hi; materials(x, y, z) gives each sample's lane."""
    lo = np.asarray(lo, np.float32)
    vox = np.float32(voxel)
    res = (np.floor((np.asarray(hi, np.float64) - lo) / float(vox) + 1e-6) + 1).astype(np.int64)
    axes = [lo[k] + vox * np.arange(res[k], dtype=np.float32) for k in range(3)]
    z, y, x = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    d = np.asarray(fn(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64)), np.float32)
    f = dict(origin=lo, voxel=(vox, vox, vox), res=res, distances=d.reshape(-1))
    f["materials"] = (np.zeros(d.size, np.uint8) if materials is None
                      else np.asarray(materials(x, y, z), np.uint8).reshape(-1))
    return f


def plane_field(height=0.0, lo=(-2.0, -1.0, -2.0), hi=(2.0, 1.0, 2.0), voxel=0.125, normal=(0.0, 1.0, 0.0),
                materials=None) -> dict:
    """The signed distance n . p - height of a plane, sampled."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    return sampled_field(lambda x, y, z: n[0] * x + n[1] * y + n[2] * z - height, lo, hi, voxel, materials)


def sphere_field(radius=0.25, lo=-0.35, hi=0.35, res=71) -> dict:
    """A sphere about the origin, res samples per axis on [lo, hi]^3."""
    voxel = (hi - lo) / (res - 1)
    return sampled_field(lambda x, y, z: np.sqrt(x * x + y * y + z * z) - radius, (lo, lo, lo), (hi, hi, hi),
                         voxel)


def contact_cube(cells: int = 40, seed: int = 0) -> dict:
    """Contact generation workload: the cells^3 cube (6 cells^3 tets, bottom
    face pinned) falling at 0.5 m/s (+-0.05 noise) onto a plane 2 mm below
    it while three 0.25 m sampled spheres (71^3 samples) move into its top
    and two sides (one spinning at 2 rad/s) and a fourth stays far away;
    dt 1/60 s, bias 2 mm, threshold 2 cm, gravity, duplicates within 1e-2 of
    the cell size.  Returns mesh, materials, rest, velocity, fields, shapes
    (field = index into fields) and the query arguments."""
    sc = cube(cells)
    mesh, mats = build_mesh(sc.bodies[0])
    rest = mesh["rest_positions"].astype(np.float64)
    rng = np.random.default_rng(seed)
    vel = np.zeros_like(rest)
    vel[:, 1] = -0.5
    vel += 0.05 * rng.normal(size=vel.shape)
    vel[mesh["anchored"] != 0] = 0.0
    fields = [plane_field(height=-0.002, lo=(-0.5, -0.5, -0.5), hi=(1.5, 0.5, 1.5), voxel=0.05), sphere_field()]
    shapes = [dict(field=0),
              dict(field=1, center=(0.5, 1.27, 0.5), linear=(0.0, -1.0, 0.0)),
              dict(field=1, center=(1.26, 0.5, 0.5), linear=(-1.0, 0.0, 0.0), angular=(0.0, 2.0, 0.0)),
              dict(field=1, center=(0.5, 0.5, -0.26), linear=(0.0, 0.0, 1.0)),
              dict(field=1, center=(-0.8, 0.5, 0.5))]
    h = 1.0 / cells
    query = dict(dt=1.0 / 60.0, bias=0.002, threshold=0.02, gravity=(0.0, -9.81, 0.0), voxel=(h, h, h))
    return dict(mesh=mesh, materials=mats, rest=rest, velocity=vel, fields=fields, shapes=shapes, query=query,
                name=f"contact_cube{cells}")
