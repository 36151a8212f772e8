"""The deformed-field re-bake in the oracle (oracle/field_oracle.c), pinned
by the reference's own deformed-field tests and closed forms.

Reference tests mirrored (tests/deformedfield_tests.cc):
  a lattice at rest rebuilds its own field          82-97
  a lattice moved and turned rigidly rebuilds it    99-120
  a lattice whose top sinks gets a lower top        122-154
  a current field on another lattice is refused     177-200
plus: the trilinear probe reproduces a linear field inside the lattice and
adds the distance to the lattice outside it; a uniformly stretched lattice
reads the rest field at the unstretched point; a crack face caps the
distance at minus its depth times the cell's rest extent.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import field as FD
from oracle import hex as H

VOXEL = 0.025
SINK = 0.005
HALF = np.array([0.1, 0.05, 0.1])


def box_sdf(p, half):
    q = np.abs(p) - half
    outside = np.linalg.norm(np.maximum(q, 0.0), axis=-1)
    return outside + np.minimum(q.max(axis=-1), 0.0)


def sample_points(field):
    res = field["res"]
    z, y, x = np.meshgrid(*(np.arange(r) for r in res[::-1]), indexing="ij")
    idx = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float64)
    return np.asarray(field["origin"], np.float64) + np.asarray(field["voxel"], np.float64) * idx


def box_field(half=HALF, voxel=VOXEL, margin=2):
    reach = (half + margin * voxel).astype(np.float32)
    res = np.round(2.0 * reach / np.float32(voxel)).astype(np.uint32) + 1
    field = {"origin": -reach, "voxel": np.full(3, voxel, np.float32), "res": res}
    field["distances"] = box_sdf(sample_points(field), half).astype(np.float32)
    return field


def jelly():
    m = oracle.Material()
    m.friction, m.density, m.strength = 0.7, 1000.0, math.inf
    m.youngs_modulus, m.poissons_ratio = 2e8, 0.3
    m.yield_stress, m.damping_ratio, m.tear_strain = math.inf, math.inf, math.inf
    m.plastic_max, m.crush_stress, m.crush_limit = math.inf, math.inf, 1.0
    return m


def box_mesh(half=HALF, voxel=VOXEL):
    cells = np.round(2 * half / voxel).astype(int)
    return H.build_hex_lattice([c + 1 for c in cells], [voxel] * 3, tuple(-half), [jelly()], [])


def largest_change(a, b):
    return float(np.abs(a - b).max())


def test_probe_is_trilinear_inside_and_adds_the_gap_outside():
    field = {"origin": np.array([-0.1, 0.0, 0.2], np.float32), "voxel": np.array([0.05, 0.04, 0.03], np.float32),
             "res": np.array([5, 4, 6], np.uint32)}
    a, c = np.array([0.3, -1.2, 0.7]), 0.05
    field["distances"] = (sample_points(field) @ a + c).astype(np.float32)
    rng = np.random.default_rng(1)
    lo = field["origin"].astype(np.float64)
    hi = lo + field["voxel"] * (field["res"] - 1)
    for _ in range(20):
        p = lo + rng.uniform(0, 1, 3) * (hi - lo)
        assert FD.sdf_probe(field, p) == pytest.approx(p @ a + c, abs=1e-6)
    p = hi + np.array([0.03, -0.5 * (hi[1] - lo[1]), 0.04])
    clamped = np.clip(p, lo, hi)
    assert FD.sdf_probe(field, p) == pytest.approx(clamped @ a + c + np.linalg.norm(p - clamped), abs=1e-6)


def test_a_lattice_at_rest_rebuilds_its_own_field():
    field, mesh = box_field(), box_mesh()
    baked, (lo, ext) = FD.bake_deformed(mesh, field, mesh["rest"])
    assert (ext == 0).all()
    assert largest_change(baked, field["distances"]) <= 1e-5


def test_a_lattice_moved_and_turned_rigidly_rebuilds_its_own_field():
    field, mesh = box_field(), box_mesh()
    axis = np.array([0.2, 1.0, 0.1])
    axis /= np.linalg.norm(axis)
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
    R = np.eye(3) + math.sin(0.5) * K + (1 - math.cos(0.5)) * K @ K
    t = np.array([0.3, 0.2, -0.1])
    model = np.hstack([R, t[:, None]])
    baked, _ = FD.bake_deformed(mesh, field, mesh["rest"] @ R.T + t, model)
    assert largest_change(baked, field["distances"]) <= 1e-5


def test_a_lattice_whose_top_sinks_rebuilds_a_field_with_a_lower_top():
    field, mesh = box_field(), box_mesh()
    X = mesh["rest"]
    top = X[:, 1].max()
    pos = X.copy()
    pos[np.abs(X[:, 1] - top) < 1e-5, 1] -= SINK
    baked, (lo, ext) = FD.bake_deformed(mesh, field, pos)
    cur = dict(field, distances=baked)
    assert abs(FD.sdf_probe(cur, [0.0, top - SINK, 0.0])) <= 2e-4
    assert FD.sdf_probe(cur, [0.0, top, 0.0]) > 0.8 * SINK
    bottom = [0.0, -top, 0.0]
    assert abs(FD.sdf_probe(cur, bottom) - FD.sdf_probe(field, bottom)) <= 1e-6
    assert 0 < ext[1] < field["res"][1]


def test_a_stretched_lattice_reads_the_rest_field_at_the_unstretched_point():
    field, mesh = box_field(), box_mesh()
    s = np.array([1.2, 1.0, 0.9])
    pos = mesh["rest"] * s
    pts = sample_points(field)
    inside = np.all(np.abs(pts) <= HALF * s - 1e-9, axis=1)
    checked = 0
    for p in pts[inside][::7]:
        found = FD.probe_deformed(mesh, field, pos, p)
        assert found is not None
        assert found[0] == pytest.approx(FD.sdf_probe(field, p / s), abs=1e-9)
        checked += 1
    assert checked > 10


def test_a_crack_face_caps_the_distance_by_its_depth():
    field, mesh = box_field(), box_mesh()
    X = mesh["rest"]
    corner0 = X[mesh["hex"][:, 0]]
    # an interior cell: corner 0 at (-0.025, -0.025, -0.025)
    cell = int(np.argmin(np.linalg.norm(corner0 - np.array([-0.025, -0.025, -0.025]), axis=1)))
    crack = np.zeros(len(mesh["hex"]), np.uint8)
    crack[cell] = 1 << 1  # axis 0, side 1
    p = corner0[cell] + np.array([0.8, 0.5, 0.5]) * VOXEL
    plain = FD.probe_deformed(mesh, field, X, p)
    cracked = FD.probe_deformed(mesh, field, X, p, crack)
    assert plain[1] == cell and cracked[1] == cell
    assert plain[0] == pytest.approx(FD.sdf_probe(field, p), abs=1e-12)
    assert plain[0] < -0.02
    assert cracked[0] == pytest.approx(-(1.0 - 0.8) * VOXEL, rel=1e-6)  # float rest positions


def test_a_current_field_on_another_lattice_is_refused():
    field, mesh = box_field(), box_mesh()
    shifted = dict(field, origin=field["origin"] + np.float32(VOXEL))
    with pytest.raises(ValueError):
        FD.bake_deformed(mesh, field, mesh["rest"], current=shifted)
    finer = dict(field, voxel=field["voxel"] * np.float32(0.5))
    with pytest.raises(ValueError):
        FD.bake_deformed(mesh, field, mesh["rest"], current=finer)


def test_carving_beyond_reach_lifts_unplaced_samples_to_the_pad():
    field, mesh = box_field(), box_mesh()
    pos = mesh["rest"] + np.array([5.0, 0.0, 0.0])  # the lattice leaves the field
    baked, _ = FD.bake_deformed(mesh, field, pos, carve=True)
    pad = np.float32(1.5 * VOXEL)
    assert np.array_equal(baked, np.maximum(field["distances"], pad))
    kept, (_, ext) = FD.bake_deformed(mesh, field, pos)
    assert np.array_equal(kept, field["distances"]) and (ext == 0).all()


def test_surface_thickness_is_minus_the_mean_field_distance_at_the_corners():
    """surfaceThickness (solver.cc:342-351) with field::distance in float:
    a linear field is reproduced by the trilinear read, points beyond the
    lattice read their distance to it; the library equals the oracle bit for
    bit."""
    from talus_amd import api
    field = {"origin": np.array([-0.2, -0.1, -0.15], np.float32),
             "voxel": np.array([0.05, 0.04, 0.05], np.float32), "res": np.array([9, 6, 7], np.uint32)}
    a, c = np.array([0.4, -0.9, 0.2]), -0.02
    field["distances"] = (sample_points(field) @ a + c).astype(np.float32)
    rng = np.random.default_rng(3)
    lo = field["origin"].astype(np.float64)
    hi = lo + field["voxel"] * (field["res"] - 1)
    rest = lo + rng.uniform(0.0, 1.0, (40, 3)) * (hi - lo)
    rest[:3] = hi + np.array([[0.1, 0.0, 0.0], [0.0, 0.2, 0.0], [0.05, 0.05, 0.05]])
    rest = rest.astype(np.float32).astype(np.float64)
    tri = rng.integers(0, 40, (30, 3)).astype(np.uint32)
    tri[0] = [0, 1, 2]
    lib = api.surface_thickness(field, rest, tri)
    orc = FD.surface_thickness(field, rest, tri)
    assert np.array_equal(lib, orc)
    inside = np.all(tri >= 3, axis=1)
    expect = -(rest[tri] @ a + c).mean(axis=1)
    assert np.abs(lib[inside] - expect[inside]).max() <= 1e-6
    out = rest[:3] - np.clip(rest[:3], lo, hi)
    assert lib[0] == pytest.approx(-np.linalg.norm(out, axis=1).mean(), rel=1e-5)
