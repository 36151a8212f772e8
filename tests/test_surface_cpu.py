"""Boundary surface and crack faces of tet meshes (host side, no GPU): the
library's talus_mi_mesh_surface against the oracle's independent
restatement, and what the geometry fixes -- a closed outward surface
encloses the mesh volume (divergence theorem: sum over triangles of
a.(b x c)/6 = sum of tet volumes), a box's surface has the box's area and
Euler characteristic 2, every edge of a closed surface bounds exactly two
triangles, and a cut through a strip adds two crack faces per cut face and
twice the cut area."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from talus_amd import api, scenes

from test_tear_cpu import check_tear, pieces, strip, tearable


def surfaces(mesh):
    a = api.mesh_surface(mesh)
    b = O.surface(mesh["tet_nodes"], mesh["rest_positions"], mesh["tet_material"])
    for key in a:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    return a


def enclosed_volume(srf, rest):
    t = srf["triangles"]
    a, b, c = rest[t[:, 0]], rest[t[:, 1]], rest[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def area(tri, rest):
    a, b, c = rest[tri[:, 0]], rest[tri[:, 1]], rest[tri[:, 2]]
    return float(0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum())


def edge_uses(srf):
    t = np.sort(srf["triangles"], axis=1)
    e = np.concatenate([t[:, [0, 1]], t[:, [0, 2]], t[:, [1, 2]]])
    _, counts = np.unique(e, axis=0, return_counts=True)
    return counts


@pytest.mark.parametrize("cells", [(1, 1, 1), (3, 2, 2), (5, 4, 3)])
def test_box_surface_is_closed_outward_and_has_the_box_area(cells):
    h = 0.1
    body = scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(h, h, h), materials=tearable())
    mesh, _ = scenes.build_mesh(body)
    rest = mesh["rest_positions"].astype(np.float64)
    srf = surfaces(mesh)
    L = rest.max(axis=0) - rest.min(axis=0)  # the box as stored (FP32 rest positions)
    assert np.allclose(L, np.array(cells) * h, rtol=1e-6)
    assert abs(enclosed_volume(srf, rest) - L.prod()) <= 1e-12 * L.prod()
    assert abs(area(srf["triangles"], rest) - 2 * (L[0] * L[1] + L[1] * L[2] + L[0] * L[2])) <= 1e-12
    assert np.all(edge_uses(srf) == 2)
    V, E, F = srf["nodes"].size, srf["edges"].shape[0], srf["triangles"].shape[0]
    assert V - E + F == 2
    on_box = np.any((rest == rest.min(axis=0)) | (rest == rest.max(axis=0)), axis=1)
    np.testing.assert_array_equal(srf["nodes"], np.flatnonzero(on_box))
    assert srf["crack_faces"].sum() == 0
    assert np.array_equal(srf["triangle_lane"], mesh["tet_material"][srf["triangle_tet"]])


def test_cut_strip_surfaces_every_piece_and_pairs_its_crack_faces():
    _, mesh, mats = strip()
    rest = mesh["rest_positions"].astype(np.float64)
    before = surfaces(mesh)
    x = rest.copy()
    x[:, 0] *= 1.01
    cent = rest[mesh["tet_nodes"]].mean(axis=1)
    torn = np.nonzero((cent[:, 0] > 0.4) & (cent[:, 0] < 0.6))[0]
    new, _ = check_tear(mesh, mats, x, torn)
    parts = pieces(new["tet_nodes"], new["anchored"].size)
    assert parts >= 2
    nrest = new["rest_positions"].astype(np.float64)
    srf = surfaces(new)
    vol = float(new["volume"].astype(np.float64).sum())
    assert abs(enclosed_volume(srf, nrest) - vol) <= 1e-6 * vol
    assert np.all(edge_uses(srf) == 2)
    V, E, F = srf["nodes"].size, srf["edges"].shape[0], srf["triangles"].shape[0]
    assert V - E + F == 2 * parts  # one sphere per piece
    # the new surface is exactly the crack faces, which come in twins (the
    # two sides of each face the tear opened)
    corner = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
    cracked = [(e, f) for e in range(new["tet_nodes"].shape[0]) for f in range(4) if srf["crack_faces"][e] >> f & 1]
    assert len(cracked) > 0 and len(cracked) % 2 == 0
    crack_tris = np.array([[new["tet_nodes"][e][c] for c in corner[f]] for e, f in cracked])
    added = area(srf["triangles"], nrest) - area(before["triangles"], rest)
    assert abs(added - area(crack_tris, nrest)) <= 1e-9
    assert F - before["triangles"].shape[0] == len(cracked)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_tears_keep_a_closed_outward_surface(seed):
    rng = np.random.default_rng(seed)
    cells, h = (5, 4, 3), 0.1
    res = tuple(c + 1 for c in cells)
    iz, iy, ix = np.meshgrid(*(np.arange(r) for r in res[::-1]), indexing="ij")
    pos = (np.stack([ix, iy, iz], -1).reshape(-1, 3) * h
           + rng.uniform(-0.05, 0.05, (ix.size, 3)) * h).astype(np.float32)
    body = scenes.Body(resolution=res, voxel=(h, h, h), sample_positions=pos, materials=tearable())
    mesh, mats = scenes.build_mesh(body)
    rest = mesh["rest_positions"].astype(np.float64)
    x = rest + 0.02 * rng.standard_normal(rest.shape)
    torn = np.sort(rng.choice(mesh["tet_nodes"].shape[0], 40, replace=False)).astype(np.uint32)
    new, _ = check_tear(mesh, mats, x, torn)
    srf = surfaces(new)
    nrest = new["rest_positions"].astype(np.float64)
    vol = float(new["volume"].astype(np.float64).sum())
    assert abs(enclosed_volume(srf, nrest) - vol) <= 1e-6 * vol
    assert srf["crack_faces"].sum() > 0


def test_bad_meshes_are_rejected():
    body = scenes.Body(resolution=(2, 2, 2), voxel=(1.0, 1.0, 1.0), materials=tearable())
    mesh, _ = scenes.build_mesh(body)
    bad = dict(mesh)
    bad["tet_nodes"] = mesh["tet_nodes"].copy()
    bad["tet_nodes"][0, 0] = 10 ** 6
    with pytest.raises(api.TalusError):
        api.mesh_surface(bad)
    with pytest.raises(ValueError):
        O.surface(bad["tet_nodes"], mesh["rest_positions"])


def _morton(cell):
    """bit i of each 10-bit axis index to bit 3 i + axis"""
    key = 0
    for axis in range(3):
        for bit in range(10):
            key |= ((int(cell[axis]) >> bit) & 1) << (3 * bit + axis)
    return key


@pytest.mark.parametrize("cells", [(1, 1, 1), (6, 3, 2)])
def test_surface_clusters_follow_the_morton_order_of_the_face_centres(cells):
    """clusterSurface (contact.cc:540-597): every triangle once, in order of
    the Morton key of its rest centre over the centres' box (ties by
    index), 32 to a cluster; the library equals the oracle."""
    body = scenes.Body(resolution=tuple(c + 1 for c in cells), voxel=(0.1, 0.07, 0.05), materials=tearable())
    mesh, _ = scenes.build_mesh(body)
    srf = surfaces(mesh)
    tri = srf["triangles"]
    rest = mesh["rest_positions"].astype(np.float32)
    centre = (rest[tri[:, 0]] + rest[tri[:, 1]] + rest[tri[:, 2]]) / np.float32(3.0)
    lo, hi = centre.min(axis=0), centre.max(axis=0)
    span = np.maximum(hi - lo, np.finfo(np.float32).tiny)
    cell = (np.clip((centre - lo) / span, 0, 1) * np.float32(1023.0)).astype(np.uint32)
    keys = np.array([_morton(c) for c in cell], dtype=np.uint64)
    expect = np.lexsort((np.arange(len(keys)), keys))
    np.testing.assert_array_equal(srf["cluster_faces"], expect)
    n = len(tri)
    np.testing.assert_array_equal(srf["cluster_starts"], np.append(np.arange(0, n, 32), n))


@pytest.mark.parametrize("cells", [(1, 1, 1), (4, 2, 3)])
def test_hex_surface_is_closed_outward_and_matches_the_oracle(cells):
    """collectSurface's hex faces (solver.cc:252-303): two triangles per
    unshared cell face, split at the lowest node, facing outward; the
    surface of a box lattice encloses the box, covers its area, uses every
    edge twice and has Euler characteristic 2; the library equals the
    oracle, clusters included."""
    from oracle import field as FD
    voxel = (0.1, 0.08, 0.05)
    mats = [api.PhysicalMaterial(youngs_modulus=1e5, density=1000.0, poissons_ratio=0.3, friction=0.7)]
    mesh = api.build_hex_lattice([c + 1 for c in cells], voxel, (0.0, 0.0, 0.0), mats, [])
    srf = api.hex_mesh_surface(mesh)
    om = {"rest": mesh["rest_positions"].astype(np.float64), "mass": mesh["node_mass"],
          "anchored": mesh["anchored"], "hex": mesh["hex_nodes"], "volume": mesh["volume"],
          "lane": mesh["hex_material"], "voxel": mesh["voxel_size"]}
    orc = FD.hex_surface(om)
    for key in orc:
        np.testing.assert_array_equal(srf[key], orc[key], err_msg=key)
    rest = om["rest"]
    L = np.array(cells) * np.array(voxel)
    assert abs(enclosed_volume(srf, rest) - L.prod()) <= 1e-6 * L.prod()
    assert abs(area(srf["triangles"], rest) - 2 * (L[0] * L[1] + L[1] * L[2] + L[0] * L[2])) <= 1e-6
    assert np.all(edge_uses(srf) == 2)
    V, E, F = srf["nodes"].size, srf["edges"].shape[0], srf["triangles"].shape[0]
    assert V - E + F == 2 and F == 4 * (cells[0] * cells[1] + cells[1] * cells[2] + cells[0] * cells[2])
    on_box = np.any(np.isclose(rest, rest.min(axis=0)) | np.isclose(rest, rest.max(axis=0)), axis=1)
    np.testing.assert_array_equal(srf["nodes"], np.flatnonzero(on_box))
    # every triangle is split at its quad's lowest node and starts there
    assert all(min(t) == t[0] for t in srf["triangles"])
